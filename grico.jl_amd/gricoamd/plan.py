"""compile(problem) -> AssemblyPlan: host lowering + one upload of the tables to HBM.

Reference call chain (src/plans.jl:170-231): field layout, integration
compile (src/integration.jl:2145-2191 _compile_cell: points, W_q = w_q*detJ,
J^-1; 2677-2742 basis tables), Dirichlet compile (src/plans.jl:486-972),
traversal batches (307-332) and the reduced map (src/assembly.jl:112-314).
Here, in addition, every cell accumulator is lowered to per-point
coefficient tables by probing `cell_accumulate` with unit trial channels
(the contract of src/accumulators.jl:523-539 is linear in the trial
channels), so the device kernels never call user code.
"""

import ctypes

import numpy as np

from . import _devabi, _native
from .problem import (AffineProblem, PointQuadrature, QuadraturePoints, TestChannels, TraceTestChannels,
                      TraceTrialChannels, TrialChannels)
from .quadrature import basis_tables, gauss_legendre_rule

FORM_DIFFUSION = 0
FORM_GENERAL = 1


def _expand_ranges(starts, counts):
    """Concatenation of arange(s, s+c) over (s, c) pairs, vectorized."""
    starts = np.asarray(starts, dtype=np.int64)
    counts = np.asarray(counts, dtype=np.int64)
    total = int(counts.sum())
    if total == 0:
        return np.zeros(0, dtype=np.int64)
    before = np.repeat(np.cumsum(counts) - counts, counts)
    return np.repeat(starts, counts) + (np.arange(total, dtype=np.int64) - before)


def _channels(result, shape, dim):
    if isinstance(result, TestChannels):
        v = np.broadcast_to(np.asarray(result.value, dtype=np.float64), shape)
        if len(result.gradient) != dim:
            raise ValueError("TestChannels gradient must have one entry per axis")
        g = [np.broadcast_to(np.asarray(c, dtype=np.float64), shape) for c in result.gradient]
        return v, g
    v = np.broadcast_to(np.asarray(result, dtype=np.float64), shape)
    return v, [np.zeros(shape)] * dim


class _TermLists:
    """term_cell / term_box / term_dof / term_coef (one entry per term of every
    active mode) built on first access from the sources _gather_maps keeps."""

    _TERM_NAMES = ("term_cell", "term_box", "term_dof", "term_coef")

    def __getattr__(self, name):
        if name in _TermLists._TERM_NAMES and "_terms_fn" in self.__dict__:
            self.__dict__.pop("_terms_fn")(self)
            return self.__dict__[name]
        if name in _TermLists._TERM_NAMES and "_terms_src" in self.__dict__:
            cell_of_mode, box, first, tcount, indices, coefficients = self.__dict__.pop("_terms_src")
            self.term_cell = np.repeat(cell_of_mode, tcount)
            self.term_box = np.repeat(box, tcount)
            sel = _expand_ranges(first, tcount)
            self.term_dof = indices[sel]
            self.term_coef = coefficients[sel]
            return self.__dict__[name]
        raise AttributeError(name)


def brick_grouping(grid, cells):
    """2x2x2 groups of cells by lattice position for the device brick kernel:
    cells of one refinement level (per axis) with coordinates 2i, 2i+1 along
    every axis form a brick, position x + 2y + 4z inside it. Returns a
    (bricks x 8) int32 array of indices into `cells` (-1: no cell of the batch
    at that position), bricks in the order of their first cell, or None when
    the batch already consists of consecutive complete bricks (the device then
    groups cells 8b..8b+7 by itself)."""
    cells = np.asarray(cells, dtype=np.int64)
    n = len(cells)
    if n == 0:
        return None
    lev = grid.levels()[cells].astype(np.int64)
    crd = grid.coords()[cells].astype(np.int64)
    key = np.concatenate([lev, crd // 2], axis=1)
    pos = crd[:, 0] % 2 + 2 * (crd[:, 1] % 2) + 4 * (crd[:, 2] % 2)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    rank = np.empty(len(first), dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(first))
    out = np.full((len(first), 8), -1, dtype=np.int32)
    out[rank[inv.reshape(-1)], pos] = np.arange(n, dtype=np.int32)
    if len(first) * 8 == n and np.array_equal(out.reshape(-1), np.arange(n)):
        return None
    return out


class Bucket(_TermLists):
    """Cells sharing degree tuple and quadrature shape (one device kernel batch)."""

    def __init__(self, space, leaf_idx, degrees, qshape):
        dom = space.domain
        self.dim = D = space.dim
        self.leaf_idx = np.asarray(leaf_idx, dtype=np.int64)  # positions in space.leaves
        self.cells = space.leaves[self.leaf_idx]
        self.degrees = tuple(int(p) for p in degrees)
        self.modes = tuple(p + 1 for p in self.degrees)
        self.points = tuple(int(q) for q in qshape)
        self.n_modes = int(np.prod(self.modes))
        self.n_points = int(np.prod(self.points))
        nc = len(self.cells)
        self.cell_count = nc
        # 1D tables per axis.
        self.B = [basis_tables(self.degrees[a], self.points[a])[0] for a in range(D)]
        self.G = [basis_tables(self.degrees[a], self.points[a])[1] for a in range(D)]
        rules = [gauss_legendre_rule(self.points[a]) for a in range(D)]
        # Tensor reference points / weights, axis 1 fastest.
        grids = np.meshgrid(*[r[0] for r in rules][::-1], indexing="ij")
        self.ref_points = [g.reshape(-1) for g in grids[::-1]]
        wgrid = np.meshgrid(*[r[1] for r in rules][::-1], indexing="ij")
        w = np.ones_like(wgrid[0])
        for g in wgrid:
            w = w * g
        self.ref_weights = w.reshape(-1)
        # Cell geometry (src/geometry.jl).
        sizes = dom.cell_sizes(self.cells)
        centers, halves = dom.cell_centers_and_half_sizes(self.cells)
        detj = np.ldexp(np.prod(sizes, axis=1), -D)
        self.sizes = sizes
        self.weights = self.ref_weights[None, :] * detj[:, None]
        self.phys_points = [centers[:, a:a + 1] + self.ref_points[a][None, :] * halves[:, a:a + 1]
                            for a in range(D)]
        self.inverse_jacobian = np.ascontiguousarray(1.0 / halves)
        self._gather_maps(space)
        self.brick_cells = brick_grouping(dom.grid, self.cells) if D == 3 else None

    def _gather_maps(self, space):
        """local_index / local_coefficient / multi-term CSR over the full mode box."""
        nc, nb = self.cell_count, self.n_modes
        lo = space.mode_offsets[self.leaf_idx]
        hi = space.mode_offsets[self.leaf_idx + 1]
        counts = hi - lo
        mode_ids = _expand_ranges(lo, counts)
        cell_of_mode = np.repeat(np.arange(nc), counts)
        box = space.mode_box[mode_ids].astype(np.int64)
        tcount = space.term_offsets[mode_ids + 1] - space.term_offsets[mode_ids]
        local_index = np.full((nc, nb), -1, dtype=np.int64)
        local_coef = np.zeros((nc, nb))
        single = tcount == 1
        first = space.term_offsets[mode_ids]
        local_index[cell_of_mode[single], box[single]] = space.term_indices[first[single]]
        local_coef[cell_of_mode[single], box[single]] = space.term_coefficients[first[single]]
        multi = np.nonzero(tcount > 1)[0]
        self.multi_offsets = np.zeros(len(multi) + 1, dtype=np.int64)
        if len(multi):
            self.multi_offsets[1:] = np.cumsum(tcount[multi])
            sel = _expand_ranges(space.term_offsets[mode_ids[multi]], tcount[multi])
            self.multi_indices = space.term_indices[sel].astype(np.int32)
            self.multi_coefficients = space.term_coefficients[sel].copy()
            local_index[cell_of_mode[multi], box[multi]] = -(np.arange(len(multi)) + 2)
        else:
            self.multi_indices = np.zeros(0, dtype=np.int32)
            self.multi_coefficients = np.zeros(0)
        if space.dof_count >= 2**31 - 1:
            raise ValueError("dof counts above 2^31-1 are not supported by the device index maps")
        self.local_index = np.ascontiguousarray(local_index.astype(np.int32))
        self.local_coefficient = np.ascontiguousarray(local_coef)
        self.unit_coefficients = bool(np.all(local_coef[local_index >= 0] == 1.0))
        # Expanded term list (cell, box, dof, coefficient) for host-side
        # scatters, built on first use (_TermLists).
        self._terms_src = (cell_of_mode, box, space.term_offsets[mode_ids], tcount, space.term_indices,
                           space.term_coefficients)

    def quadrature(self, lo=0, hi=None):
        hi = self.cell_count if hi is None else hi
        return QuadraturePoints([p[lo:hi] for p in self.phys_points], self.weights[lo:hi],
                                [self.sizes[lo:hi, a:a + 1] for a in range(self.dim)])

    def _responses(self, operators, lo, hi, check_linear):
        """Callback responses to the D+1 unit trial channels on cells lo..hi,
        summed over the operators: [(v, [g_1..g_D]) per channel] (unweighted)."""
        D = self.dim
        shape = (hi - lo, self.n_points)
        q = self.quadrature(lo, hi)
        zeros = np.zeros(shape)
        ones = np.ones(shape)
        out = [(np.zeros(shape), [np.zeros(shape) for _ in range(D)]) for _ in range(D + 1)]
        for op in operators:
            window = getattr(op, "set_window", None)
            if window is not None:  # operators holding per-point data of the whole bucket
                window(lo, hi)
            if check_linear:
                base_v, base_g = _channels(op.cell_accumulate(q, TrialChannels(0, zeros, (zeros,) * D), 0),
                                           shape, D)
                if np.any(base_v != 0) or any(np.any(g != 0) for g in base_g):
                    raise ValueError("cell_accumulate must be linear in the trial channels")
            for ch in range(D + 1):
                val = ones if ch == 0 else zeros
                grad = tuple(ones if ch == a + 1 else zeros for a in range(D))
                v, g = _channels(op.cell_accumulate(q, TrialChannels(0, val, grad), 0), shape, D)
                out[ch][0][...] += v
                for a in range(D):
                    out[ch][1][a][...] += g[a]
        return out

    def _probe(self, operators, lo, hi, check_linear=True, resp=None):
        """(D+1)x(D+1) action per point of cells lo..hi (unweighted)."""
        D = self.dim
        resp = resp if resp is not None else self._responses(operators, lo, hi, check_linear)
        M = np.empty((hi - lo, self.n_points, D + 1, D + 1))
        for ch, (v, g) in enumerate(resp):
            M[..., 0, ch] = v
            for a in range(D):
                M[..., a + 1, ch] = g[a]
        return M

    @staticmethod
    def _diffusion_coefficient(resp, D):
        """c when every point's action is diag(0, c I), else None."""
        v0, g0 = resp[0]
        if np.any(v0 != 0) or any(np.any(g != 0) for g in g0):
            return None
        c = resp[1][1][0]
        for a in range(D):
            v, g = resp[a + 1]
            if np.any(v != 0):
                return None
            for b in range(D):
                if b == a:
                    if not np.array_equal(g[b], c):
                        return None
                elif np.any(g[b] != 0):
                    return None
        return c

    def lower_operators(self, operators, chunk=8192):
        """Probe cell_accumulate with unit trial channels; returns (form, point_data).

        DIFFUSION when every point's action is diag(0, c I): one coefficient
        W_q c_q per point; otherwise GENERAL with the full W_q M_q matrix. The
        action matrix is only materialized for the general form; linearity in
        the trial channels is checked on the first chunk."""
        D = self.dim
        nc = self.cell_count
        coef = np.zeros((nc, self.n_points))
        general = None
        for lo in range(0, nc, chunk):
            hi = min(nc, lo + chunk)
            W = self.weights[lo:hi]
            resp = None
            if general is None:
                resp = self._responses(operators, lo, hi, check_linear=lo == 0)
                c = self._diffusion_coefficient(resp, D)
                if c is not None:
                    coef[lo:hi] = W * c
                    continue
                general = np.zeros((nc, self.n_points * (D + 1) ** 2))
                for lo2 in range(0, lo, chunk):
                    hi2 = min(nc, lo2 + chunk)
                    M2 = self._probe(operators, lo2, hi2, check_linear=False)
                    general[lo2:hi2] = (self.weights[lo2:hi2][..., None, None] * M2).reshape(hi2 - lo2, -1)
            M = self._probe(operators, lo, hi, check_linear=lo == 0, resp=resp)
            general[lo:hi] = (W[..., None, None] * M).reshape(hi - lo, -1)
        if general is None:
            return FORM_DIFFUSION, np.ascontiguousarray(coef)
        return FORM_GENERAL, np.ascontiguousarray(general)

    def interpolate(self, coefficients_box, derivative_axis=None):
        """Sum-factorized evaluation on all cells: (nc, N box) -> (nc, Q); host numpy."""
        D = self.dim
        t = coefficients_box.reshape((self.cell_count,) + self.modes[::-1])
        for a in range(D):
            table = self.G[a] if a == derivative_axis else self.B[a]
            ax = D - a  # numpy axis of mode axis a (axis 1 fastest = last)
            t = np.moveaxis(np.tensordot(t, table, axes=([ax], [0])), -1, ax)
        out = t.reshape(self.cell_count, -1)
        if derivative_axis is not None:
            out = out * self.inverse_jacobian[:, derivative_axis:derivative_axis + 1]
        return out

    def project(self, point_values, derivative_axis=None):
        """Transpose of interpolate: (nc, Q) -> (nc, N box)."""
        D = self.dim
        vals = point_values
        if derivative_axis is not None:
            vals = vals * self.inverse_jacobian[:, derivative_axis:derivative_axis + 1]
        t = vals.reshape((self.cell_count,) + self.points[::-1])
        for a in range(D):
            table = self.G[a] if a == derivative_axis else self.B[a]
            ax = D - a
            t = np.moveaxis(np.tensordot(t, table, axes=([ax], [1])), -1, ax)
        return t.reshape(self.cell_count, -1)


def _basis_at(x, degree):
    """psi_0..psi_p and derivatives at points x (n,) -> (n, p+1) each."""
    from .quadrature import fe_basis_values_and_derivatives
    return fe_basis_values_and_derivatives(x, degree)


class CloudBucket(_TermLists):
    """Items that carry their own reference points: cells with a custom
    quadrature (src/accumulators.jl:1022-1054) or embedded-surface pieces
    (1133-1214, src/integration.jl:2338-2368). One device batch per degree
    tuple; the point action is always the general (D+1) x (D+1) form."""

    cloud = True

    def __init__(self, space, leaf_idx, degrees, rules, normals=None):
        dom = space.domain
        self.dim = D = space.dim
        self.surface = normals is not None
        self.leaf_idx = np.asarray(leaf_idx, dtype=np.int64)
        self.cells = space.leaves[self.leaf_idx]
        self.degrees = tuple(int(p) for p in degrees)
        self.modes = tuple(p + 1 for p in self.degrees)
        self.n_modes = int(np.prod(self.modes))
        self.cell_count = nc = len(self.cells)
        counts = np.asarray([r.point_count() for r in rules], dtype=np.int64)
        self.point_offsets = np.zeros(nc + 1, dtype=np.int64)
        self.point_offsets[1:] = np.cumsum(counts)
        self.n_points_total = int(self.point_offsets[-1])
        self.ref = np.ascontiguousarray(np.concatenate([r.points for r in rules], axis=0).reshape(-1, D)) \
            if nc else np.zeros((0, D))
        ref_w = np.concatenate([r.weights for r in rules]) if nc else np.zeros(0)
        self.item_of_point = np.repeat(np.arange(nc), counts)
        sizes = dom.cell_sizes(self.cells)
        centers, halves = dom.cell_centers_and_half_sizes(self.cells)
        self.sizes = sizes
        self.inverse_jacobian = np.ascontiguousarray(1.0 / halves)
        h = halves[self.item_of_point]
        detj = np.prod(h, axis=1)
        self.phys_points = [centers[self.item_of_point, a] + self.ref[:, a] * h[:, a] for a in range(D)]
        if self.surface:
            n_ref = np.concatenate(normals, axis=0).reshape(-1, D)
            # src/integration.jl:2501-2516: weight scale det J / |J n_ref|, normal J^-1 n_ref normalized
            mapped = np.linalg.norm(h * n_ref, axis=1)
            self.weights = ref_w * detj / mapped
            t = n_ref / h
            self.normals = t / np.linalg.norm(t, axis=1)[:, None]
        else:
            self.weights = ref_w * detj
            self.normals = None
        Bucket._gather_maps(self, space)

    face_axis = None
    face_side = None

    def quadrature(self):
        return QuadraturePoints(self.phys_points, self.weights,
                                [self.sizes[self.item_of_point, a] for a in range(self.dim)],
                                normal=None if self.normals is None else [self.normals[:, a] for a in range(self.dim)],
                                face_axis=self.face_axis, face_side=self.face_side)

    def lower(self, operators, surface=False, callback=None):
        """Probe the callbacks with unit trial channels at every point:
        point_data (P x (D+1)^2) = W_q M_q, row-major [test channel][trial channel]."""
        D = self.dim
        P = self.n_points_total
        q = self.quadrature()
        ones, zeros = np.ones(P), np.zeros(P)
        M = np.zeros((P, D + 1, D + 1))
        callback = callback or ("surface_accumulate" if surface else "cell_accumulate")
        for op in operators:
            call = getattr(op, callback)
            for ch in range(D + 1):
                val = ones if ch == 0 else zeros
                grad = tuple(ones if ch == a + 1 else zeros for a in range(D))
                v, g = _channels(call(q, TrialChannels(0, val, grad), 0), (P,), D)
                M[:, 0, ch] += v
                for a in range(D):
                    M[:, a + 1, ch] += g[a]
        self.form = FORM_GENERAL
        self.point_data = np.ascontiguousarray((self.weights[:, None, None] * M).reshape(P, -1))
        return self.form, self.point_data

    def tables(self):
        """Full basis tables at the points: values (P x N box) and physical
        gradients (D, P, N box), axis 1 fastest in the box (cached)."""
        if getattr(self, "_tables", None) is None:
            self._tables = self._build_tables()
        return self._tables

    def _build_tables(self):
        D = self.dim
        vals, ders = [], []
        for a in range(D):
            v, d = _basis_at(self.ref[:, a], self.degrees[a])
            vals.append(v)
            ders.append(d)
        P = self.n_points_total
        phi = np.ones((P, 1))
        for a in range(D):  # box index = i0 + N0 (i1 + N1 i2): axis 0 fastest
            phi = (vals[a][:, :, None] * phi[:, None, :]).reshape(P, -1)
        grads = []
        jinv = self.inverse_jacobian[self.item_of_point]
        for ax in range(D):
            g = np.ones((P, 1))
            for a in range(D):
                f = ders[a] if a == ax else vals[a]
                g = (f[:, :, None] * g[:, None, :]).reshape(P, -1)
            grads.append(g * jinv[:, ax:ax + 1])
        return phi, np.stack(grads)

    def interpolate(self, coefficients_box, derivative_axis=None):
        """(items, N box) -> (P,) values at the points (host)."""
        phi, grads = self.tables()
        t = phi if derivative_axis is None else grads[derivative_axis]
        return np.einsum("pn,pn->p", t, coefficients_box[self.item_of_point])

    def project(self, point_values, derivative_axis=None):
        """(P,) -> (items, N box): sum_q f_q phi_i(q) per item (host)."""
        phi, grads = self.tables()
        t = phi if derivative_axis is None else grads[derivative_axis]
        out = np.zeros((self.cell_count, self.n_modes))
        np.add.at(out, self.item_of_point, t * point_values[:, None])
        return out


def _tangential_rule(shape, axis):
    """Gauss tensor rule over the tangential axes of a face (first tangential
    axis fastest, src/integration.jl:2866-2884): (points (n x D-1), weights)."""
    D = len(shape)
    tang = [b for b in range(D) if b != axis]
    if not tang:
        return np.zeros((1, 0)), np.ones(1)
    rules = [gauss_legendre_rule(int(shape[b])) for b in tang]
    grids = np.meshgrid(*[r[0] for r in rules[::-1]], indexing="ij")
    wgrids = np.meshgrid(*[r[1] for r in rules[::-1]], indexing="ij")
    pts = np.stack([g.reshape(-1) for g in grids[::-1]], axis=1)
    w = np.ones(pts.shape[0])
    for g in wgrids:
        w = w * g.reshape(-1)
    return pts, w


def boundary_face_rule(shape, axis, side):
    """Reference points and weights on one face of the biunit cell
    (src/integration.jl:2846-2862); the face axis of `shape` is ignored."""
    D = len(shape)
    t, w = _tangential_rule(shape, axis)
    pts = np.zeros((t.shape[0], D))
    tang = [b for b in range(D) if b != axis]
    for k, b in enumerate(tang):
        pts[:, b] = t[:, k]
    pts[:, axis] = -1.0 if side == 0 else 1.0
    return PointQuadrature(pts, w)


class InterfaceBucket(_TermLists):
    """Interior interfaces between two leaves (src/integration.jl:2262-2316
    _compile_interface, 2800-2841 _interface_face_points): items share the
    face axis, both sides' degree tuples and the tangential point shape. The
    points are Gauss points of the shared face patch; each side evaluates its
    own basis at its reference coordinates. The point action is lowered by
    probing `interface_accumulate` with the 2(D+1) unit trace channels:
    point_data (P x (2(D+1))^2) = W_q M_q over channels s(D+1) + c."""

    cloud = True
    sides = 2
    form = FORM_GENERAL

    def __init__(self, space, minus_idx, plus_idx, axis, degrees_minus, degrees_plus, tshape):
        dom = space.domain
        self.dim = D = space.dim
        self.axis = axis = int(axis)
        self.minus_idx = np.asarray(minus_idx, dtype=np.int64)
        self.plus_idx = np.asarray(plus_idx, dtype=np.int64)
        self.degrees = tuple(int(p) for p in degrees_minus)
        self.plus_degrees = tuple(int(p) for p in degrees_plus)
        self.modes = tuple(p + 1 for p in self.degrees)
        self.plus_modes = tuple(p + 1 for p in self.plus_degrees)
        self.n_minus = int(np.prod(self.modes))
        self.n_plus = int(np.prod(self.plus_modes))
        self.n_modes = self.n_minus + self.n_plus
        self.cell_count = nc = len(self.minus_idx)
        cm, hm = dom.cell_centers_and_half_sizes(space.leaves[self.minus_idx])
        cp, hp = dom.cell_centers_and_half_sizes(space.leaves[self.plus_idx])
        tang = [b for b in range(D) if b != axis]
        lo = np.maximum(cm - hm, cp - hp)
        up = np.minimum(cm + hm, cp + hp)
        half = 0.5 * (up - lo)
        if tang and np.any(half[:, tang] <= 0):
            raise ValueError("interface leaves do not share a face patch")
        t, w = _tangential_rule(tshape, axis)
        n = t.shape[0]
        self.point_offsets = np.arange(nc + 1, dtype=np.int64) * n
        self.n_points_total = P = nc * n
        self.item_of_point = np.repeat(np.arange(nc), n)
        phys = np.zeros((P, D))
        for k, b in enumerate(tang):
            phys[:, b] = (lo[:, b:b + 1] + half[:, b:b + 1] * (1.0 + t[None, :, k])).reshape(-1)
        phys[:, axis] = np.repeat((cm + hm)[:, axis], n)
        scale = np.prod(half[:, tang], axis=1) if tang else np.ones(nc)
        self.weights = np.repeat(scale, n) * np.tile(w, nc)
        ref = np.zeros((P, 2 * D))
        ip = self.item_of_point
        for b in range(D):
            ref[:, b] = (phys[:, b] - cm[ip, b]) / hm[ip, b]
            ref[:, D + b] = (phys[:, b] - cp[ip, b]) / hp[ip, b]
        ref[:, axis] = 1.0
        ref[:, D + axis] = -1.0
        self.ref = np.ascontiguousarray(np.clip(ref, -1.0, 1.0))
        self.phys_points = [phys[:, b] for b in range(D)]
        self.inverse_jacobian = np.ascontiguousarray(np.concatenate([1.0 / hm, 1.0 / hp], axis=1))
        self.sizes = np.minimum(2.0 * hm, 2.0 * hp)
        self.normals = np.zeros((P, D))
        self.normals[:, axis] = 1.0
        self._gather_maps(space)

    def _gather_maps(self, space):
        class _Side(_TermLists):
            pass

        sides = []
        for idx, modes in ((self.minus_idx, self.modes), (self.plus_idx, self.plus_modes)):
            sd = _Side()
            sd.leaf_idx, sd.cell_count, sd.n_modes = idx, len(idx), int(np.prod(modes))
            Bucket._gather_maps(sd, space)
            sides.append(sd)
        m, p = sides
        nm = len(m.multi_offsets) - 1
        pl = p.local_index.astype(np.int64)
        pl = np.where(pl <= -2, pl - nm, pl)
        self.local_index = np.ascontiguousarray(np.concatenate([m.local_index, pl.astype(np.int32)], axis=1))
        self.local_coefficient = np.ascontiguousarray(np.concatenate([m.local_coefficient, p.local_coefficient], 1))
        self.unit_coefficients = m.unit_coefficients and p.unit_coefficients
        self.multi_offsets = np.concatenate([m.multi_offsets, p.multi_offsets[1:] + m.multi_offsets[-1]])
        self.multi_indices = np.concatenate([m.multi_indices, p.multi_indices]).astype(np.int32)
        self.multi_coefficients = np.concatenate([m.multi_coefficients, p.multi_coefficients])

        def terms(obj):
            obj.term_cell = np.concatenate([m.term_cell, p.term_cell])
            obj.term_box = np.concatenate([m.term_box, p.term_box + obj.n_minus])
            obj.term_dof = np.concatenate([m.term_dof, p.term_dof])
            obj.term_coef = np.concatenate([m.term_coef, p.term_coef])

        self._terms_fn = terms

    def quadrature(self):
        return QuadraturePoints(self.phys_points, self.weights,
                                [self.sizes[self.item_of_point, a] for a in range(self.dim)],
                                normal=[self.normals[:, a] for a in range(self.dim)], face_axis=self.axis)

    def trace_channels(self, result):
        """(value, gradient) arrays of the minus and plus test channels of one callback result."""
        if not isinstance(result, TraceTestChannels):
            raise TypeError("interface callbacks must return TraceTestChannels")
        shape = (self.n_points_total,)
        return [_channels(side, shape, self.dim) for side in (result.minus, result.plus)]

    def lower(self, operators):
        D = self.dim
        C = D + 1
        P = self.n_points_total
        q = self.quadrature()
        ones, zeros = np.ones(P), np.zeros(P)
        Mt = np.zeros((2 * C, 2 * C, P))  # [test channel][trial channel][point]: contiguous fills
        for op in operators:
            for ch in range(2 * C):
                sides = []
                for s in range(2):
                    k = ch - s * C
                    val = ones if k == 0 else zeros
                    grad = tuple(ones if k == a + 1 else zeros for a in range(D))
                    sides.append(TrialChannels(0, val, grad))
                res = op.interface_accumulate(q, TraceTrialChannels(0, sides[0], sides[1]), 0)
                for s, (v, g) in enumerate(self.trace_channels(res)):
                    Mt[s * C, ch] += v
                    for a in range(D):
                        Mt[s * C + a + 1, ch] += g[a]
        Mt *= self.weights[None, None, :]
        self.point_data = np.ascontiguousarray(Mt.transpose(2, 0, 1).reshape(P, -1))
        return self.form, self.point_data

    def tables(self):
        """Per side: full basis values (P x N_s) and physical gradients (D, P, N_s) (cached)."""
        if getattr(self, "_tables", None) is None:
            D = self.dim
            P = self.n_points_total
            out = []
            for s, degs in enumerate((self.degrees, self.plus_degrees)):
                vals, ders = [], []
                for a in range(D):
                    v, d = _basis_at(self.ref[:, s * D + a], degs[a])
                    vals.append(v)
                    ders.append(d)
                phi = np.ones((P, 1))
                for a in range(D):
                    phi = (vals[a][:, :, None] * phi[:, None, :]).reshape(P, -1)
                jinv = self.inverse_jacobian[self.item_of_point, s * D:(s + 1) * D]
                grads = []
                for ax in range(D):
                    g = np.ones((P, 1))
                    for a in range(D):
                        f = ders[a] if a == ax else vals[a]
                        g = (f[:, :, None] * g[:, None, :]).reshape(P, -1)
                    grads.append(g * jinv[:, ax:ax + 1])
                out.append((phi, np.stack(grads)))
            self._tables = out
        return self._tables

    def project_sides(self, channels):
        """[(v, g) minus, (v, g) plus] point channels (weights included) ->
        (items, N_minus + N_plus) local vector (host)."""
        out = np.zeros((self.cell_count, self.n_modes))
        for s, ((phi, grads), (v, g)) in enumerate(zip(self.tables(), channels)):
            r = phi * np.asarray(v)[:, None]
            for a in range(self.dim):
                r = r + grads[a] * np.asarray(g[a])[:, None]
            lo = 0 if s == 0 else self.n_minus
            block = np.zeros((self.cell_count, r.shape[1]))
            np.add.at(block, self.item_of_point, r)
            out[:, lo:lo + r.shape[1]] += block
        return out


class AssemblyPlan:
    """Frozen host tables plus the device plan handle."""

    def __init__(self, problem: AffineProblem):
        self.problem = problem
        space = problem.field.space
        self.space = space
        self.dim = space.dim
        self.full_dof_count = space.dof_count
        keys = {}
        custom = {}
        explicit = getattr(problem, "cell_quadratures", {})
        for leaf in explicit:
            if leaf not in space.leaf_index:
                raise ValueError(f"cell quadrature leaf {leaf} is not an active leaf of the space")
        # Physical domains: automatic cut-cell rules, explicit rules override
        # them (src/integration.jl:2420-2453).
        auto = getattr(space.domain, "assembly_cell_quadratures", None)
        cq = dict(auto(space)) if auto is not None else {}
        cq.update(explicit)
        for li in range(len(space.leaves)):
            leaf = int(space.leaves[li])
            if leaf in cq:
                custom.setdefault(tuple(space.degrees[li]), []).append(li)
                continue
            key = (tuple(space.degrees[li]), tuple(space.quadrature_shapes[li]))
            keys.setdefault(key, []).append(li)
        self.buckets = []
        ops = [op for (_, _, op) in problem.cell_operators]
        for (deg, qs), lis in keys.items():
            b = Bucket(space, lis, deg, qs)
            if ops:
                b.form, b.point_data = b.lower_operators(ops)
            else:
                b.form, b.point_data = FORM_DIFFUSION, np.zeros((b.cell_count, b.n_points))
            self.buckets.append(b)
        # Cells with their own rule (cut cells) and embedded-surface pieces.
        self.cloud_buckets = []
        for deg, lis in custom.items():
            rules = [cq[int(space.leaves[li])] for li in lis]
            cb = CloudBucket(space, lis, deg, rules)
            cb.lower(ops)
            self.cloud_buckets.append(cb)
        sops = [op for (_, _, op) in getattr(problem, "surface_operators", [])]
        surf = {}
        pieces = list(getattr(problem, "surface_quadratures", []))
        for es in getattr(problem, "embedded_surfaces", []):
            pieces.extend(es.surface_quadratures(space))
        for sq in pieces:
            if sq.leaf not in space.leaf_index:
                raise ValueError(f"surface quadrature leaf {sq.leaf} is not an active leaf of the space")
            li = space.leaf_index[sq.leaf]
            surf.setdefault(tuple(space.degrees[li]), []).append((li, sq))
        self.surface_buckets = []
        for deg, items in surf.items():
            sb = CloudBucket(space, [li for li, _ in items], deg, [sq.quadrature for _, sq in items],
                             normals=[sq.normals for _, sq in items])
            sb.lower(sops, surface=True)
            self.surface_buckets.append(sb)
        self._compile_faces(problem)
        self._compile_dirichlet()
        self._compile_mean_constraints()
        self._compile_reduced_map()
        self._rhs = None
        self._device = None

    # Boundary-face and interface items (src/integration.jl:2195-2316).
    def _compile_faces(self, problem):
        space = self.space
        D = self.dim
        self.boundary_buckets = []  # (bucket, bilinear operators, rhs operators)
        sides = {}
        for (bf, _, _, op) in getattr(problem, "boundary_operators", []):
            sides.setdefault((bf.axis, bf.side), ([], []))[0].append(op)
        for (bf, _, op) in getattr(problem, "boundary_rhs", []):
            sides.setdefault((bf.axis, bf.side), ([], []))[1].append(op)
        for (axis, side), (ops, rops) in sorted(sides.items()):
            sel = np.nonzero((space.boundary_axis == axis) & (space.boundary_side == side))[0]
            groups = {}
            for f in sel:
                li = int(space.boundary_leaf[f])
                key = (tuple(space.degrees[li]), tuple(space.quadrature_shapes[li]))
                groups.setdefault(key, []).append(li)
            for (deg, qs), lis in groups.items():
                rule = boundary_face_rule(qs, axis, side)
                normal = np.zeros((rule.point_count(), D))
                normal[:, axis] = -1.0 if side == 0 else 1.0
                bk = CloudBucket(space, lis, deg, [rule] * len(lis), normals=[normal] * len(lis))
                bk.face_axis, bk.face_side = axis, side
                if ops:
                    bk.lower(ops, callback="boundary_accumulate")
                self.boundary_buckets.append((bk, ops, rops))
        self.interface_buckets = []
        iops = [op for (_, _, op) in getattr(problem, "interface_operators", [])]
        irops = [op for (_, op) in getattr(problem, "interface_rhs", [])]
        if (iops or irops) and len(space.interface_minus):
            mi, pi, ax = space.interface_minus, space.interface_plus, space.interface_axis
            qs = space.quadrature_shapes
            # Interfaces of equal sub-face position on both sides (position
            # and extent of the shared patch within each leaf's face, per
            # tangential axis; conforming: 0 and 1) form their own batches: in
            # each, every side's points are one fixed tangential tensor grid.
            cm, hm = space.domain.cell_centers_and_half_sizes(space.leaves[mi])
            cp, hp = space.domain.cell_centers_and_half_sizes(space.leaves[pi])
            lo = np.maximum(cm - hm, cp - hp)
            half = 0.5 * (np.minimum(cm + hm, cp + hp) - lo)
            sub = []
            for c, h in ((cm, hm), (cp, hp)):
                pos = (lo - (c - h)) / (2.0 * h)
                scl = half / h
                pos[np.arange(len(ax)), ax] = 0.0
                scl[np.arange(len(ax)), ax] = 0.0
                sub += [np.rint(pos * 2 ** 24).astype(np.int64), np.rint(scl * 2 ** 24).astype(np.int64)]
            key = np.concatenate([ax[:, None], space.degrees[mi], space.degrees[pi],
                                  np.maximum(qs[mi], qs[pi])] + sub, axis=1)
            uniq, inv = np.unique(key, axis=0, return_inverse=True)
            for g, k in enumerate(uniq):
                sel = np.nonzero(inv.reshape(-1) == g)[0]
                axis = int(k[0])
                tshape = tuple(int(v) for v in k[1 + 2 * D:1 + 3 * D])
                ib = InterfaceBucket(space, mi[sel], pi[sel], axis, k[1:1 + D], k[1 + D:1 + 2 * D], tshape)
                if iops:
                    ib.lower(iops)
                self.interface_buckets.append(ib)
        self.interface_ops, self.interface_rhs_ops = iops, irops

    # Dirichlet data (src/plans.jl:486-514, 884-972).
    def _compile_dirichlet(self):
        """Fixed trace dofs and affine relations. Zero data on every selected
        face fixes every active trace dof to zero (the trace systems built here
        are full rank); other data go through the boundary L2 projection."""
        space = self.space
        cons = self.problem.constraints
        self.affine_rows = []  # (pivot, indices, coefficients, rhs): u_p - sum c_j u_j = rhs
        fixed_dofs, fixed_values = np.zeros(0, dtype=np.int64), np.zeros(0)
        if cons:
            for con in cons:
                sel = (space.boundary_axis == con.boundary.axis) & (space.boundary_side == con.boundary.side)
                if not np.any(sel):
                    raise ValueError("Dirichlet constraint does not match any boundary face")
            dofs = self._dirichlet_trace_dofs(cons)
            if all((not callable(c.data)) and float(c.data) == 0.0 for c in cons):
                fixed_dofs, fixed_values = dofs, np.zeros(len(dofs))
            else:
                fixed_dofs, fixed_values, self.affine_rows = _boundary_l2_projection(self, cons, dofs)
        order = np.argsort(fixed_dofs, kind="stable")
        self.fixed_dofs = np.asarray(fixed_dofs, dtype=np.int64)[order]
        self.fixed_values = np.asarray(fixed_values, dtype=np.float64)[order]
        self.affine_rows.sort(key=lambda row: row[0])

    # Mean-value rows (src/plans.jl:750-805, 976-1003).
    def _compile_mean_constraints(self):
        """One explicit linear row per MeanValue: sum_j c_j u_j = target *
        measure with c_j = integral of the dof's basis function over the
        compiled cells; fixed dofs move to the right-hand side and the pivot is
        the smallest free dof not already eliminated or used."""
        self.mean_rows = []
        means = getattr(self.problem, "mean_constraints", [])
        if not means:
            return
        if self.cloud_buckets:
            raise NotImplementedError("mean-value constraints on plans with cut-cell quadratures are not supported")
        n = self.full_dof_count
        coef = np.zeros(n)
        touched = np.zeros(n, dtype=bool)
        measure = 0.0
        for bk in self.buckets:
            phi = np.ones((1, 1))
            for a in range(bk.dim - 1, -1, -1):  # mode and point boxes are axis-1 fastest
                phi = np.kron(phi, bk.B[a])
            ints = bk.weights @ phi.T  # cells x modes: integral of each mode over each cell
            measure += float(bk.weights.sum())
            np.add.at(coef, bk.term_dof, bk.term_coef * ints[bk.term_cell, bk.term_box])
            touched[bk.term_dof] = True
        fixed_value = dict(zip(self.fixed_dofs.tolist(), self.fixed_values.tolist()))
        blocked = set(self.fixed_dofs.tolist()) | {row[0] for row in self.affine_rows}
        used = set()
        for con in means:
            idx = np.nonzero(touched)[0]
            is_fixed = np.isin(idx, self.fixed_dofs)
            rhs = con.target * measure - float(sum(coef[j] * fixed_value[j] for j in idx[is_fixed].tolist()))
            indices = idx[~is_fixed]
            if len(indices) == 0:
                raise ValueError("mean-value constraint leaves no free dofs")
            free = [j for j in indices.tolist() if j not in blocked]
            if not free:
                raise ValueError("mean-value constraint leaves no free dofs")
            pivot = min(free)
            if pivot in used:
                raise ValueError("mean-value pivots must be unique")
            used.add(pivot)
            self.mean_rows.append((pivot, indices.astype(np.int64), coef[indices].copy(), rhs))

    def _trace_terms(self, con):
        """Terms of every mode with nonzero trace on the constraint's boundary
        side, as arrays (dof, coef, leaf index, box), vectorized over faces."""
        space = self.space
        a, side = con.boundary.axis, con.boundary.side
        sel = np.nonzero((space.boundary_axis == a) & (space.boundary_side == side))[0]
        li = space.boundary_leaf[sel].astype(np.int64)
        lo = space.mode_offsets[li]
        counts = space.mode_offsets[li + 1] - lo
        modes = _expand_ranges(lo, counts)
        leaf = np.repeat(li, counts)
        box = space.mode_box[modes].astype(np.int64)
        shape = space.degrees[leaf].astype(np.int64) + 1
        stride = np.ones(len(modes), dtype=np.int64)
        for b in range(a):
            stride = stride * shape[:, b]
        idx_a = (box // stride) % shape[:, a]
        keep = (shape[:, a] < 2) | (idx_a == side)
        modes, leaf, box = modes[keep], leaf[keep], box[keep]
        tcount = space.term_offsets[modes + 1] - space.term_offsets[modes]
        terms = _expand_ranges(space.term_offsets[modes], tcount)
        return (space.term_indices[terms].astype(np.int64), space.term_coefficients[terms],
                np.repeat(leaf, tcount), np.repeat(box, tcount))

    def _dirichlet_trace_dofs(self, cons):
        dofs = []
        for con in cons:
            dof, coef, _, _ = self._trace_terms(con)
            dofs.append(dof[coef != 0.0])
        return np.unique(np.concatenate(dofs)).astype(np.int64) if dofs else np.zeros(0, dtype=np.int64)

    # Reduced map (src/assembly.jl:112-154, 156-314).
    def _compile_reduced_map(self):
        """u_full = shift + P u_red as a CSR over full dofs. Fixed dofs carry
        their value in the shift; affine-relation and mean-row pivots are
        expanded into the free dofs (reduced) through their rows, the mean rows
        solved together (dense LU) after the Dirichlet relations are
        substituted; the other dofs are the reduced unknowns, in order."""
        n = self.full_dof_count
        fixed = np.zeros(n, dtype=bool)
        fixed[self.fixed_dofs] = True
        eliminated = np.zeros(n, dtype=bool)
        for row in self.affine_rows:
            eliminated[row[0]] = True
        mean_pivot = np.zeros(n, dtype=bool)
        for row in self.mean_rows:
            mean_pivot[row[0]] = True
        constrained = fixed | eliminated | mean_pivot
        self.solve_dofs = np.nonzero(~constrained)[0].astype(np.int64)
        self.reduced_index = np.full(n, -1, dtype=np.int64)
        self.reduced_index[self.solve_dofs] = np.arange(len(self.solve_dofs))
        self.reduced_dof_count = len(self.solve_dofs)
        self.shift = np.zeros(n)
        self.shift[self.fixed_dofs] = self.fixed_values
        counts = (~constrained).astype(np.int64)
        pivots = np.nonzero(eliminated | mean_pivot)[0]
        if len(pivots) == 0:
            self.map_offsets = np.zeros(n + 1, dtype=np.int64)
            self.map_offsets[1:] = np.cumsum(counts)
            self.map_indices = np.arange(self.reduced_dof_count, dtype=np.int64)
            self.map_coefficients = np.ones(self.reduced_dof_count)
            self.selection_map = True
            return
        expansions = _reduced_expansions(self, fixed)
        rows = {}
        for d in pivots.tolist():
            shift, red = expansions[d]
            self.shift[d] = shift
            keys = sorted(k for k, v in red.items() if v != 0.0)
            rows[d] = (np.asarray(keys, dtype=np.int64), np.asarray([red[k] for k in keys]))
            counts[d] = len(keys)
        self.map_offsets = np.zeros(n + 1, dtype=np.int64)
        self.map_offsets[1:] = np.cumsum(counts)
        self.map_indices = np.zeros(int(self.map_offsets[-1]), dtype=np.int64)
        self.map_coefficients = np.zeros(int(self.map_offsets[-1]))
        start = self.map_offsets[self.solve_dofs]
        self.map_indices[start] = np.arange(self.reduced_dof_count)
        self.map_coefficients[start] = 1.0
        for d, (keys, vals) in rows.items():
            lo = int(self.map_offsets[d])
            self.map_indices[lo:lo + len(keys)] = keys
            self.map_coefficients[lo:lo + len(keys)] = vals
        self.selection_map = False

    def expand_host(self, x_red, include_shift=True):
        """shift + P x_red on the host (src/assembly.jl:851-874)."""
        x_red = np.asarray(x_red, dtype=np.float64)
        rows = np.repeat(np.arange(self.full_dof_count), np.diff(self.map_offsets))
        out = self.shift.copy() if include_shift else np.zeros(self.full_dof_count)
        np.add.at(out, rows, self.map_coefficients * x_red[self.map_indices])
        return out

    def project_host(self, y_full):
        """P^T y on the host (src/assembly.jl:876-891)."""
        y_full = np.asarray(y_full, dtype=np.float64)
        rows = np.repeat(np.arange(self.full_dof_count), np.diff(self.map_offsets))
        out = np.zeros(self.reduced_dof_count)
        np.add.at(out, self.map_indices, self.map_coefficients * y_full[rows])
        return out

    # Host-side right-hand side (src/assembly.jl:488-507 rhs!).
    @property
    def all_buckets(self):
        """Every batch the device applies: tensor cells, point-cloud cells,
        embedded surfaces, boundary faces and interfaces with an operator."""
        faces = [bk for (bk, ops, _) in self.boundary_buckets if ops]
        faces += list(self.interface_buckets) if self.interface_ops else []
        # no cell action next to face or surface batches: the cell batches
        # would apply zeros and are left out
        cells = list(self.buckets) + list(self.cloud_buckets) if self.device_cells else []
        return cells + list(self.surface_buckets) + faces

    def rhs_physical(self):
        """Host right-hand side: cell loads on tensor and cut cells, surface
        loads on embedded pieces (src/assembly.jl:488-507, 586-601)."""
        if self._rhs is None:
            b = np.zeros(self.full_dof_count)
            groups = [(self.problem.cell_rhs, list(self.buckets) + list(self.cloud_buckets), False),
                      (getattr(self.problem, "surface_rhs", []), list(self.surface_buckets), True)]
            for ops, buckets, surface in groups:
                for (_, op) in ops:
                    for bk in buckets:
                        cloud = getattr(bk, "cloud", False)
                        shape = (bk.n_points_total,) if cloud else (bk.cell_count, bk.n_points)
                        call = op.surface_rhs_accumulate if surface else op.cell_rhs_accumulate
                        v, g = _channels(call(bk.quadrature(), 0), shape, bk.dim)
                        local = bk.project(bk.weights * v)
                        for a in range(bk.dim):
                            if np.any(g[a] != 0):
                                local = local + bk.project(bk.weights * g[a], derivative_axis=a)
                        contrib = bk.term_coef * local[bk.term_cell, bk.term_box]
                        b += np.bincount(bk.term_dof, weights=contrib, minlength=self.full_dof_count)
            # Boundary-face loads (src/accumulators.jl:1230-1274) and interface
            # loads (1666-1698): both traces projected.
            for bk, _, rops in self.boundary_buckets:
                for op in rops:
                    v, g = _channels(op.boundary_rhs_accumulate(bk.quadrature(), 0), (bk.n_points_total,), bk.dim)
                    local = bk.project(bk.weights * v)
                    for a in range(bk.dim):
                        if np.any(g[a] != 0):
                            local = local + bk.project(bk.weights * g[a], derivative_axis=a)
                    b += np.bincount(bk.term_dof, weights=bk.term_coef * local[bk.term_cell, bk.term_box],
                                     minlength=self.full_dof_count)
            for bk in self.interface_buckets:
                for op in self.interface_rhs_ops:
                    ch = bk.trace_channels(op.interface_rhs_accumulate(bk.quadrature(), 0))
                    ch = [(bk.weights * v, [bk.weights * ga for ga in g]) for v, g in ch]
                    local = bk.project_sides(ch)
                    b += np.bincount(bk.term_dof, weights=bk.term_coef * local[bk.term_cell, bk.term_box],
                                     minlength=self.full_dof_count)
            self._rhs = b
        return self._rhs

    def lowered_loads(self):
        """Point loads of the tensor-product buckets for the device right-hand
        side: (flat array, channels), per cell and point (W v) or
        (W v, W g_1, .., W g_D), cells in bucket order; the load callbacks
        evaluated as rhs_physical does."""
        if getattr(self, "_loads", None) is None:
            D = self.dim
            per = []
            grad = False
            for bk in self.buckets:
                shape = (bk.cell_count, bk.n_points)
                v = np.zeros(shape)
                g = [np.zeros(shape) for _ in range(D)]
                for (_, op) in self.problem.cell_rhs:
                    vv, gg = _channels(op.cell_rhs_accumulate(bk.quadrature(), 0), shape, D)
                    v = v + bk.weights * vv
                    for a in range(D):
                        g[a] = g[a] + bk.weights * gg[a]
                grad = grad or any(np.any(ga != 0) for ga in g)
                per.append((v, g))
            channels = 1 + D if grad else 1
            flat = []
            for v, g in per:
                arr = np.stack([v] + (g if grad else []), axis=-1)
                flat.append(arr.reshape(-1))
            self._loads = (np.concatenate(flat) if flat else np.zeros(0), channels)
        return self._loads

    def _l2_layout(self):
        """Batches of the device verification sums (gmf_l2_error) in device
        bucket order: (bucket, counted) for the tensor-product cells and every
        one-sided point-cloud batch; only cells (tensor and cut) are counted."""
        if not self.device_cells:
            raise ValueError("the device plan holds no cell batches (no cell operator next to face or surface "
                             "batches); use the host l2_error")
        cut = {id(bk) for bk in self.cloud_buckets}
        out = []
        for bk in self.all_buckets:
            if not getattr(bk, "cloud", False):
                out.append((bk, True))
            elif getattr(bk, "sides", 1) == 1:
                out.append((bk, id(bk) in cut))
        return out

    def point_weights(self):
        """Quadrature weights per point in the gmf_l2_error layout: tensor
        cells, then one-sided point-cloud batches (zeros for surface pieces and
        boundary faces)."""
        parts = [bk.weights.reshape(-1) if counted else np.zeros(bk.n_points_total)
                 for bk, counted in self._l2_layout()]
        return np.concatenate(parts) if parts else np.zeros(0)

    def point_values(self, fn):
        """fn(point tuple) per point in the gmf_l2_error layout (zeros where
        point_weights is zero)."""
        parts = []
        for bk, counted in self._l2_layout():
            if not counted:
                parts.append(np.zeros(bk.n_points_total))
                continue
            shape = (bk.n_points_total,) if getattr(bk, "cloud", False) else (bk.cell_count, bk.n_points)
            parts.append(np.broadcast_to(np.asarray(fn(tuple(bk.phys_points)), dtype=np.float64), shape).reshape(-1))
        return np.concatenate(parts) if parts else np.zeros(0)

    @property
    def device_cells(self):
        """True when the device plan holds the cell batches (a cell operator,
        or no face and surface batches beside them)."""
        faces = [bk for (bk, ops, _) in self.boundary_buckets if ops]
        faces += list(self.interface_buckets) if self.interface_ops else []
        return bool(self.problem.cell_operators) or not (faces or self.surface_buckets)

    @property
    def device_rhs(self):
        """True when every load lives on a batch the device plan holds: tensor
        and cut cells (the cell batches are on the device), embedded-surface
        pieces, and boundary faces / interfaces with an operator."""
        if self.problem.cell_rhs and not self.device_cells:
            return False  # the device plan holds no cell batches
        if self.problem.cell_rhs and not self.problem.cell_operators and (self.boundary_buckets or
                                                                          self.interface_buckets):
            return False  # load-only faces make the host and device plans differ
        if any(rops and not ops for (_, ops, rops) in self.boundary_buckets):
            return False
        if self.interface_rhs_ops and not self.interface_ops:
            return False
        return True

    def face_loads(self):
        """Point loads of the device plan's point-cloud batches, in its bucket
        order: per item, point and side (W v, W g_1, .., W g_D) (gmf_rhs_faces):
        cut-cell loads (cell_rhs_accumulate), embedded-surface loads
        (surface_rhs_accumulate), boundary-face and interface loads."""
        D = self.dim
        rops_of = {id(bk): rops for (bk, _, rops) in self.boundary_buckets}
        cut = {id(bk) for bk in self.cloud_buckets}
        surface = {id(bk) for bk in self.surface_buckets}
        out = []
        for bk in self.all_buckets:
            if not getattr(bk, "cloud", False):
                continue
            P = bk.n_points_total
            q = bk.quadrature()
            if getattr(bk, "sides", 1) == 2:
                arr = np.zeros((P, 2, D + 1))
                for op in self.interface_rhs_ops:
                    for s, (v, g) in enumerate(bk.trace_channels(op.interface_rhs_accumulate(q, 0))):
                        arr[:, s, 0] += bk.weights * v
                        for a in range(D):
                            arr[:, s, a + 1] += bk.weights * g[a]
            else:
                arr = np.zeros((P, 1, D + 1))
                if id(bk) in cut:
                    calls = [op.cell_rhs_accumulate for (_, op) in self.problem.cell_rhs]
                elif id(bk) in surface:
                    calls = [op.surface_rhs_accumulate for (_, op) in getattr(self.problem, "surface_rhs", [])]
                else:
                    calls = [op.boundary_rhs_accumulate for op in rops_of.get(id(bk), [])]
                for call in calls:
                    v, g = _channels(call(q, 0), (P,), D)
                    arr[:, 0, 0] += bk.weights * v
                    for a in range(D):
                        arr[:, 0, a + 1] += bk.weights * g[a]
            out.append(arr.reshape(-1))
        return np.concatenate(out) if out else np.zeros(0)

    @property
    def tensor_only(self):
        """True when every operator and load lives on tensor-product cells
        (the device right-hand side covers exactly those)."""
        return not (self.cloud_buckets or self.surface_buckets or self.boundary_buckets or self.interface_buckets)

    def rhs(self):
        """Full-layout right-hand side with the constraint rows' values
        (fixed values, relation and mean right-hand sides)."""
        b = self.rhs_physical().copy()
        b[self.fixed_dofs] = self.fixed_values
        for row in self.affine_rows + self.mean_rows:
            b[row[0]] = row[3]
        return b

    def reduced_rhs(self):
        """P^T (b - A shift) (src/assembly.jl:917-933), A shift evaluated on the device."""
        b = self.rhs_physical().copy()
        if np.any(self.shift != 0):
            from . import ops
            b -= ops.apply_physical(self, self.shift)
        return b[self.solve_dofs] if self.selection_map else self.project_host(b)

    # Device plan.
    def device(self):
        if self._device is None:
            self._device = _devabi.DevicePlan(self)
        return self._device


def _boundary_l2_projection(plan, cons, dofs):
    """L2 projection of g onto the active trace space (src/plans.jl:884-972):
    (fixed dofs, their values, affine relations). Per boundary face the basis
    values are tensor products of per-axis tables at the face's Gauss points
    and the face's Gram block is Phi W Phi^T; the global Gram matrix is sparse
    (face blocks). Small systems take the reference's dense spectral route
    (_dirichlet_relations), larger ones SciPy's sparse LU (full rank only)."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla

    from .quadrature import fe_basis_values_and_derivatives

    space = plan.space
    n = len(dofs)
    pos = np.full(plan.full_dof_count, -1, dtype=np.int64)
    pos[dofs] = np.arange(n)
    rows, cols, vals = [], [], []
    r = np.zeros(n)
    dom = space.domain
    tables = {}
    for con in cons:
        t_dof, t_coef, t_leaf, t_box = plan._trace_terms(con)
        by_face = {}
        for dof, coef, li, box in zip(t_dof.tolist(), t_coef.tolist(), t_leaf.tolist(), t_box.tolist()):
            by_face.setdefault(li, []).append((dof, coef, box))
        a = con.boundary.axis
        side = -1.0 if con.boundary.side == 0 else 1.0
        for li, items in by_face.items():
            deg = space.degrees[li]
            qs = space.quadrature_shapes[li]
            cell = space.leaves[li]
            centers, halves = dom.cell_centers_and_half_sizes([cell])
            tang = [ax for ax in range(space.dim) if ax != a]
            rules = [gauss_legendre_rule(int(qs[ax])) for ax in tang]
            grids = np.meshgrid(*[rl[0] for rl in rules], indexing="ij") if tang else []
            wgrid = np.meshgrid(*[rl[1] for rl in rules], indexing="ij") if tang else []
            npts = int(np.prod([int(qs[ax]) for ax in tang])) if tang else 1
            w = np.ones(npts)
            for wg in wgrid:
                w = w * wg.reshape(-1)
            w = w * float(np.prod([halves[0, ax] for ax in tang])) if tang else w
            xi = {}
            for k, ax in enumerate(tang):
                xi[ax] = grids[k].reshape(-1)
            xi[a] = np.full(npts, side)
            pts = tuple(centers[0, ax] + xi[ax] * halves[0, ax] for ax in range(space.dim))
            gval = np.broadcast_to(np.asarray(con.data(pts) if callable(con.data) else con.data,
                                              dtype=np.float64), (npts,))
            # per-axis basis tables (npts x (p+1)), shared by faces of equal shape
            tab = []
            for ax in range(space.dim):
                key = (a, side, tuple(int(q) for q in qs), ax, int(deg[ax]))
                if key not in tables:
                    tables[key] = fe_basis_values_and_derivatives(xi[ax], int(deg[ax]))[0]
                tab.append(tables[key])
            shape = tuple(int(p) + 1 for p in deg)
            dof = np.array([it[0] for it in items], dtype=np.int64)
            coef = np.array([it[1] for it in items])
            idx = np.unravel_index(np.array([it[2] for it in items], dtype=np.int64), shape[::-1])[::-1]
            phi_t = np.ones((npts, len(items)))
            for ax in range(space.dim):
                phi_t = phi_t * tab[ax][:, idx[ax]]
            phi_t = phi_t * coef[None, :]
            u, inv = np.unique(dof, return_inverse=True)
            phi = np.zeros((len(u), npts))
            np.add.at(phi, inv, phi_t.T)
            gl = (phi * w[None, :]) @ phi.T
            rl = phi @ (w * gval)
            ii = pos[u]
            rows.append(np.repeat(ii, len(u)))
            cols.append(np.tile(ii, len(u)))
            vals.append(gl.reshape(-1))
            np.add.at(r, ii, rl)
    if n == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(0), []
    G = sps.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    if n <= DENSE_DIRICHLET_LIMIT:
        return _dirichlet_relations(dofs, G.toarray(), r)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("error", spla.MatrixRankWarning)
        try:
            sol = np.asarray(spla.spsolve(G, r), dtype=np.float64)
        except spla.MatrixRankWarning:
            sol = None
    # A singular trace Gram matrix means trace dofs that the selected faces do
    # not control independently; above the dense limit the relations the
    # reference builds for them (src/plans.jl:884-943) are not computed.
    resid = np.inf if sol is None or not np.all(np.isfinite(sol)) else \
        np.linalg.norm(G @ sol - r) / max(np.linalg.norm(r), np.finfo(np.float64).tiny)
    if not resid <= 1e-8:
        raise ValueError(f"rank-deficient Dirichlet trace system with {n} dofs: affine Dirichlet relations are "
                         f"built for at most {DENSE_DIRICHLET_LIMIT} trace dofs")
    return np.asarray(dofs, dtype=np.int64), sol, []


# Trace systems up to this size take the reference's dense spectral route
# (eigen decomposition, rank cut, column-pivoted QR); larger ones are solved
# sparsely and must be full rank.
DENSE_DIRICHLET_LIMIT = 2000


def _dirichlet_relations(dofs, G, r):
    """Fixed dofs and affine relations from the trace Gram system G a = r
    (src/plans.jl:884-972). Positive eigenmodes above 1000 eps scale n define
    the independent constraints; the data must be consistent on the others.
    Column-pivoted QR of the constraint matrix picks one pivot dof per
    constraint; the pivots become fixed dofs when no free trace dof couples to
    them above the rank tolerance, else relations u_p - sum c_j u_j = value."""
    import scipy.linalg as sla
    dofs = np.asarray(dofs, dtype=np.int64)
    n = len(dofs)
    lam, V = np.linalg.eigh(0.5 * (G + G.T))
    order = np.argsort(lam)[::-1]
    lam, V = lam[order], V[:, order]
    scale = float(np.max(np.abs(lam))) if n else 0.0
    tol = 0.0 if scale == 0.0 else 1000.0 * np.finfo(np.float64).eps * scale * max(n, 1)
    k = int(np.count_nonzero(lam > tol))
    tr = V.T @ r
    if np.any(np.abs(tr[k:]) > tol * max(np.linalg.norm(r), 1.0)):
        raise ValueError("Dirichlet data is inconsistent on the selected boundary faces")
    if k == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(0), []
    C = V[:, :k].T
    targets = tr[:k] / lam[:k]
    _, _, perm = sla.qr(C, mode="economic", pivoting=True)
    pivot_cols = perm[:k]
    is_pivot = np.zeros(n, dtype=bool)
    is_pivot[pivot_cols] = True
    free_cols = np.nonzero(~is_pivot)[0]
    Pm = C[:, pivot_cols]
    values = np.linalg.solve(Pm, targets)
    if len(free_cols) == 0:
        return dofs[pivot_cols], values, []
    coupling = -np.linalg.solve(Pm, C[:, free_cols])
    fixed_d, fixed_v, rows = [], [], []
    for i in range(k):
        keep = np.abs(coupling[i]) > tol
        if not np.any(keep):
            fixed_d.append(int(dofs[pivot_cols[i]]))
            fixed_v.append(float(values[i]))
        else:
            rows.append((int(dofs[pivot_cols[i]]), dofs[free_cols[keep]].astype(np.int64),
                         coupling[i][keep].copy(), float(values[i])))
    return np.asarray(fixed_d, dtype=np.int64), np.asarray(fixed_v), rows


def _reduced_expansions(plan, fixed):
    """(shift, {reduced index: coefficient}) of every relation and mean pivot
    (src/assembly.jl:156-270): relation pivots expand through their rows
    (recursively, cycles rejected); the mean rows, with relation pivots
    substituted, form a small dense system for the mean pivots, solved once for
    the shift and once per reduced column that appears."""
    red_index = plan.reduced_index
    fixed_value = np.zeros(plan.full_dof_count)
    fixed_value[plan.fixed_dofs] = plan.fixed_values
    rows = {row[0]: row for row in plan.affine_rows}
    mean_of = {row[0]: i for i, row in enumerate(plan.mean_rows)}
    base_cache, visiting = {}, set()

    def base(d):
        """(shift, reduced {}, mean {}) with mean pivots kept symbolic."""
        if d in base_cache:
            return base_cache[d]
        if d in visiting:
            raise ValueError("cyclic Dirichlet constraint relation detected")
        if fixed[d]:
            out = (fixed_value[d], {}, {})
        elif red_index[d] >= 0:
            out = (0.0, {int(red_index[d]): 1.0}, {})
        elif d in mean_of:
            out = (0.0, {}, {mean_of[d]: 1.0})
        elif d in rows:
            visiting.add(d)
            _, idx, cf, rhs = rows[d]
            shift, red, mean = rhs, {}, {}
            for j, c in zip(idx.tolist(), cf.tolist()):
                s2, r2, m2 = base(j)
                shift += c * s2
                for key, v in r2.items():
                    red[key] = red.get(key, 0.0) + c * v
                for key, v in m2.items():
                    mean[key] = mean.get(key, 0.0) + c * v
            visiting.discard(d)
            out = (shift, red, mean)
        else:
            raise ValueError(f"dof {d} has no reduced representation")
        base_cache[d] = out
        return out

    m = len(plan.mean_rows)
    mean_shift = np.zeros(m)
    mean_red = [{} for _ in range(m)]
    if m:
        A = np.zeros((m, m))
        rhs = np.zeros(m)
        cols = {}
        for i, (_, idx, cf, target) in enumerate(plan.mean_rows):
            rhs[i] = target
            for j, c in zip(idx.tolist(), cf.tolist()):
                s2, r2, m2 = base(j)
                rhs[i] -= c * s2
                for key, v in m2.items():
                    A[i, key] += c * v
                for key, v in r2.items():
                    cols.setdefault(key, np.zeros(m))[i] -= c * v
        import scipy.linalg as sla
        lu = sla.lu_factor(A)
        mean_shift = sla.lu_solve(lu, rhs)
        for key, col in cols.items():
            sol = sla.lu_solve(lu, col)
            for i in range(m):
                mean_red[i][key] = mean_red[i].get(key, 0.0) + sol[i]
    final_cache = {}

    def final(d):
        if d in final_cache:
            return final_cache[d]
        if fixed[d]:
            out = (fixed_value[d], {})
        elif red_index[d] >= 0:
            out = (0.0, {int(red_index[d]): 1.0})
        elif d in mean_of:
            i = mean_of[d]
            out = (float(mean_shift[i]), dict(mean_red[i]))
        else:
            _, idx, cf, rhs = rows[d]
            shift, red = rhs, {}
            for j, c in zip(idx.tolist(), cf.tolist()):
                s2, r2 = final(j)
                shift += c * s2
                for key, v in r2.items():
                    red[key] = red.get(key, 0.0) + c * v
            out = (shift, red)
        final_cache[d] = out
        return out

    return {d: final(d) for d in list(rows) + list(mean_of)}


def compile(problem):
    if not isinstance(problem, AffineProblem):
        raise TypeError("compile expects an AffineProblem")
    return AssemblyPlan(problem)
