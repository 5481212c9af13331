"""Manual hp adaptivity, source-to-target transitions and state transfer.

Reference: src/adaptivity.jl (AdaptivityPlan 351-415, request_h_refinement!
922-940, request_p_refinement! 972-993, request_p_derefinement! 995-1017,
_apply_p_refinement!/_apply_p_derefinement! 459-493: refined children inherit
the leaf's degree tuple, degree edits are checked against max_p / min_p),
src/transition.jl (transition 349-387, source_leaves 143) and src/transfer.jl
(transfer_state 892-1000: the target coefficients are the L2 projection of the
old field on the target space, with the projection's load integrated over
source/target overlaps).

Here the target grid is a copy of the source grid with the requested splits
applied; a merged leaf keeps its children in the grid and is listed as a leaf
of the target space. The target mass operator (value channel only, no
constraints) is compiled like any other cell accumulator and the load
u_old(x_q) is evaluated on the host at the target quadrature points (located
in the source grid by descending its refinement tree). The target quadrature
takes max(p_source, p_target) + 1 Gauss points per axis, which integrates the
mass matrix and the load exactly, so a target space that contains the source
space reproduces the old field up to the solve tolerance.

transfer_state without a linear_solve policy follows the reference's default
(src/transfer.jl:431-437, 637-804): one local L2 projection per target cell
over the leaf's basis modes, reconciled on shared dofs through the normal
equations of the local relations (host). With a linear_solve policy the
global L2 projection is solved on the device by that solver (the reference's
variational path).

A merged (h-derefined) target leaf covers several source leaves; its cell
takes a composite rule, one Gauss rule per covered source leaf, so the
piecewise polynomial old field is integrated leaf by leaf as in
src/transfer.jl:22-25, and the cell runs on the point-cloud kernel.

Interface jump and modal coefficient indicators (src/indicators.jl:260-342)
are evaluated on the host. Not built: the projection and multiresolution
indicators and the automatic planner of src/indicators.jl.
"""

import numpy as np

from .mesh import Domain
from .problem import AffineProblem, PointQuadrature, ScalarField, SPD, TestChannels
from .quadrature import gauss_legendre_rule
from .space import DegreePlusQuadrature, HpSpace, SpaceOptions


class AdaptivityLimits:
    """Bounds on h levels and degrees (src/adaptivity.jl:134-172)."""

    def __init__(self, dimension, max_h_level=None, min_p=1, max_p=None):
        D = int(dimension)
        self.dim = D
        self.max_h_level = _tuple(max_h_level if max_h_level is not None else 2**30, D)
        self.min_p = _tuple(min_p, D)
        self.max_p = _tuple(max_p if max_p is not None else 2**30, D)


def _tuple(value, D):
    if np.isscalar(value):
        return (int(value),) * D
    value = tuple(int(v) for v in value)
    if len(value) != D:
        raise ValueError(f"expected {D} entries, got {len(value)}")
    return value


def copy_domain(domain):
    """A new Domain with the same roots and the same refinement tree (cell ids
    preserved: splits are replayed in the order their children were created)."""
    out = Domain(domain.origin, domain.extent, domain.grid.root_counts)
    fc = domain.grid.first_child()
    split = domain.grid.split_axis()
    refined = np.nonzero(fc >= 0)[0]
    for cell in refined[np.argsort(fc[refined], kind="stable")]:
        first = out.refine(int(cell), int(split[cell]))
        if first != fc[cell]:
            raise RuntimeError("grid replay produced different cell ids")
    return out


class AdaptivityPlan:
    """Target leaves and degree tuples derived from a source space
    (src/adaptivity.jl:351-415). Leaves are cell ids of the target grid;
    before any split they coincide with the source space's leaves."""

    def __init__(self, space, limits=None):
        self.source_space = space
        self.dim = D = space.dim
        self.limits = limits or AdaptivityLimits(D)
        self.target_domain = copy_domain(space.domain)
        self.degrees = {int(l): tuple(int(p) for p in space.degrees[i]) for i, l in enumerate(space.leaves)}

    def _checked_leaf(self, leaf):
        leaf = int(leaf)
        if leaf not in self.degrees:
            raise ValueError(f"leaf {leaf} is not an active target leaf")
        return leaf

    def _checked_axis(self, axis):
        axis = int(axis)
        if not 0 <= axis < self.dim:
            raise ValueError(f"axis {axis} out of range for dimension {self.dim}")
        return axis

    def cell_degrees(self, leaf):
        return self.degrees[self._checked_leaf(leaf)]

    def is_empty(self):
        """True when the target equals the source (src/adaptivity.jl:435-447)."""
        src = self.source_space
        if self.target_domain.grid.cell_count() != src.domain.grid.cell_count():
            return False
        return all(self.degrees[int(l)] == tuple(int(p) for p in src.degrees[i]) for i, l in enumerate(src.leaves))

    def request_h_refinement(self, leaf, axis):
        """Split an active target leaf along `axis`; both children inherit its
        degrees (src/adaptivity.jl:889-940). Returns the first child. A cell
        that the grid already splits along `axis` (a leaf merged earlier)
        gets its existing children back."""
        leaf = self._checked_leaf(leaf)
        axis = self._checked_axis(axis)
        grid = self.target_domain.grid
        level = int(grid.levels()[leaf, axis])
        if level + 1 > self.limits.max_h_level[axis]:
            raise ValueError(f"leaf {leaf} reached max_h_level[{axis}]")
        first = int(grid.first_child()[leaf])
        if first >= 0 and int(grid.split_axis()[leaf]) != axis:
            raise NotImplementedError(f"cell {leaf} is already split along axis {int(grid.split_axis()[leaf])}")
        deg = self.degrees.pop(leaf)
        if first < 0:
            first = self.target_domain.refine(leaf, axis)
        self.degrees[first] = deg
        self.degrees[first + 1] = deg
        return first

    def request_h_derefinement(self, cell, axis):
        """Merge the two active children of `cell` split along `axis` back into
        one leaf; it takes the componentwise maximum of the child degrees
        (src/adaptivity.jl:942-970)."""
        cell = int(cell)
        axis = self._checked_axis(axis)
        grid = self.target_domain.grid
        if not 0 <= cell < grid.cell_count():
            raise ValueError(f"cell {cell} out of range")
        first = int(grid.first_child()[cell])
        if first < 0 or int(grid.split_axis()[cell]) != axis:
            raise ValueError(f"cell {cell} is not split along axis {axis}")
        children = (first, first + 1)
        for c in children:
            if c not in self.degrees:
                raise ValueError(f"child {c} of cell {cell} is not an active target leaf")
        merged = tuple(max(a, b) for a, b in zip(self.degrees[children[0]], self.degrees[children[1]]))
        for c in children:
            del self.degrees[c]
        self.degrees[cell] = merged
        return self

    def request_p_refinement(self, leaf, axis=None, increment=1):
        """Raise the degree of one leaf along `axis` (all axes if None) by
        `increment`, bounded by max_p (src/adaptivity.jl:459-474, 972-993)."""
        return self._change_degree(leaf, axis, int(increment), "increment")

    def request_p_derefinement(self, leaf, axis=None, decrement=1):
        """Lower the degree of one leaf, bounded by min_p (src/adaptivity.jl:476-493)."""
        return self._change_degree(leaf, axis, -int(decrement), "decrement")

    def _change_degree(self, leaf, axis, step, name):
        leaf = self._checked_leaf(leaf)
        if (step > 0) != (name == "increment") or step == 0:
            raise ValueError(f"{name} must be positive")
        axes = range(self.dim) if axis is None else (self._checked_axis(axis),)
        deg = list(self.degrees[leaf])
        for a in axes:
            new = deg[a] + step
            if new > self.limits.max_p[a]:
                raise ValueError(f"leaf {leaf} reached max_p[{a}]")
            if new < self.limits.min_p[a]:
                raise ValueError(f"leaf {leaf} reached min_p[{a}]")
            deg[a] = new
        self.degrees[leaf] = tuple(deg)
        return self

    def target_leaves(self):
        return np.asarray(sorted(self.degrees), dtype=np.int64)


class SpaceTransition:
    """Source space, target space and the source leaf holding each target
    leaf (src/transition.jl:349-387, source_leaves 143)."""

    def __init__(self, source_space, target_space, source_leaves_of_target):
        self.source_space = source_space
        self.target_space = target_space
        self.source_leaves_of_target = source_leaves_of_target

    def source_leaves(self, target_leaf):
        """Source leaves overlapping one target leaf: its ancestor-or-self
        source leaf, or the source leaves below a merged cell."""
        return list(self.source_leaves_of_target[int(target_leaf)])


def transition(plan):
    """Compile the target space of an AdaptivityPlan with the source space's
    basis and continuity (src/transition.jl:349-387)."""
    src = plan.source_space
    leaves = plan.target_leaves()
    degrees = np.asarray([plan.degrees[int(l)] for l in leaves], dtype=np.int32).reshape(len(leaves), plan.dim)
    options = SpaceOptions(basis=src.options.basis, degree=src.options.degree,
                           quadrature=src.options.quadrature, continuity=src.options.continuity)
    target = HpSpace(plan.target_domain, options, leaves=leaves, degrees=degrees)
    # Ancestor-or-self walk in the target grid; merged cells collect the
    # source leaves below them.
    parent = plan.target_domain.grid.parent()
    first_child = plan.target_domain.grid.first_child()
    source_set = set(int(l) for l in src.leaves)
    owner = {}
    for l in leaves:
        c = int(l)
        while c >= 0 and c not in source_set:
            c = int(parent[c])
        if c >= 0:
            owner[int(l)] = (c,)
            continue
        found, stack = [], [int(l)]
        while stack:
            c = stack.pop()
            if c in source_set:
                found.append(c)
            elif first_child[c] >= 0:
                stack.extend((int(first_child[c]) + 1, int(first_child[c])))
            else:
                raise RuntimeError(f"target leaf {int(l)} is not covered by source leaves")
        owner[int(l)] = tuple(found)
    return SpaceTransition(src, target, owner)


# ---------------------------------------------------------------------------
# Evaluation of a state at arbitrary points of its space.

def integrated_legendre(x, degree):
    """psi_0..psi_p at each x (shape (len(x), p+1)): psi_0 = (1-x)/2,
    psi_1 = (1+x)/2, psi_n = (P_n - P_{n-2}) / sqrt(4n-2); degree 0 gives the
    constant 1 (src/polynomials.jl:403-434, 238-247)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    out = np.empty((x.size, degree + 1))
    if degree == 0:
        out[:, 0] = 1.0
        return out
    out[:, 0] = 0.5 * (1.0 - x)
    out[:, 1] = 0.5 * (1.0 + x)
    if degree >= 2:
        P = [np.ones_like(x), x.copy()]
        for n in range(2, degree + 1):
            P.append(((2 * n - 1) * x * P[n - 1] - (n - 1) * P[n - 2]) / n)
            out[:, n] = (P[n] - P[n - 2]) / np.sqrt(4.0 * n - 2.0)
    return out


def locate(domain, points, leaves=None):
    """Leaf containing each point (tuple of D arrays of one shape), by
    descending the refinement tree from the root cell; the descent stops at
    the cells listed in `leaves` (default: the grid's active leaves)."""
    D = domain.dim
    pts = [np.asarray(p, dtype=np.float64).reshape(-1) for p in points]
    grid = domain.grid
    lev, crd, fc, split = grid.levels(), grid.coords(), grid.first_child(), grid.split_axis()
    roots = np.nonzero(grid.parent() < 0)[0]
    counts = grid.root_counts
    root_of = np.full(int(np.prod(counts)), -1, dtype=np.int64)
    flat = np.zeros(len(roots), dtype=np.int64)
    for a in reversed(range(D)):
        flat = flat * counts[a] + crd[roots, a]
    root_of[flat] = roots
    idx = np.zeros(pts[0].size, dtype=np.int64)
    for a in reversed(range(D)):
        h = domain.extent[a] / counts[a]
        i = np.clip(np.floor((pts[a] - domain.origin[a]) / h).astype(np.int64), 0, counts[a] - 1)
        idx = idx * counts[a] + i
    cell = root_of[idx]
    stop = fc < 0
    if leaves is not None:
        stop = np.zeros(fc.size, dtype=bool)
        stop[np.asarray(leaves, dtype=np.int64)] = True
    while True:
        inner = np.nonzero(~stop[cell])[0]
        if inner.size == 0:
            return cell
        c = cell[inner]
        if np.any(fc[c] < 0):
            raise ValueError("point outside the leaves of the space")
        a = split[c]
        k = np.arange(c.size)
        mid = domain.cell_lowers(c)[k, a] + 0.5 * domain.cell_sizes(c)[k, a]
        coord = np.stack(pts)[a, inner]
        first = fc[c]
        # The child on the upper side of the split holds the larger coordinate.
        upper_child = np.where(crd[first, a] > crd[first + 1, a], first, first + 1)
        lower_child = np.where(crd[first, a] > crd[first + 1, a], first + 1, first)
        cell[inner] = np.where(coord >= mid, upper_child, lower_child)


def evaluate(state, points, leaves=None):
    """u_h at arbitrary points of the state's domain (tuple of D arrays of one
    shape); returns an array of that shape. `leaves` (same shape) names the
    leaf whose polynomial to evaluate, e.g. one side of a face; by default the
    leaf containing each point."""
    plan = state.plan
    space = plan.space
    dom = space.domain
    D = dom.dim
    shape = np.shape(points[0])
    pts = [np.asarray(p, dtype=np.float64).reshape(-1) for p in points]
    if leaves is None:
        leaf = locate(dom, pts, space.leaves)
    else:
        leaf = np.asarray(leaves, dtype=np.int64).reshape(-1)
    out = np.zeros(pts[0].size)
    row_of = np.full(dom.grid.cell_count(), -1, dtype=np.int64)
    cells = [cb for cb in getattr(plan, "cloud_buckets", []) if not getattr(cb, "surface", False)]
    for b in list(plan.buckets) + cells:
        box = np.zeros((b.cell_count, b.n_modes))
        np.add.at(box, (b.term_cell, b.term_box), b.term_coef * state.coefficients[b.term_dof])
        row_of[:] = -1
        row_of[b.cells] = np.arange(b.cell_count)
        rows = row_of[leaf]
        sel = np.nonzero(rows >= 0)[0]
        if sel.size == 0:
            continue
        r = rows[sel]
        centers, halves = dom.cell_centers_and_half_sizes(b.cells[r])
        # Box layout: axis 1 fastest, so C order is (axis D, ..., axis 1).
        acc = box[r].reshape((sel.size,) + tuple(reversed(b.modes)))
        for a in range(D):
            xi = (pts[a][sel] - centers[:, a]) / halves[:, a]
            psi = integrated_legendre(xi, b.degrees[a])
            # contract the last (axis a) index
            acc = np.einsum("n...i,ni->n...", acc, psi)
        out[sel] = acc
    return out.reshape(shape)


class _MassOperator:
    def cell_accumulate(self, q, trial, test_component):
        return TestChannels(trial.value, tuple(0.0 * g for g in trial.gradient))


class _TransferLoad:
    def __init__(self, state):
        self.state = state

    def cell_rhs_accumulate(self, q, test_component):
        return evaluate(self.state, q.point)


def projection_problem(transition, state):
    """Target L2 projection problem (src/transfer.jl:27-47): mass operator on the
    target field and the old field as load, no constraints."""
    tgt = transition.target_space
    src = transition.source_space
    if getattr(state.plan, "space", None) is not src:
        raise ValueError("state does not live on the transition's source space")
    extra = 1
    merged = []
    for l in tgt.leaves:
        pt = np.asarray(tgt.degrees[tgt.leaf_index[int(l)]])
        covered = transition.source_leaves(l)
        if len(covered) > 1 or covered[0] != int(l) and _is_descendant(tgt.domain, covered[0], int(l)):
            merged.append((int(l), covered))
            continue
        drop = np.asarray(src.degrees[src.leaf_index[covered[0]]]) - pt
        extra = max(extra, 1 + int(drop.max()))
    options = SpaceOptions(basis=tgt.options.basis, degree=tgt.options.degree,
                           quadrature=DegreePlusQuadrature(extra), continuity=tgt.options.continuity)
    space = HpSpace(tgt.domain, options, leaves=tgt.leaves, degrees=tgt.degrees)
    u = ScalarField(space, "u")
    problem = AffineProblem(u, SPD())
    problem.add_cell_accumulator(u, u, _MassOperator())
    problem.add_cell_accumulator(u, _TransferLoad(state))
    for leaf, covered in merged:
        problem.add_cell_quadrature(leaf, composite_rule(tgt, src, leaf, covered))
    return problem


def _is_descendant(domain, cell, ancestor):
    parent = domain.grid.parent()
    while cell >= 0:
        if cell == ancestor:
            return True
        cell = int(parent[cell])
    return False


def composite_rule(target_space, source_space, leaf, covered):
    """Quadrature on the reference cell of a merged target leaf made of one
    Gauss rule per covered source leaf, max(p_source, p_target) + 1 points per
    axis mapped into that leaf's sub-box: exact for the target mass matrix and
    for the load of the piecewise polynomial old field (src/transfer.jl:22-25)."""
    dom = target_space.domain
    D = dom.dim
    pt = target_space.degrees[target_space.leaf_index[leaf]]
    lo_t = dom.cell_lowers([leaf])[0]
    size_t = dom.cell_sizes([leaf])[0]
    points, weights = [], []
    for s in covered:
        ps = source_space.degrees[source_space.leaf_index[s]]
        lo = (dom.cell_lowers([s])[0] - lo_t) / size_t * 2.0 - 1.0
        hi = lo + dom.cell_sizes([s])[0] / size_t * 2.0
        axes = []
        for a in range(D):
            x, w = gauss_legendre_rule(max(int(ps[a]), int(pt[a])) + 1)
            axes.append((lo[a] + 0.5 * (x + 1.0) * (hi[a] - lo[a]), 0.5 * w * (hi[a] - lo[a])))
        # axis 1 fastest
        grids = np.meshgrid(*[ax[0] for ax in axes][::-1], indexing="ij")
        wgrid = np.meshgrid(*[ax[1] for ax in axes][::-1], indexing="ij")
        points.append(np.stack([g.reshape(-1) for g in grids[::-1]], axis=1))
        weights.append(np.prod(np.stack([w.reshape(-1) for w in wgrid]), axis=0))
    return PointQuadrature(np.concatenate(points), np.concatenate(weights))


def _regularized_cholesky_solve(M, r):
    """Batched dense solves (items x n x n) with the reference's relative
    diagonal shift 1000 eps max|M_ii| (src/transfer.jl:497-525); items whose
    matrix is zero get zero."""
    scale = np.abs(np.diagonal(M, axis1=1, axis2=2)).max(axis=1)
    out = np.zeros_like(r)
    live = scale > 0.0
    if not np.any(live):
        return out
    n = M.shape[1]
    A = M[live] + (1000.0 * np.finfo(np.float64).eps * scale[live])[:, None, None] * np.eye(n)[None]
    L = np.linalg.cholesky(A)
    y = np.linalg.solve(L, r[live][..., None])
    out[live] = np.linalg.solve(np.transpose(L, (0, 2, 1)), y)[..., 0]
    return out


def _cloud_basis(bucket):
    """psi values of the bucket's full mode box at every point (points x modes,
    axis-1-fastest boxes)."""
    D = bucket.dim
    phi = np.ones((bucket.n_points_total, 1))
    for a in range(D - 1, -1, -1):
        va = integrated_legendre(bucket.ref[:, a], bucket.degrees[a])
        phi = (phi[:, :, None] * va[:, None, :]).reshape(bucket.n_points_total, -1)
    return phi


def _active_modes(space, bucket):
    """cells x modes mask of the leaves' basis modes."""
    lo = space.mode_offsets[bucket.leaf_idx]
    counts = space.mode_offsets[bucket.leaf_idx + 1] - lo
    from .plan import _expand_ranges
    modes = _expand_ranges(lo, counts)
    mask = np.zeros((bucket.cell_count, bucket.n_modes), dtype=bool)
    mask[np.repeat(np.arange(bucket.cell_count), counts), space.mode_box[modes].astype(np.int64)] = True
    return mask


def local_projections(plan, state):
    """Per batch of the projection plan, the local L2 projection of the old
    field on each target cell (cells x modes; src/transfer.jl:778-798): the
    cell's mass matrix and the old field's load on the cell's own rule (the
    composite per-source-leaf rule on merged cells), solved per cell."""
    out = []
    for bk in list(plan.buckets) + list(plan.cloud_buckets):
        if getattr(bk, "cloud", False):
            phi = _cloud_basis(bk)
            v = evaluate(state, tuple(bk.phys_points))
            wphi = bk.weights[:, None] * phi
            M = np.zeros((bk.cell_count, bk.n_modes, bk.n_modes))
            np.add.at(M, bk.item_of_point, wphi[:, :, None] * phi[:, None, :])
            rhs = np.zeros((bk.cell_count, bk.n_modes))
            np.add.at(rhs, bk.item_of_point, wphi * v[:, None])
        else:
            phi = np.ones((1, 1))
            for a in range(bk.dim - 1, -1, -1):
                phi = np.kron(phi, bk.B[a])
            v = evaluate(state, tuple(p.reshape(-1) for p in bk.phys_points)).reshape(bk.weights.shape)
            M = np.einsum("iq,cq,jq->cij", phi, bk.weights, phi)
            rhs = (bk.weights * v) @ phi.T
        active = _active_modes(plan.space, bk)
        if not np.all(active):  # modes outside the basis: identity rows, zero load
            idx = np.arange(bk.n_modes)
            M = np.where(active[:, :, None] & active[:, None, :], M, 0.0)
            M[:, idx, idx] = np.where(active, M[:, idx, idx], 1.0)
            rhs = np.where(active, rhs, 0.0)
        out.append((bk, _regularized_cholesky_solve(M, rhs)))
    return out


def reconcile(plan, parts):
    """Target coefficients from the local projections (src/transfer.jl:637-776):
    the normal equations of the local relations sum_t c_t u_(g_t) = s_(cell,
    mode) over every (cell, mode) and its terms; dofs coupled by multi-term
    relations form connected components solved densely (regularized Cholesky),
    the others are rhs / diagonal (zero where the diagonal is below 1000 eps)."""
    import scipy.sparse as sps
    from scipy.sparse.csgraph import connected_components
    n = plan.full_dof_count
    diag = np.zeros(n)
    rhs = np.zeros(n)
    rows, cols, vals = [], [], []
    for bk, s in parts:
        diag_add = bk.term_coef * bk.term_coef
        np.add.at(diag, bk.term_dof, diag_add)
        np.add.at(rhs, bk.term_dof, bk.term_coef * s[bk.term_cell, bk.term_box])
        key = bk.term_cell * bk.n_modes + bk.term_box
        order = np.argsort(key, kind="stable")
        k_sorted = key[order]
        starts = np.nonzero(np.r_[True, k_sorted[1:] != k_sorted[:-1]])[0]
        counts = np.diff(np.r_[starts, len(k_sorted)])
        for st, ct in zip(starts[counts > 1].tolist(), counts[counts > 1].tolist()):
            terms = order[st:st + ct]
            for i in range(ct):
                for j in range(i + 1, ct):
                    a, b = int(bk.term_dof[terms[i]]), int(bk.term_dof[terms[j]])
                    if a == b:
                        continue
                    v = float(bk.term_coef[terms[i]] * bk.term_coef[terms[j]])
                    rows += [a, b]
                    cols += [b, a]
                    vals += [v, v]
    u = np.zeros(n)
    coupled = np.zeros(n, dtype=bool)
    if rows:
        C = sps.csr_matrix((vals, (rows, cols)), shape=(n, n))
        C.sum_duplicates()
        coupled[np.unique(rows)] = True
    tol = 1000.0 * np.finfo(np.float64).eps
    lone = ~coupled & (diag > tol)
    u[lone] = rhs[lone] / diag[lone]
    if rows:
        _, labels = connected_components(C, directed=False)
        for comp in np.unique(labels[coupled]).tolist():
            dofs = np.nonzero(coupled & (labels == comp))[0]
            A = C[dofs][:, dofs].toarray()
            A[np.arange(len(dofs)), np.arange(len(dofs))] = diag[dofs]
            u[dofs] = _regularized_cholesky_solve(A[None], rhs[dofs][None])[0]
    return u


def transfer_state(transition, state, linear_solve=None, relative_tolerance=1e-13, maxiter=None):
    """Old field on the transition's target space (src/transfer.jl:892-1000).

    Default (src/transfer.jl:431-437, 778-804): a local L2 projection on every
    target cell, reconciled on shared dofs through the normal equations of the
    local relations (for DG targets this is the cellwise projection). With a
    linear_solve policy (e.g. CGSolver(preconditioner=JacobiPreconditioner())),
    the global L2 projection (the reference's variational path) is solved on
    the device with it. Returns a State of the projection plan (full layout on
    the target space)."""
    from .plan import compile as compile_plan
    from .solve import State, solve

    plan = compile_plan(projection_problem(transition, state))
    if linear_solve is None:
        return State(plan, reconcile(plan, local_projections(plan, state)))
    return solve(plan, solver=linear_solve, relative_tolerance=relative_tolerance, maxiter=maxiter)


def interface_jump_indicators(state):
    """Per-leaf, per-axis interface jump indicators (src/indicators.jl:260-324):
    for every interface patch normal to axis a, J = int_F |u- - u+|^2 dS is
    added to both leaves with their own scaling (p_a + 1)^2 / h_a; the result
    is sqrt of the sums, shape (leaf count, D) in the space's leaf order. The
    patch is the overlap of the two faces (the smaller face on hanging
    interfaces), integrated by max(p-, p+) + 1 Gauss points per tangential
    axis, which is exact for the squared jump."""
    space = state.plan.space
    dom = space.domain
    D = dom.dim
    out = np.zeros((len(space.leaves), D))
    # the space's interface lists hold positions in space.leaves
    im, ip, axes = space.interface_minus, space.interface_plus, space.interface_axis
    if len(im) == 0:
        return out
    minus, plus = space.leaves[im], space.leaves[ip]
    lo_m, lo_p = dom.cell_lowers(minus), dom.cell_lowers(plus)
    hi_m, hi_p = lo_m + dom.cell_sizes(minus), lo_p + dom.cell_sizes(plus)
    for k in range(len(minus)):
        a = int(axes[k])
        n = int(max(space.degrees[im[k]].max(), space.degrees[ip[k]].max())) + 1
        x, w = gauss_legendre_rule(n)
        lo = np.maximum(lo_m[k], lo_p[k])
        hi = np.minimum(hi_m[k], hi_p[k])
        tang = [b for b in range(D) if b != a]
        grids = np.meshgrid(*[lo[b] + 0.5 * (x + 1.0) * (hi[b] - lo[b]) for b in tang], indexing="ij")
        wts = np.ones(1)
        for b in tang:
            wts = np.multiply.outer(wts, 0.5 * w * (hi[b] - lo[b]))
        wts = wts.reshape(-1)
        m = wts.size
        pts = [None] * D
        for j, b in enumerate(tang):
            pts[b] = grids[j].reshape(-1)
        pts[a] = np.full(m, hi_m[k][a])
        jump = evaluate(state, tuple(pts), np.full(m, minus[k])) - evaluate(state, tuple(pts), np.full(m, plus[k]))
        energy = float(np.sum(wts * jump * jump))
        if not np.isfinite(energy):
            raise ValueError("interface jump indicators must be finite")
        for idx, leaf in ((im[k], minus[k]), (ip[k], plus[k])):
            h = dom.cell_sizes([leaf])[0][a]
            out[idx, a] += (space.degrees[idx][a] + 1) ** 2 / h * energy
    return np.sqrt(out)


def coefficient_coarsening_indicators(state):
    """Per-leaf, per-axis normalized top-layer modal energy sqrt(E_top / E_total)
    (src/indicators.jl:326-342, layer energies 61-108). E_total sums the squared
    local mode amplitudes of the leaf; E_top sums those whose index along the
    axis equals the degree. At degree 1 each (0, 1) endpoint pair along the
    axis is read as constant + linear content and E_top sums the squared linear
    parts (upper - lower) / 2; at degree 0 it is zero. Shape (leaf count, D),
    space leaf order."""
    plan = state.plan
    space = plan.space
    D = space.dim
    out = np.zeros((len(space.leaves), D))
    for b in plan.buckets:
        amp = np.zeros((b.cell_count, b.n_modes))
        np.add.at(amp, (b.term_cell, b.term_box), b.term_coef * state.coefficients[b.term_dof])
        active = np.zeros((b.cell_count, b.n_modes), dtype=bool)
        for r, li in enumerate(b.leaf_idx):
            active[r, space.mode_box[space.mode_offsets[li]:space.mode_offsets[li + 1]]] = True
        amp = np.where(active, amp, 0.0)
        total = np.sum(amp * amp, axis=1)
        shape = tuple(reversed(b.modes))  # C order: axis D .. axis 1
        A = amp.reshape((b.cell_count,) + shape)
        M = active.reshape((b.cell_count,) + shape)
        for a in range(D):
            p = b.degrees[a]
            ax = D - a  # array axis of mesh axis a
            if p == 0:
                continue
            if p == 1:
                lower = np.take(A, 0, axis=ax)
                upper = np.take(A, 1, axis=ax)
                both = np.take(M, 0, axis=ax) & np.take(M, 1, axis=ax)
                lin = np.where(both, 0.5 * (upper - lower), 0.0)
                top = np.sum((lin * lin).reshape(b.cell_count, -1), axis=1)
            else:
                t = np.take(A, p, axis=ax)
                top = np.sum((t * t).reshape(b.cell_count, -1), axis=1)
            with np.errstate(invalid="ignore", divide="ignore"):
                out[b.leaf_idx, a] = np.where(total > 0, np.sqrt(top / np.where(total > 0, total, 1.0)), 0.0)
    if not np.all(np.isfinite(out)):
        raise ValueError("modal indicator energies must be finite")
    return out
