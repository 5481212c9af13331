"""ORACLE (test infrastructure only): Python driver of oracle/oracle.c.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module. It restates the reference's per-cell pipeline: gather ->
interpolate (value + gradients) -> user callback on the actual trial channels
at every point -> weighted projection -> scatter skipping fixed rows ->
constraint rows (src/assembly.jl:540-564, 1160-1174;
src/accumulators.jl:1955-2013). It reads the plan's host tables (term maps,
points, weights) but none of the device lowering.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

P_f64 = ctypes.POINTER(ctypes.c_double)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_u8 = ctypes.POINTER(ctypes.c_uint8)


class OracleBucket(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("N", ctypes.c_int * 3), ("Q", ctypes.c_int * 3), ("cells", ctypes.c_int64),
        ("values", P_f64 * 3), ("derivatives", P_f64 * 3), ("inverse_jacobian", P_f64), ("weights", P_f64),
        ("term_offsets", P_i64), ("term_indices", P_i64), ("term_coefficients", P_f64),
    ]


OPERATOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p, P_f64, P_f64)

_lib = None


def build():
    """Compile liboracle.so from oracle.c (gcc, OpenMP)."""
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall",
                    "-o", LIB, os.path.join(HERE, "oracle.c"), "-lm"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        B = ctypes.POINTER(OracleBucket)
        L.oracle_interpolate.argtypes = [B, P_f64, P_f64, P_f64]
        L.oracle_project.argtypes = [B, P_f64, P_f64, P_u8, P_f64]
        L.oracle_fixed_rows.argtypes = [P_i64, ctypes.c_int64, P_f64, P_f64]
        L.oracle_apply_lowered.argtypes = [B, P_f64, P_f64, P_u8, P_f64, P_i64, P_i64, ctypes.c_int, ctypes.c_int]
        L.oracle_apply_basis.argtypes = [B, P_f64, P_f64, P_u8, P_f64]
        L.oracle_diagonal.argtypes = [B, P_f64, P_f64]
        L.oracle_color_cells.argtypes = [B, ctypes.c_int64, P_i32]
        L.oracle_color_cells.restype = ctypes.c_int
        L.oracle_cg.argtypes = [OPERATOR, ctypes.c_void_p, ctypes.c_int64, P_f64, P_f64, P_f64,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int64]
        L.oracle_cg.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(t))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(t))


class _Bucket:
    """ctypes view of one plan bucket with its arrays kept alive."""

    def __init__(self, bk):
        self.bk = bk
        nc, nb = bk.cell_count, bk.n_modes
        order = np.lexsort((bk.term_box, bk.term_cell))
        entry = bk.term_cell[order] * nb + bk.term_box[order]
        counts = np.bincount(entry, minlength=nc * nb)
        self.offsets = np.zeros(nc * nb + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum(counts)
        self.indices = np.ascontiguousarray(bk.term_dof[order], dtype=np.int64)
        self.coefs = np.ascontiguousarray(bk.term_coef[order], dtype=np.float64)
        self.values = [np.ascontiguousarray(t) for t in bk.B]
        self.derivs = [np.ascontiguousarray(t) for t in bk.G]
        self.jinv = np.ascontiguousarray(bk.inverse_jacobian)
        self.weights = np.ascontiguousarray(bk.weights)
        s = OracleBucket()
        s.dim = bk.dim
        for a in range(3):
            s.N[a] = bk.modes[a] if a < bk.dim else 1
            s.Q[a] = bk.points[a] if a < bk.dim else 1
            if a < bk.dim:
                s.values[a] = _p(self.values[a], ctypes.c_double)
                s.derivatives[a] = _p(self.derivs[a], ctypes.c_double)
        s.cells = nc
        s.inverse_jacobian = _p(self.jinv, ctypes.c_double)
        s.weights = _p(self.weights, ctypes.c_double)
        s.term_offsets = _p(self.offsets, ctypes.c_int64)
        s.term_indices = _p(self.indices, ctypes.c_int64)
        s.term_coefficients = _p(self.coefs, ctypes.c_double)
        self.struct = s

    def ref(self):
        return ctypes.byref(self.struct)


def _fixed_mask(plan):
    """Rows the full apply writes from the constraints instead of the
    operator: fixed dofs, affine-relation pivots and mean-row pivots
    (src/plans.jl:457-480 masks, src/assembly.jl:1788-1803 scatter skip)."""
    m = np.zeros(plan.full_dof_count, dtype=np.uint8)
    m[plan.fixed_dofs] = 1
    for row in getattr(plan, "affine_rows", []) + getattr(plan, "mean_rows", []):
        m[row[0]] = 1
    return m


def _constraint_rows(plan, x, y):
    """src/assembly.jl:1893-1920: fixed rows copy x; relation rows
    y_p = x_p - sum c_j x_j; mean rows y_p = sum c_j x_j."""
    for (p, idx, cf, _) in getattr(plan, "affine_rows", []):
        v = x[p]
        for j, c in zip(idx.tolist(), cf.tolist()):
            v -= c * x[j]
        y[p] = v
    for (p, idx, cf, _) in getattr(plan, "mean_rows", []):
        v = 0.0
        for j, c in zip(idx.tolist(), cf.tolist()):
            v += c * x[j]
        y[p] = v


def expand(plan, x_red, shift=True):
    """shift + P x_red from the plan's CSR map (src/assembly.jl:851-874)."""
    out = np.array(plan.shift, dtype=np.float64) if shift else np.zeros(plan.full_dof_count)
    off, idx, cf = plan.map_offsets, plan.map_indices, plan.map_coefficients
    for f in np.nonzero(np.diff(off))[0].tolist():
        for t in range(off[f], off[f + 1]):
            out[f] += cf[t] * x_red[idx[t]]
    return out


def project(plan, y):
    """P^T y from the plan's CSR map (src/assembly.jl:876-891)."""
    out = np.zeros(plan.reduced_dof_count)
    off, idx, cf = plan.map_offsets, plan.map_indices, plan.map_coefficients
    for f in np.nonzero(np.diff(off))[0].tolist():
        for t in range(off[f], off[f + 1]):
            out[idx[t]] += cf[t] * y[f]
    return out


def _trial(value, gradient):
    from gricoamd.problem import TrialChannels
    return TrialChannels(0, value, tuple(gradient))


def _test_channels(result, shape, D):
    from gricoamd.problem import TestChannels
    if isinstance(result, TestChannels):
        v = np.broadcast_to(np.asarray(result.value, dtype=np.float64), shape)
        g = [np.broadcast_to(np.asarray(c, dtype=np.float64), shape) for c in result.gradient]
        return v, g
    return np.broadcast_to(np.asarray(result, dtype=np.float64), shape), [np.zeros(shape)] * D


_CALLBACK = {False: "cell_accumulate", True: "surface_accumulate", "boundary": "boundary_accumulate"}


def _cloud_groups(plan):
    """(bucket, operators, kind) for every point-cloud batch of the plan; kind
    False = cells, True = embedded surface, "boundary" = boundary faces,
    "interface" = interfaces."""
    cells = [op for (_, _, op) in plan.problem.cell_operators]
    surf = [op for (_, _, op) in getattr(plan.problem, "surface_operators", [])]
    iops = [op for (_, _, op) in getattr(plan.problem, "interface_operators", [])]
    return ([(b, cells, False) for b in getattr(plan, "cloud_buckets", [])] +
            [(b, surf, True) for b in getattr(plan, "surface_buckets", [])] +
            [(b, ops, "boundary") for (b, ops, _) in getattr(plan, "boundary_buckets", []) if ops] +
            ([(b, iops, "interface") for b in getattr(plan, "interface_buckets", [])] if iops else []))


def _gather_local(bk, x):
    loc = np.zeros((bk.cell_count, bk.n_modes))
    np.add.at(loc, (bk.term_cell, bk.term_box), bk.term_coef * x[bk.term_dof])
    return loc


def _scatter_local(bk, rloc, y, fixed, squared=False):
    c = bk.term_coef ** 2 if squared else bk.term_coef
    contrib = c * rloc[bk.term_cell, bk.term_box]
    keep = np.ones(len(contrib), dtype=bool) if fixed is None else fixed[bk.term_dof] == 0
    np.add.at(y, bk.term_dof[keep], contrib[keep])


def _interface_apply(bk, ops, x, y, fixed):
    """Interface items (src/accumulators.jl:1576-1616 interface_apply!): the
    two-sided trial trace from the local coefficients (minus box, then plus
    box), the callback on the actual trace channels, each returned side
    projected onto that side's modes, scatter skipping fixed rows. Host numpy."""
    from gricoamd.problem import TraceTrialChannels
    D = bk.dim
    loc = _gather_local(bk, x)
    up = loc[bk.item_of_point]
    trial = []
    for s, (phi, grads) in enumerate(bk.tables()):
        lo = 0 if s == 0 else bk.n_minus
        u = up[:, lo:lo + phi.shape[1]]
        trial.append(_trial(np.einsum("pn,pn->p", phi, u), [np.einsum("pn,pn->p", grads[a], u) for a in range(D)]))
    q = bk.quadrature()
    acc = [(np.zeros(bk.n_points_total), [np.zeros(bk.n_points_total) for _ in range(D)]) for _ in range(2)]
    for op in ops:
        res = op.interface_accumulate(q, TraceTrialChannels(0, trial[0], trial[1]), 0)
        for s, (v, g) in enumerate(bk.trace_channels(res)):
            acc[s][0][:] += bk.weights * v
            for a in range(D):
                acc[s][1][a][:] += bk.weights * g[a]
    _scatter_local(bk, bk.project_sides(acc), y, fixed)


def _interface_diagonal(bk, ops, d):
    """src/accumulators.jl:1618-1664 interface_diagonal!: each trace basis
    function as the trial (the other side zero, 723-736), its own side's test
    contribution, scattered with squared coefficients."""
    from gricoamd.problem import TraceTrialChannels
    D = bk.dim
    P = bk.n_points_total
    q = bk.quadrature()
    zeros = np.zeros(P)
    rloc = np.zeros((bk.cell_count, bk.n_modes))
    for s, (phi, grads) in enumerate(bk.tables()):
        lo = 0 if s == 0 else bk.n_minus
        for i in range(phi.shape[1]):
            active = _trial(phi[:, i], [grads[a][:, i] for a in range(D)])
            inactive = _trial(zeros, [zeros] * D)
            tr = TraceTrialChannels(0, active, inactive) if s == 0 else TraceTrialChannels(0, inactive, active)
            contrib = np.zeros(P)
            for op in ops:
                v, g = bk.trace_channels(op.interface_accumulate(q, tr, 0))[s]
                c = v * phi[:, i]
                for a in range(D):
                    c = c + g[a] * grads[a][:, i]
                contrib += bk.weights * c
            np.add.at(rloc[:, lo + i], bk.item_of_point, contrib)
    _scatter_local(bk, rloc, d, None, squared=True)


def _cloud_apply(bk, ops, surface, x, y, fixed):
    """Basis path of the reference (src/accumulators.jl:1022-1054, one-sided
    surface and boundary-face items 1133-1161, 1202-1214): gather, full tables
    at the points, the callback on the actual trial channels, weighted
    projection, scatter skipping fixed rows. Host numpy."""
    if surface == "interface":
        return _interface_apply(bk, ops, x, y, fixed)
    D = bk.dim
    phi, grads = bk.tables()
    loc = np.zeros((bk.cell_count, bk.n_modes))
    np.add.at(loc, (bk.term_cell, bk.term_box), bk.term_coef * x[bk.term_dof])
    up = loc[bk.item_of_point]
    val = np.einsum("pn,pn->p", phi, up)
    grad = [np.einsum("pn,pn->p", grads[a], up) for a in range(D)]
    q = bk.quadrature()
    P = bk.n_points_total
    a0 = np.zeros(P)
    ag = np.zeros((D, P))
    for op in ops:
        call = getattr(op, _CALLBACK[surface])
        v, g = _test_channels(call(q, _trial(val, grad), 0), (P,), D)
        a0 += bk.weights * v
        for a in range(D):
            ag[a] += bk.weights * g[a]
    r = phi * a0[:, None]
    for a in range(D):
        r = r + grads[a] * ag[a][:, None]
    rloc = np.zeros((bk.cell_count, bk.n_modes))
    np.add.at(rloc, bk.item_of_point, r)
    contrib = bk.term_coef * rloc[bk.term_cell, bk.term_box]
    keep = np.ones(len(contrib), dtype=bool) if fixed is None else fixed[bk.term_dof] == 0
    np.add.at(y, bk.term_dof[keep], contrib[keep])


def _cloud_diagonal(bk, ops, surface, d):
    """Local diagonal sum_q sum_{c,k} W M_q[c][k] chi_c chi_k scattered with c^2."""
    if surface == "interface":
        return _interface_diagonal(bk, ops, d)
    D = bk.dim
    phi, grads = bk.tables()
    q = bk.quadrature()
    P = bk.n_points_total
    ones, zeros = np.ones(P), np.zeros(P)
    M = np.zeros((P, D + 1, D + 1))
    for op in ops:
        call = getattr(op, _CALLBACK[surface])
        for j in range(D + 1):
            v, g = _test_channels(call(q, _trial(ones if j == 0 else zeros,
                                                 [ones if j == a + 1 else zeros for a in range(D)]), 0), (P,), D)
            M[:, 0, j] += v
            for a in range(D):
                M[:, a + 1, j] += g[a]
    M *= bk.weights[:, None, None]
    chi = np.concatenate([phi[None], grads], axis=0)  # (D+1, P, N)
    dl = np.einsum("pck,cpn,kpn->pn", M, chi, chi)
    rloc = np.zeros((bk.cell_count, bk.n_modes))
    np.add.at(rloc, bk.item_of_point, dl)
    contrib = bk.term_coef ** 2 * rloc[bk.term_cell, bk.term_box]
    np.add.at(d, bk.term_dof, contrib)


def apply(plan, x, physical=False):
    """Reference-algorithm full-layout apply (callback evaluated on the actual
    trial channels). physical=True skips the constraint handling."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(plan.full_dof_count)
    fixed = None if physical else _fixed_mask(plan)
    for bk in plan.buckets:
        ob = _Bucket(bk)
        nc, nq, D = bk.cell_count, bk.n_points, bk.dim
        vals = np.zeros((nc, nq))
        grads = np.zeros((nc, nq, D))
        L.oracle_interpolate(ob.ref(), _p(x, ctypes.c_double), _p(vals, ctypes.c_double), _p(grads, ctypes.c_double))
        q = bk.quadrature()
        a0 = np.zeros((nc, nq))
        ag = np.zeros((nc, nq, D))
        for (_, _, op) in plan.problem.cell_operators:
            v, g = _test_channels(op.cell_accumulate(q, _trial(vals, [grads[..., a] for a in range(D)]), 0),
                                  (nc, nq), D)
            a0 += bk.weights * v
            for a in range(D):
                ag[..., a] += bk.weights * g[a]
        a0 = np.ascontiguousarray(a0)
        ag = np.ascontiguousarray(ag)
        L.oracle_project(ob.ref(), _p(a0, ctypes.c_double), _p(ag, ctypes.c_double),
                         _p(fixed, ctypes.c_uint8), _p(y, ctypes.c_double))
    for bk, ops, surface in _cloud_groups(plan):
        _cloud_apply(bk, ops, surface, x, y, fixed)
    if not physical:
        fd = np.ascontiguousarray(plan.fixed_dofs, dtype=np.int64)
        L.oracle_fixed_rows(_p(fd, ctypes.c_int64), len(fd), _p(x, ctypes.c_double), _p(y, ctypes.c_double))
        _constraint_rows(plan, x, y)
    return y


def action_matrices(plan, bk):
    """Per-point (D+1)x(D+1) action of the plan's accumulators, probed with
    unit trial channels (the accumulator contract is linear in the trial)."""
    nc, nq, D = bk.cell_count, bk.n_points, bk.dim
    q = bk.quadrature()
    M = np.zeros((nc, nq, D + 1, D + 1))
    ones, zeros = np.ones((nc, nq)), np.zeros((nc, nq))
    for (_, _, op) in plan.problem.cell_operators:
        for j in range(D + 1):
            val = ones if j == 0 else zeros
            grad = [ones if j == a + 1 else zeros for a in range(D)]
            v, g = _test_channels(op.cell_accumulate(q, _trial(val, grad), 0), (nc, nq), D)
            M[..., 0, j] += v
            for a in range(D):
                M[..., a + 1, j] += g[a]
    return np.ascontiguousarray(M.reshape(nc * nq, (D + 1) ** 2))


def apply_lowered(plan, x, threads=0, matrices=None, colors=None, physical=False):
    """Whole-pipeline C apply with probed action matrices; threaded over
    DOF-disjoint colors. Returns y."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(plan.full_dof_count)
    fixed = None if physical else _fixed_mask(plan)
    for i, bk in enumerate(plan.buckets):
        ob = _Bucket(bk)
        M = matrices[i] if matrices is not None else action_matrices(plan, bk)
        co, cc = colors[i] if colors is not None else color_batches(plan, ob)
        L.oracle_apply_lowered(ob.ref(), _p(M, ctypes.c_double), _p(x, ctypes.c_double), _p(fixed, ctypes.c_uint8),
                               _p(y, ctypes.c_double), _p(co, ctypes.c_int64), _p(cc, ctypes.c_int64),
                               len(co) - 1, int(threads))
    for bk, ops, surface in _cloud_groups(plan):
        _cloud_apply(bk, ops, surface, x, y, fixed)
    if not physical:
        fd = np.ascontiguousarray(plan.fixed_dofs, dtype=np.int64)
        L.oracle_fixed_rows(_p(fd, ctypes.c_int64), len(fd), _p(x, ctypes.c_double), _p(y, ctypes.c_double))
    return y


def color_batches(plan, ob):
    L = lib()
    col = np.zeros(ob.bk.cell_count, dtype=np.int32)
    k = L.oracle_color_cells(ob.ref(), plan.full_dof_count, _p(col, ctypes.c_int32))
    if k < 0:
        raise RuntimeError("more than 64 colors")
    order = np.argsort(col, kind="stable").astype(np.int64)
    offsets = np.zeros(k + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(np.bincount(col, minlength=k))
    return offsets, np.ascontiguousarray(order)


def apply_basis(plan, x, physical=False):
    """Brute-force full-table path (src/accumulators.jl:1022-1054)."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(plan.full_dof_count)
    fixed = None if physical else _fixed_mask(plan)
    for bk in plan.buckets:
        ob = _Bucket(bk)
        M = action_matrices(plan, bk)
        L.oracle_apply_basis(ob.ref(), _p(M, ctypes.c_double), _p(x, ctypes.c_double), _p(fixed, ctypes.c_uint8),
                             _p(y, ctypes.c_double))
    for bk, ops, surface in _cloud_groups(plan):
        _cloud_apply(bk, ops, surface, x, y, fixed)
    if not physical:
        fd = np.ascontiguousarray(plan.fixed_dofs, dtype=np.int64)
        L.oracle_fixed_rows(_p(fd, ctypes.c_int64), len(fd), _p(x, ctypes.c_double), _p(y, ctypes.c_double))
    return y


def diagonal_physical(plan):
    L = lib()
    d = np.zeros(plan.full_dof_count)
    for bk in plan.buckets:
        ob = _Bucket(bk)
        M = action_matrices(plan, bk)
        L.oracle_diagonal(ob.ref(), _p(M, ctypes.c_double), _p(d, ctypes.c_double))
    for bk, ops, surface in _cloud_groups(plan):
        _cloud_diagonal(bk, ops, surface, d)
    return d


def _state_channel_arrays(L, bk, u):
    ob = _Bucket(bk)
    nc, nq, D = bk.cell_count, bk.n_points, bk.dim
    vals = np.zeros((nc, nq))
    grads = np.zeros((nc, nq, D))
    L.oracle_interpolate(ob.ref(), _p(np.ascontiguousarray(u, dtype=np.float64), ctypes.c_double),
                         _p(vals, ctypes.c_double), _p(grads, ctypes.c_double))
    return ob, vals, [grads[..., a] for a in range(D)]


def _project_channels(L, ob, bk, results, fixed, y):
    D = bk.dim
    shape = (bk.cell_count, bk.n_points)
    a0 = np.zeros(shape)
    ag = np.zeros(shape + (D,))
    for res in results:
        v, g = _test_channels(res, shape, D)
        a0 += bk.weights * v
        for a in range(D):
            ag[..., a] += bk.weights * g[a]
    L.oracle_project(ob.ref(), _p(np.ascontiguousarray(a0), ctypes.c_double),
                     _p(np.ascontiguousarray(ag), ctypes.c_double), _p(fixed, ctypes.c_uint8),
                     _p(y, ctypes.c_double))


def residual(rplan, u_full):
    """Reference residual! (src/assembly.jl:997-1026 with the tensor residual
    kernel src/accumulators.jl:2015-2060): state channels interpolated from the
    gathered coefficients, the residual callback on them, weighted projection,
    scatter skipping fixed rows, then the constraint rows u[d] - g_d
    (1922-1940). Full layout."""
    from gricoamd.nonlinear import StateChannels
    L = lib()
    plan = rplan.plan
    y = np.zeros(plan.full_dof_count)
    fixed = _fixed_mask(plan)
    for bk in plan.buckets:
        ob, v, g = _state_channel_arrays(L, bk, u_full)
        q = bk.quadrature()
        st = StateChannels(v, g)
        _project_channels(L, ob, bk, [op.cell_residual_accumulate(q, st, 0)
                                      for (_, _, op) in rplan.problem.cell_operators], fixed, y)
    y[plan.fixed_dofs] = np.asarray(u_full)[plan.fixed_dofs] - plan.fixed_values
    return y


def tangent_apply(rplan, u_full, increment):
    """Reference tangent_apply! (src/assembly.jl:1050-1080, tensor tangent
    kernel src/accumulators.jl:2062-2148): state and increment channels at the
    points, the tangent callback on the actual increment channels, weighted
    projection, scatter skipping fixed rows; fixed rows copy the increment."""
    from gricoamd.nonlinear import StateChannels
    L = lib()
    plan = rplan.plan
    y = np.zeros(plan.full_dof_count)
    fixed = _fixed_mask(plan)
    for bk in plan.buckets:
        ob, v, g = _state_channel_arrays(L, bk, u_full)
        _, dv, dg = _state_channel_arrays(L, bk, increment)
        q = bk.quadrature()
        st = StateChannels(v, g)
        _project_channels(L, ob, bk, [op.cell_tangent_accumulate(q, st, _trial(dv, dg), 0)
                                      for (_, _, op) in rplan.problem.cell_operators], fixed, y)
    y[plan.fixed_dofs] = np.asarray(increment)[plan.fixed_dofs]
    return y


def newton_dense(rplan, u0=None, rtol=1e-10, atol=1e-10, maxiter=20):
    """Newton on the free dofs with the dense tangent from oracle tangent
    applies of unit vectors (small plans): u_free -= J^-1 R(u)."""
    plan = rplan.plan
    n = plan.reduced_dof_count
    u = plan.shift.copy() if u0 is None else np.array(u0, dtype=np.float64)
    norms = []
    for _ in range(maxiter + 1):
        r = residual(rplan, u)[plan.solve_dofs]
        norms.append(float(np.linalg.norm(r)))
        if norms[-1] <= max(atol, rtol * norms[0]):
            return u, norms
        J = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(plan.full_dof_count)
            e[plan.solve_dofs[j]] = 1.0
            J[:, j] = tangent_apply(rplan, u, e)[plan.solve_dofs]
        u[plan.solve_dofs] += np.linalg.solve(J, -r)
    raise RuntimeError("oracle Newton did not converge")


def reduced_apply(plan, x_red):
    """P^T A P x (src/assembly.jl:905-915); selection maps by indexing."""
    if getattr(plan, "selection_map", True):
        full = np.zeros(plan.full_dof_count)
        full[plan.solve_dofs] = x_red
        return apply(plan, full, physical=True)[plan.solve_dofs]
    return project(plan, apply(plan, expand(plan, x_red, shift=False), physical=True))


def cg(plan, b_red, inv_diag=None, rtol=1e-10, atol=0.0, maxiter=10000, lowered=True, threads=0):
    """Reference CG on the reduced system with the oracle operator."""
    L = lib()
    n = plan.reduced_dof_count
    mats = [action_matrices(plan, bk) for bk in plan.buckets] if lowered else None
    cols = [color_batches(plan, _Bucket(bk)) for bk in plan.buckets] if lowered else None

    def op(_ctx, xp, yp):
        xr = np.ctypeslib.as_array(xp, shape=(n,))
        yr = np.ctypeslib.as_array(yp, shape=(n,))
        if not getattr(plan, "selection_map", True):
            yr[:] = reduced_apply(plan, xr)
            return
        full = np.zeros(plan.full_dof_count)
        full[plan.solve_dofs] = xr
        if lowered:
            out = apply_lowered(plan, full, threads=threads, matrices=mats, colors=cols, physical=True)
        else:
            out = apply(plan, full, physical=True)
        yr[:] = out[plan.solve_dofs]

    cb = OPERATOR(op)
    b = np.ascontiguousarray(b_red, dtype=np.float64)
    x = np.zeros(n)
    inv = None if inv_diag is None else np.ascontiguousarray(inv_diag, dtype=np.float64)
    it = L.oracle_cg(cb, None, n, _p(inv, ctypes.c_double), _p(b, ctypes.c_double), _p(x, ctypes.c_double),
                     float(rtol), float(atol), int(maxiter))
    return x, int(it)


# ---- hp geometric multigrid (src/multigrid.jl:858-1087), dense, small levels ----

def reduced_matrix(plan):
    """Dense P^T A P of a small plan from oracle applies of unit vectors."""
    n = plan.reduced_dof_count
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = reduced_apply(plan, e)
    return A


def _probe_points(ops, callback, q, P, D):
    """W-less (D+1)x(D+1) action per point of `ops` probed with unit trial
    channels through `callback` (linear in the trial channels)."""
    M = np.zeros((P, D + 1, D + 1))
    ones, zeros = np.ones(P), np.zeros(P)
    for op in ops:
        call = getattr(op, callback)
        for j in range(D + 1):
            v, g = _test_channels(call(q, _trial(ones if j == 0 else zeros,
                                                 [ones if j == a + 1 else zeros for a in range(D)]), 0), (P,), D)
            M[:, 0, j] += v
            for a in range(D):
                M[:, a + 1, j] += g[a]
    return M


def assembled_reduced_matrix(plan):
    """Sparse P^T A P from dense element matrices E_c = sum_q chi_q^T (W_q M_q)
    chi_q, chi = (value, physical gradient) of the full basis at the points and
    M_q the accumulators probed at the points (the reference's basis path,
    src/accumulators.jl:1022-1054, 1133-1161), scattered through the term maps.
    Tensor cells, point-cloud cells and embedded-surface pieces; P is the
    plan's reduced map (a selection: fixed dofs drop out with their rows and
    columns; else the coefficient CSR of the affine and mean relations)."""
    import scipy.sparse as sps
    if getattr(plan, "boundary_buckets", []) or getattr(plan, "interface_buckets", []):
        raise ValueError("assembled_reduced_matrix covers cells and surface pieces")
    n = plan.full_dof_count
    D = plan.dim
    A = sps.csr_matrix((n, n))
    cell_ops = [op for (_, _, op) in plan.problem.cell_operators]
    surf_ops = [op for (_, _, op) in getattr(plan.problem, "surface_operators", [])]
    groups = [(b, False) for b in plan.buckets] + [(b, True) for b in getattr(plan, "cloud_buckets", [])] + \
             [(b, True) for b in getattr(plan, "surface_buckets", [])]
    for bk, cloud in groups:
        if bk.cell_count == 0:
            continue
        if cloud:
            phi, grads = bk.tables()
            chi = np.concatenate([phi[None], np.asarray(grads)], axis=0)  # (D+1, P, N)
            ops, cb = (surf_ops, "surface_accumulate") if bk.surface else (cell_ops, "cell_accumulate")
            M = _probe_points(ops, cb, bk.quadrature(), bk.n_points_total, D) * bk.weights[:, None, None]
            E = np.zeros((bk.cell_count, bk.n_modes, bk.n_modes))
            for c in range(bk.cell_count):
                s0, s1 = bk.point_offsets[c], bk.point_offsets[c + 1]
                ch = chi[:, s0:s1, :]
                E[c] = np.einsum("cpn,pck,kpm->nm", ch, M[s0:s1], ch)
        else:
            Phi, dPhi = _dense_tables(bk)
            M = action_matrices(plan, bk).reshape(bk.cell_count, bk.n_points, D + 1, D + 1)
            M = M * bk.weights[:, :, None, None]
            chi = [np.broadcast_to(Phi, (bk.cell_count,) + Phi.shape)] + \
                  [dPhi[a][None] * bk.inverse_jacobian[:, a][:, None, None] for a in range(D)]
            E = np.zeros((bk.cell_count, bk.n_modes, bk.n_modes))
            for c1 in range(D + 1):
                for c2 in range(D + 1):
                    E += np.einsum("cqn,cq,cqm->cnm", chi[c1], M[:, :, c1, c2], chi[c2])
        S = sps.csr_matrix((bk.term_coef, (bk.term_cell * bk.n_modes + bk.term_box, bk.term_dof)),
                           shape=(bk.cell_count * bk.n_modes, n))
        A = A + S.T @ sps.block_diag(list(E), format="csr") @ S
    if getattr(plan, "selection_map", True):
        sd = plan.solve_dofs
        return sps.csr_matrix(A)[sd][:, sd]
    rows = np.repeat(np.arange(n), np.diff(plan.map_offsets))
    Pm = sps.csr_matrix((plan.map_coefficients, (rows, plan.map_indices)), shape=(n, plan.reduced_dof_count))
    return sps.csr_matrix(Pm.T @ A @ Pm)


def _chebyshev(A, dinv, x, b, x_zero, lam_max, degree, rng):
    """Chebyshev smoothing in S A, S = diag(dinv), or S = dinv(r) when dinv
    is callable (the block smoother)."""
    S = dinv if callable(dinv) else (lambda v: dinv * v)
    lam_min = lam_max / rng
    theta, delta = 0.5 * (lam_max + lam_min), 0.5 * (lam_max - lam_min)
    sigma = theta / delta
    r = b.copy() if x_zero else b - A @ x
    d = S(r) / theta
    x = d.copy() if x_zero else x + d
    rho = 1.0 / sigma
    for _ in range(1, degree):
        r = r - A @ d
        rho_new = 1.0 / (2.0 * sigma - rho)
        d = rho_new * rho * d + (2.0 * rho_new / delta) * S(r)
        x = x + d
        rho = rho_new
    return x


def cell_block_dofs(plan, all_cells=False, max_dofs=512):
    """Reduced dofs touched by each cell that carries its own rule (the
    plan's point-cloud cells), in cell order: the blocks of the cell-block
    smoother; all_cells: every tensor and point-cloud cell with at most
    max_dofs of them."""
    out = []
    buckets = (list(plan.buckets) if all_cells else []) + list(getattr(plan, "cloud_buckets", []))
    for cb in buckets:
        red = plan.reduced_index[cb.term_dof]
        for c in range(cb.cell_count):
            d = np.unique(red[(cb.term_cell == c) & (red >= 0)])
            if len(d) and (not all_cells or len(d) <= max_dofs):
                out.append(d)
    return out


def block_smoother(A, blocks):
    """S r = D_u^-1 r + sum_b R_b^T (R_b A R_b^T)^-1 R_b r: dense inverses of the
    level matrix restricted to each block's dofs, point Jacobi on the dofs no
    block covers."""
    d = np.diag(A)
    covered = np.zeros(len(d), dtype=bool)
    inv = []
    for dofs in blocks:
        dofs = np.asarray(dofs, dtype=np.int64)
        inv.append((dofs, np.linalg.inv(A[np.ix_(dofs, dofs)])))
        covered[dofs] = True
    dinv_u = np.where(covered, 0.0, 1.0 / d)

    def S(r):
        z = dinv_u * r
        for dofs, Bi in inv:
            z[dofs] += Bi @ r[dofs]
        return z
    return S


def gmres_smooth(A, dinv, x, b, steps, restart, omega):
    """One fixed GMRES smoothing cycle (src/multigrid.jl:963-1035): m =
    min(steps, restart, n) Arnoldi steps on A D^-1 from r = b - A x (modified
    Gram-Schmidt), then x += omega Z y with y the least-squares minimizer of
    |beta e1 - H y| (solved here with numpy's lstsq instead of the
    reference's Givens rotations)."""
    n = len(b)
    r = b - A @ x
    beta = float(np.linalg.norm(r))
    m = min(int(steps), int(restart), n)
    if beta == 0.0 or m == 0:
        return x
    V = np.zeros((n, m + 1))
    Z = np.zeros((n, m))
    H = np.zeros((m + 1, m))
    V[:, 0] = r / beta
    k = 0
    for j in range(m):
        k = j + 1
        Z[:, j] = dinv * V[:, j]
        w = A @ Z[:, j]
        for i in range(j + 1):
            H[i, j] = w @ V[:, i]
            w = w - H[i, j] * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        if H[j + 1, j] == 0.0:
            break
        V[:, j + 1] = w / H[j + 1, j]
    e = np.zeros(k + 1)
    e[0] = beta
    y = np.linalg.lstsq(H[:k + 1, :k], e, rcond=None)[0]
    return x + omega * (Z[:, :k] @ y)


def vcycle(mats, prolongations, r, smoother="jacobi", pre=2, post=2, omega=0.8, lambdas=None,
           cheb_degree=4, cheb_range=20.0, restart=8, blocks=None):
    """z = M^-1 r for the V-cycle: damped Jacobi sweeps (src/multigrid.jl:943-956)
    or Chebyshev applications, residual restriction with P^T, recursive coarse
    correction, exact dense coarse solve (891-899), prolongation (1060-1072).
    mats[l]: dense reduced operator of level l (0 = coarsest); prolongations[l]:
    dense (n_{l+1} x n_l) matrix. blocks[l] (optional): dof lists of level l's
    block smoother (block_smoother), which replaces D^-1 in the Jacobi and
    Chebyshev smoothers."""
    dinvs = [1.0 / np.diag(A) if A.size else np.zeros(0) for A in mats]
    if blocks is not None:
        for l, bl in enumerate(blocks):
            if l > 0 and bl:
                dinvs[l] = block_smoother(mats[l], bl)

    def smooth(l, x, b, steps, x_zero):
        A, dinv = mats[l], dinvs[l]
        if A.size == 0:
            return x
        if smoother == "gmres":
            return gmres_smooth(A, dinv, x, b, steps, restart, omega) if steps else x
        for k in range(steps):
            first = x_zero and k == 0
            if smoother == "jacobi":
                res = b if first else b - A @ x
                x = (0.0 if first else x) + omega * (dinv(res) if callable(dinv) else dinv * res)
            else:
                x = _chebyshev(A, dinv, x, b, first, lambdas[l], cheb_degree, cheb_range)
        return x

    def cycle(l, b):
        if l == 0:
            return np.linalg.solve(mats[0], b) if b.size else b.copy()
        x = np.zeros_like(b)
        x = smooth(l, x, b, pre, True)
        res = b - mats[l] @ x
        xc = cycle(l - 1, prolongations[l - 1].T @ res)
        x = x + prolongations[l - 1] @ xc
        return smooth(l, x, b, post, False)

    return cycle(len(mats) - 1, np.asarray(r, dtype=np.float64))


# ---- right-hand side and verification sums (dense per-cell basis matrices) ----

def _dense_tables(bk):
    """Phi (Q x N) and reference-derivative tables dPhi[d] of one tensor bucket:
    Kronecker products of the 1D tables B[a] / G[a] (modes x points), axis 0
    fastest for both modes and points (src/integration.jl:2677-2742 full tables)."""
    D = bk.dim
    Phi = np.ones((1, 1))
    dPhi = [np.ones((1, 1)) for _ in range(D)]
    for a in range(D):
        Ba, Ga = bk.B[a].T, bk.G[a].T  # (Q_a x N_a)
        Phi = np.kron(Ba, Phi)
        for d in range(D):
            dPhi[d] = np.kron(Ga if d == a else Ba, dPhi[d])
    return Phi, dPhi


def rhs_from_loads(plan, loads, channels):
    """Physical right-hand side from lowered point loads (per tensor cell and
    point: W v, or W v and W g_d): r_i = sum_q phi_i(x_q) (W v)_q + sum_d
    J^-1_d dphi_i/dxi_d (W g_d)_q, scattered through the term maps
    (src/assembly.jl:488-507; src/accumulators.jl:1230-1274 for the one-sided
    form), with dense basis matrices instead of sum factorization."""
    y = np.zeros(plan.full_dof_count)
    off = 0
    for bk in plan.buckets:
        nc, Q = bk.cell_count, bk.n_points
        L = np.asarray(loads[off:off + nc * Q * channels]).reshape(nc, Q, channels)
        off += nc * Q * channels
        Phi, dPhi = _dense_tables(bk)
        local = L[:, :, 0] @ Phi
        for d in range(bk.dim if channels > 1 else 0):
            local = local + (L[:, :, 1 + d] * bk.inverse_jacobian[:, d:d + 1]) @ dPhi[d]
        contrib = bk.term_coef * local[bk.term_cell, bk.term_box]
        y += np.bincount(bk.term_dof, weights=contrib, minlength=plan.full_dof_count)
    return y


def l2_sums(plan, u_full, weights, exact):
    """(sum_q w_q (u_h(x_q) - e_q)^2, sum_q w_q e_q^2) over the tensor cells
    (src/verification.jl:201-230), u_h from the full-layout coefficients
    through the term maps and the dense value table."""
    s0 = s1 = 0.0
    off = 0
    u = np.asarray(u_full, dtype=np.float64)
    for bk in plan.buckets:
        nc, Q = bk.cell_count, bk.n_points
        Phi, _ = _dense_tables(bk)
        box = np.zeros((nc, bk.n_modes))
        np.add.at(box, (bk.term_cell, bk.term_box), bk.term_coef * u[bk.term_dof])
        uh = box @ Phi.T
        w = np.asarray(weights[off:off + nc * Q]).reshape(nc, Q)
        e = np.asarray(exact[off:off + nc * Q]).reshape(nc, Q)
        off += nc * Q
        s0 += float(np.sum(w * (uh - e) ** 2))
        s1 += float(np.sum(w * e ** 2))
    return s0, s1
