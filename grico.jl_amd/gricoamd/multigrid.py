"""hp geometric multigrid: hierarchy, reduced transfers and the device V-cycle.

Reference: src/multigrid.jl. The policy mirrors GeometricMultigridPreconditioner
(51-96); the hierarchy first halves the degrees on the fine mesh
(_hp_multigrid_spaces 443-469, _coarsened_degree 473-499) and then merges one
binary refinement frontier per level (_h_coarsened_space 514-561,
_multigrid_h_coarsening_candidates 568-586, parent degrees = min of the
children 563-566). Every level is the problem rediscretized on its space
(_problem_on_fields 596-612). Transfers are built by the host library
(gh_transfer_create, src/multigrid.jl:614-857) and the V-cycle runs on the
device (gmf_precond_create_mg, src/multigrid.jl:858-1087).

`h_steps_per_level` (no reference counterpart; default 1, the reference's
hierarchy) merges that many binary refinement frontiers into one h-level: the
intermediate levels are compiled for their transfers only, and the kept
levels' prolongation is the product of the intermediate ones (exact, since the
spaces are nested). With 3 on a uniformly refined 3D mesh each h-level halves
the mesh along every axis, which removes the small semi-coarsened levels whose
V-cycle work is launch-latency bound.

Smoothers: "jacobi" is the reference's damped Jacobi V(pre, post) (943-956);
"chebyshev" is a Chebyshev polynomial in D^-1 A on [lambda_max / range,
lambda_max] with lambda_max estimated by Lanczos (no reference counterpart).
"""

import ctypes
import math

import numpy as np

from . import _native
from .plan import AssemblyPlan, compile as compile_plan
from .problem import SPD, AffineProblem, Dirichlet, MeanValue, ScalarField
from .space import HpSpace, SpaceOptions, StoredDegrees

_DEFAULT_COARSE_DIRECT_DOF_LIMIT = 512
DENSE_COARSE_MAX = 46340  # largest n with n^2 < 2^31 (rocSOLVER 32-bit sizes)

SMOOTHERS = {"jacobi": 0, "chebyshev": 1, "gmres": 2}


def _positive_int(v, name):
    if int(v) != v or int(v) < 1:
        raise ValueError(f"{name} must be a positive integer")
    return int(v)


def _nonnegative_int(v, name):
    if int(v) != v or int(v) < 0:
        raise ValueError(f"{name} must be a non-negative integer")
    return int(v)


def _tolerance(v, name):
    v = float(v)
    if not (math.isfinite(v) and v >= 0.0):
        raise ValueError(f"{name} must be finite and non-negative")
    return v


class GeometricMultigridPreconditioner:
    """hp geometric multigrid policy (src/multigrid.jl:51-96).

    Additions without a reference counterpart: the Chebyshev-Jacobi smoother
    (`chebyshev_degree`, `chebyshev_range`, `eigen_iterations`), `graph`
    (replay the V-cycle as a captured HIP graph), `precision`: "fp64", or
    "fp32" for smoothers, transfers and level operators in FP32 (coarse solve,
    diagonals and eigenvalue estimates in FP64) inside the FP64 outer Krylov
    method, `h_steps_per_level` and `cell_blocks`: "cut" smooths with one
    dense block per cell that carries its own rule (finite cells cut by the
    boundary) on every smoothing level, point Jacobi elsewhere; "cells" with
    one block per cell of at most 512 reduced dofs (larger cells keep point
    Jacobi)."""

    def __init__(self, min_degree=1, max_levels=16, p_sequence="bisect", pre_smoothing_steps=None,
                 post_smoothing_steps=None, smoother_damping=0.8, smoother="auto", smoother_restart=8,
                 coarse_relative_tolerance=1.0e-12, coarse_absolute_tolerance=0.0, coarse_maxiter=10_000,
                 coarse_direct_dof_limit=_DEFAULT_COARSE_DIRECT_DOF_LIMIT, chebyshev_degree=4,
                 chebyshev_range=20.0, eigen_iterations=12, graph=True, precision="fp64", h_steps_per_level=1,
                 cell_blocks=None):
        self.min_degree = _nonnegative_int(min_degree, "min_degree")
        self.max_levels = _positive_int(max_levels, "max_levels")
        if p_sequence not in ("bisect", "decrease_by_one", "go_to_one"):
            raise ValueError("p_sequence must be 'bisect', 'decrease_by_one', or 'go_to_one'")
        self.p_sequence = p_sequence
        if smoother not in ("auto", "jacobi", "gmres", "chebyshev"):
            raise ValueError("smoother must be 'auto', 'jacobi', 'gmres', or 'chebyshev'")
        self.smoother = smoother
        default_steps = 1 if smoother == "chebyshev" else 2
        self.pre_smoothing_steps = _nonnegative_int(
            default_steps if pre_smoothing_steps is None else pre_smoothing_steps, "pre_smoothing_steps")
        self.post_smoothing_steps = _nonnegative_int(
            default_steps if post_smoothing_steps is None else post_smoothing_steps, "post_smoothing_steps")
        damping = float(smoother_damping)
        if not (math.isfinite(damping) and damping > 0.0):
            raise ValueError("smoother_damping must be finite and positive")
        self.smoother_damping = damping
        self.smoother_restart = _positive_int(smoother_restart, "smoother_restart")
        self.coarse_relative_tolerance = _tolerance(coarse_relative_tolerance, "coarse_relative_tolerance")
        self.coarse_absolute_tolerance = _tolerance(coarse_absolute_tolerance, "coarse_absolute_tolerance")
        self.coarse_maxiter = _positive_int(coarse_maxiter, "coarse_maxiter")
        self.coarse_direct_dof_limit = _nonnegative_int(coarse_direct_dof_limit, "coarse_direct_dof_limit")
        self.chebyshev_degree = _positive_int(chebyshev_degree, "chebyshev_degree")
        if not (math.isfinite(float(chebyshev_range)) and float(chebyshev_range) > 1.0):
            raise ValueError("chebyshev_range must be greater than one")
        self.chebyshev_range = float(chebyshev_range)
        self.eigen_iterations = _positive_int(eigen_iterations, "eigen_iterations")
        self.graph = bool(graph)
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self.h_steps_per_level = _positive_int(h_steps_per_level, "h_steps_per_level")
        if cell_blocks not in (None, "cut", "cells"):
            raise ValueError("cell_blocks must be None, 'cut' or 'cells'")
        self.cell_blocks = cell_blocks

    def resolved_smoother(self, operator_class):
        if self.smoother == "auto":
            return "jacobi" if isinstance(operator_class, SPD) else "gmres"
        return self.smoother


# ---- hierarchy of spaces ----------------------------------------------------------

def _coarsened_degree(degree, lower, sequence):
    if degree <= lower:
        return degree
    if sequence == "bisect":
        return max(degree // 2, lower)
    if sequence == "decrease_by_one":
        return max(degree - 1, lower)
    return lower


def coarsened_degrees(degrees, continuity, min_degree, sequence):
    """Per-leaf degree tuples one p-level down (src/multigrid.jl:473-499)."""
    out = np.array(degrees, dtype=np.int32, copy=True)
    for a, kind in enumerate(continuity):
        floor = max(min_degree, 1 if kind == "cg" else 0)
        for i in range(out.shape[0]):
            d = int(out[i, a])
            out[i, a] = _coarsened_degree(d, min(d, floor), sequence)
    return out


def _space_with(space, leaves, degrees):
    opts = space.options
    options = SpaceOptions(basis=opts.basis, degree=StoredDegrees(leaves, degrees), quadrature=opts.quadrature,
                           continuity=opts.continuity)
    return HpSpace(space.domain, options, leaves=leaves, degrees=degrees)


def h_coarsening_candidates(space):
    """(parents, first children) of binary splits whose two children are both
    active leaves of `space` (src/multigrid.jl:568-586)."""
    grid = space.domain.grid
    first_child = np.asarray(grid.first_child())
    active = np.zeros(len(first_child), dtype=bool)
    active[space.leaves] = True
    cells = np.nonzero(first_child >= 0)[0]
    fc = first_child[cells]
    ok = active[fc] & active[fc + 1]
    return cells[ok], fc[ok]


def _balanced(space, parents, firsts):
    """Mask of the candidates whose merge keeps every neighbour within a factor
    two per axis: all leaves touching the parent's box have the children's
    size. Physical domains need it: leaves whose sibling is outside never
    merge, so without the check size ratios across a face grow by a factor two
    per level, and the continuity constraints of such faces (a leaf hanging on
    a neighbour four or more times its size) no longer give nested spaces.
    Parent boxes are compared on a lattice of the finest leaf size."""
    dom = space.domain
    leaves = space.leaves
    origin = np.asarray(dom.origin, dtype=np.float64)
    sizes = dom.cell_sizes(leaves)
    unit = sizes.min(axis=0)
    lo = np.rint((dom.cell_lowers(leaves) - origin) / unit).astype(np.int64)
    ext = np.rint(sizes / unit).astype(np.int64)
    shape = (lo + ext).max(axis=0) + 2
    owner = np.full(tuple(shape), -1, dtype=np.int64)
    D = space.dim
    for e in np.unique(ext, axis=0):
        sel = np.nonzero(np.all(ext == e, axis=1))[0]
        offs = np.stack(np.meshgrid(*[np.arange(int(v)) for v in e], indexing="ij"), axis=-1).reshape(-1, D)
        cells = lo[sel][:, None, :] + offs[None, :, :]
        owner[tuple(cells.reshape(-1, D).T + 1)] = np.repeat(sel, len(offs))
    ok = np.zeros(len(parents), dtype=bool)
    psize = dom.cell_sizes(parents)
    plo = np.rint((dom.cell_lowers(parents) - origin) / unit).astype(np.int64)
    pext = np.rint(psize / unit).astype(np.int64)
    csize = dom.cell_sizes(firsts)
    for e in np.unique(pext, axis=0):
        sel = np.nonzero(np.all(pext == e, axis=1))[0]
        box = np.stack(np.meshgrid(*[np.arange(-1, int(v) + 1) for v in e], indexing="ij"), axis=-1).reshape(-1, D)
        inner = np.all((box >= 0) & (box < e), axis=1)
        shell = box[~inner]
        cells = plo[sel][:, None, :] + shell[None, :, :] + 1
        cells = np.clip(cells, 0, shape - 1)
        nb = owner[tuple(cells.reshape(-1, D).T)].reshape(len(sel), len(shell))
        nb_size = np.where(nb[..., None] >= 0, sizes[np.maximum(nb, 0)], csize[sel][:, None, :])
        ok[sel] = np.all(np.isclose(nb_size, csize[sel][:, None, :], rtol=1e-12, atol=0.0), axis=(1, 2))
    return ok


def h_coarsened_space(space):
    parents, firsts = h_coarsening_candidates(space)
    if parents.size and getattr(space.domain, "classify", None) is not None:
        keep = _balanced(space, parents, firsts)
        parents, firsts = parents[keep], firsts[keep]
    if parents.size == 0:
        return None
    index = space.leaf_index
    deg = space.degrees
    child0 = np.fromiter((index[int(c)] for c in firsts), dtype=np.int64, count=len(firsts))
    child1 = np.fromiter((index[int(c) + 1] for c in firsts), dtype=np.int64, count=len(firsts))
    parent_degrees = np.minimum(deg[child0], deg[child1])
    drop = np.zeros(len(space.leaves), dtype=bool)
    drop[child0] = True
    drop[child1] = True
    leaves = np.concatenate([space.leaves[~drop], parents.astype(np.int64)])
    degrees = np.concatenate([deg[~drop], parent_degrees]).astype(np.int32)
    # Physical domains keep only domain-active leaves (src/adaptivity.jl:526-551).
    classify = getattr(space.domain, "classify", None)
    if classify is not None:
        from .embedded import OUTSIDE
        keep = classify(leaves) != OUTSIDE
        leaves, degrees = leaves[keep], degrees[keep]
    return _space_with(space, leaves, degrees)


def space_chain(fine_space, policy):
    """Coarse-to-fine list of (space, kept): the p-levels, then one entry per
    merged binary refinement frontier; with h_steps_per_level = k only every
    k-th h-entry (and the coarsest) is a level of the hierarchy, the others
    only carry transfers (src/multigrid.jl:443-469 for k = 1)."""
    chain = [(fine_space, True)]
    kept = 1
    current = fine_space
    degrees = np.asarray(fine_space.degrees, dtype=np.int32)
    while kept < policy.max_levels:
        nxt = coarsened_degrees(degrees, current.continuity, policy.min_degree, policy.p_sequence)
        if np.array_equal(nxt, degrees):
            break
        current = _space_with(current, current.leaves, nxt)
        chain.append((current, True))
        kept += 1
        degrees = nxt
    while kept < policy.max_levels:
        group = []
        cur = current
        for _ in range(policy.h_steps_per_level):
            nxt = h_coarsened_space(cur)
            if nxt is None:
                break
            group.append(nxt)
            cur = nxt
        if not group:
            break
        chain.extend((sp, False) for sp in group[:-1])
        chain.append((group[-1], True))
        kept += 1
        current = group[-1]
    chain.reverse()
    return chain


def multigrid_spaces(fine_space, policy):
    """Coarse-to-fine list of level spaces (src/multigrid.jl:443-469)."""
    return [sp for sp, kept in space_chain(fine_space, policy) if kept]


def has_coarsening_level(space, policy):
    if policy.max_levels <= 1:
        return False
    degrees = np.asarray(space.degrees, dtype=np.int32)
    if not np.array_equal(coarsened_degrees(degrees, space.continuity, policy.min_degree, policy.p_sequence),
                          degrees):
        return True
    return h_coarsening_candidates(space)[0].size > 0


def problem_on_space(problem, space):
    """The same operators and constraints on another space (src/multigrid.jl:596-612):
    cell, boundary-face, interface and surface operators, surface pieces,
    embedded surfaces and cell rules carry over."""
    old = problem.field
    field = ScalarField(space, old.name)
    out = AffineProblem(field, problem.operator_class)
    for (_, _, op) in problem.cell_operators:
        out.add_cell_accumulator(field, field, op)
    for (_, op) in problem.cell_rhs:
        out.add_cell_accumulator(field, op)
    for (bf, _, _, op) in getattr(problem, "boundary_operators", []):
        out.add_boundary_accumulator(bf, field, field, op)
    for (bf, _, op) in getattr(problem, "boundary_rhs", []):
        out.add_boundary_accumulator(bf, field, op)
    for (_, _, op) in getattr(problem, "interface_operators", []):
        out.add_interface_accumulator(field, field, op)
    for (_, op) in getattr(problem, "interface_rhs", []):
        out.add_interface_accumulator(field, op)
    for (_, _, op) in getattr(problem, "surface_operators", []):
        out.add_surface_accumulator(field, field, op)
    for (_, op) in getattr(problem, "surface_rhs", []):
        out.add_surface_accumulator(field, op)
    # Explicit pieces and rules are tied to leaves: compile() rejects them on a
    # level where their leaf is not active (as the reference does); embedded
    # surfaces and a physical domain's cut-cell rules follow every level.
    for sq in getattr(problem, "surface_quadratures", []):
        out.add_surface_quadrature(sq)
    for es in getattr(problem, "embedded_surfaces", []):
        out.add_embedded_surface(es)
    for leaf, rule in getattr(problem, "cell_quadratures", {}).items():
        out.add_cell_quadrature(leaf, rule)
    for con in problem.constraints:
        out.add_constraint(Dirichlet(field, con.boundary, con.data))
    for con in getattr(problem, "mean_constraints", []):
        out.add_constraint(MeanValue(field, con.target))
    return out


# ---- transfers ------------------------------------------------------------------------

class _SpaceView(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("leaf_count", ctypes.c_int64),
        ("leaves", _native.P_i64), ("degrees", _native.P_i32), ("mode_offsets", _native.P_i64),
        ("mode_box", _native.P_i32), ("term_offsets", _native.P_i64), ("term_indices", _native.P_i64),
        ("term_coefficients", _native.P_f64), ("full_dof_count", ctypes.c_int64),
        ("map_offsets", _native.P_i64), ("map_indices", _native.P_i64), ("map_coefficients", _native.P_f64),
        ("reduced_index", _native.P_i64),
    ]


def _declare_transfer(lib):
    if getattr(lib, "_transfer_declared", False):
        return
    lib.gh_transfer_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(_SpaceView), ctypes.POINTER(_SpaceView),
                                       ctypes.POINTER(ctypes.c_void_p)]
    lib.gh_transfer_destroy.argtypes = [ctypes.c_void_p]
    lib.gh_transfer_nnz.argtypes = [ctypes.c_void_p]
    lib.gh_transfer_nnz.restype = ctypes.c_int64
    lib.gh_transfer_export.argtypes = [ctypes.c_void_p, _native.P_i64, _native.P_i64, _native.P_f64]
    lib._transfer_declared = True


def _view(plan, keep):
    sp = plan.space
    arrays = [np.ascontiguousarray(a) for a in (
        sp.leaves.astype(np.int64), sp.degrees.astype(np.int32), sp.mode_offsets.astype(np.int64),
        sp.mode_box.astype(np.int32), sp.term_offsets.astype(np.int64), sp.term_indices.astype(np.int64),
        sp.term_coefficients.astype(np.float64), plan.map_offsets.astype(np.int64),
        plan.map_indices.astype(np.int64), plan.map_coefficients.astype(np.float64),
        plan.reduced_index.astype(np.int64))]
    keep.extend(arrays)
    v = _SpaceView()
    v.dim = sp.dim
    v.leaf_count = len(sp.leaves)
    P = _native.ptr
    v.leaves = P(arrays[0], _native.c_int64)
    v.degrees = P(arrays[1], _native.c_int32)
    v.mode_offsets = P(arrays[2], _native.c_int64)
    v.mode_box = P(arrays[3], _native.c_int32)
    v.term_offsets = P(arrays[4], _native.c_int64)
    v.term_indices = P(arrays[5], _native.c_int64)
    v.term_coefficients = P(arrays[6], _native.c_double)
    v.full_dof_count = plan.full_dof_count
    v.map_offsets = P(arrays[7], _native.c_int64)
    v.map_indices = P(arrays[8], _native.c_int64)
    v.map_coefficients = P(arrays[9], _native.c_double)
    v.reduced_index = P(arrays[10], _native.c_int64)
    return v


class ReducedTransfer:
    """Prolongation from `coarse_plan` to `fine_plan` as COO entries sorted by
    (fine, coarse) (src/multigrid.jl:115-121)."""

    def __init__(self, coarse_plan, fine_plan):
        lib = _native.host()
        _declare_transfer(lib)
        keep = []
        cv, fv = _view(coarse_plan, keep), _view(fine_plan, keep)
        handle = ctypes.c_void_p()
        _native.check_host(lib.gh_transfer_create(coarse_plan.space.domain.grid.handle, ctypes.byref(cv),
                                                  ctypes.byref(fv), ctypes.byref(handle)))
        try:
            nnz = lib.gh_transfer_nnz(handle)
            self.coarse_indices = np.zeros(nnz, dtype=np.int64)
            self.fine_indices = np.zeros(nnz, dtype=np.int64)
            self.coefficients = np.zeros(nnz, dtype=np.float64)
            _native.check_host(lib.gh_transfer_export(handle, _native.ptr(self.coarse_indices, _native.c_int64),
                                                      _native.ptr(self.fine_indices, _native.c_int64),
                                                      _native.ptr(self.coefficients, _native.c_double)))
        finally:
            lib.gh_transfer_destroy(handle)
        self.coarse_plan = coarse_plan
        self.fine_plan = fine_plan

    @classmethod
    def compose(cls, transfers):
        """The prolongation from the first transfer's coarse plan to the last
        one's fine plan through the intermediate levels (product of the
        sparse factors), as COO entries sorted by (fine, coarse)."""
        import scipy.sparse as sps
        mats = [sps.csr_matrix((t.coefficients, (t.fine_indices, t.coarse_indices)),
                               shape=(t.fine_plan.reduced_dof_count, t.coarse_plan.reduced_dof_count))
                for t in transfers]
        P = mats[0]
        for m in mats[1:]:
            P = m @ P
        P = sps.csr_matrix(P)
        P.eliminate_zeros()
        P.sort_indices()
        out = object.__new__(cls)
        out.fine_indices = np.repeat(np.arange(P.shape[0], dtype=np.int64), np.diff(P.indptr))
        out.coarse_indices = P.indices.astype(np.int64)
        out.coefficients = P.data.astype(np.float64)
        out.coarse_plan = transfers[0].coarse_plan
        out.fine_plan = transfers[-1].fine_plan
        return out

    @property
    def nnz(self):
        return len(self.coefficients)

    def prolongate_add(self, fine, coarse):
        """fine += P coarse on host arrays (src/multigrid.jl:1060-1072)."""
        np.add.at(fine, self.fine_indices, self.coefficients * coarse[self.coarse_indices])
        return fine

    def restrict(self, fine):
        """P^T fine on host arrays (src/multigrid.jl:1074-1087)."""
        out = np.zeros(self.coarse_plan.reduced_dof_count)
        np.add.at(out, self.coarse_indices, self.coefficients * fine[self.fine_indices])
        return out

    def matrix(self):
        P = np.zeros((self.fine_plan.reduced_dof_count, self.coarse_plan.reduced_dof_count))
        P[self.fine_indices, self.coarse_indices] = self.coefficients
        return P


# ---- compiled hierarchy ------------------------------------------------------------

def _check_smoother(policy, smoother):
    """src/multigrid.jl:350-357: the GMRES smoother's Krylov length covers the
    larger smoothing count. It runs on FP64 levels only."""
    if smoother == "gmres":
        if policy.smoother_restart < max(policy.pre_smoothing_steps, policy.post_smoothing_steps):
            raise ValueError("GMRES multigrid smoother requires smoother_restart to be at least "
                             "max(pre_smoothing_steps, post_smoothing_steps)")
        if policy.precision != "fp64":
            raise ValueError("the GMRES multigrid smoother runs on FP64 levels only")


class MultigridHierarchy:
    """Level plans (coarse to fine), transfers and, once requested, the device
    V-cycle (src/multigrid.jl:309-345)."""

    def __init__(self, problem, policy, fine_plan=None):
        if not isinstance(problem, AffineProblem):
            raise TypeError("GeometricMultigridPreconditioner requires the original AffineProblem so every level "
                            "can rediscretize the operators")
        fine_space = problem.field.space
        if not has_coarsening_level(fine_space, policy):
            raise ValueError("GeometricMultigridPreconditioner requires at least one admissible h or p coarsening "
                             "level; use an explicit non-GMG preconditioner for this problem")
        self.policy = policy
        self.problem = problem
        self.smoother = policy.resolved_smoother(problem.operator_class)
        _check_smoother(policy, self.smoother)
        chain = space_chain(fine_space, policy)
        plans = []
        for i, (sp, _) in enumerate(chain):
            if i == len(chain) - 1 and fine_plan is not None:
                plans.append(fine_plan)
            elif i == len(chain) - 1:
                plans.append(compile_plan(problem))
            else:
                plans.append(compile_plan(problem_on_space(problem, sp)))
        steps = [ReducedTransfer(plans[i], plans[i + 1]) for i in range(len(plans) - 1)]
        kept = [i for i, (_, k) in enumerate(chain) if k]
        self.spaces = [chain[i][0] for i in kept]
        self.plans = [plans[i] for i in kept]
        self.transfers = [steps[a] if b == a + 1 else ReducedTransfer.compose(steps[a:b])
                          for a, b in zip(kept[:-1], kept[1:])]
        self._device = None

    def with_policy(self, policy):
        """The same level plans and transfers under another smoother, step
        count, precision or coarse-solve limit. A smaller max_levels keeps the
        finest levels (the hierarchy is generated from the fine end, so this is
        the hierarchy that max_levels builds); the other coarsening parameters
        must agree."""
        for name in ("min_degree", "p_sequence", "h_steps_per_level"):
            if getattr(policy, name) != getattr(self.policy, name):
                raise ValueError(f"with_policy keeps the hierarchy: {name} must agree")
        if policy.max_levels > self.policy.max_levels and len(self.plans) == self.policy.max_levels:
            raise ValueError("with_policy cannot add levels: max_levels must not grow")
        smoother = policy.resolved_smoother(self.problem.operator_class)
        _check_smoother(policy, smoother)
        other = object.__new__(MultigridHierarchy)
        other.__dict__.update(self.__dict__)
        keep = min(len(self.plans), policy.max_levels)
        if keep < len(self.plans):
            other.spaces = self.spaces[-keep:]
            other.plans = self.plans[-keep:]
            other.transfers = self.transfers[len(self.plans) - keep:]
        other.policy = policy
        other.smoother = smoother
        other._device = None
        return other

    @property
    def level_count(self):
        return len(self.plans)

    def reduced_dof_counts(self):
        return [p.reduced_dof_count for p in self.plans]

    def coarse_solver_kind(self):
        n = self.plans[0].reduced_dof_count
        return "dense_cholesky" if 0 < n <= self.policy.coarse_direct_dof_limit else "krylov"

    def _check_dense_coarse(self):
        """The device dense inverse holds n^2 < 2^31 entries (rocSOLVER's
        32-bit sizes): larger dense coarse levels are rejected up front."""
        n = self.plans[0].reduced_dof_count
        if self.coarse_solver_kind() == "dense_cholesky" and n > DENSE_COARSE_MAX:
            raise ValueError(f"dense coarse multigrid level has {n} dofs; the device dense inverse supports at most "
                             f"{DENSE_COARSE_MAX} (lower coarse_direct_dof_limit or keep more levels)")

    def device_preconditioner(self):
        """Compile the V-cycle on the device (gmf_precond_create_mg)."""
        if self._device is None:
            self._check_dense_coarse()
            from . import _devabi, ops
            from .solve import DevicePreconditioner
            lib = _native.device()
            _devabi.declare_mg(lib)
            wss = [ops._ws(p).ws for p in self.plans]
            levels = (ctypes.c_void_p * len(wss))(*[w.value for w in wss])
            tds = (_devabi.TransferDesc * len(self.transfers))()
            keep = [levels, tds]
            for i, t in enumerate(self.transfers):
                ci, fi, cf = (np.ascontiguousarray(t.coarse_indices), np.ascontiguousarray(t.fine_indices),
                              np.ascontiguousarray(t.coefficients))
                keep += [ci, fi, cf]
                tds[i].nnz = t.nnz
                tds[i].coarse_indices = _native.ptr(ci, _native.c_int64)
                tds[i].fine_indices = _native.ptr(fi, _native.c_int64)
                tds[i].coefficients = _native.ptr(cf, _native.c_double)
            pol = self.policy
            d = _devabi.MgDesc()
            d.level_count = len(wss)
            d.levels = levels
            d.transfers = tds
            d.smoother = SMOOTHERS[self.smoother]
            d.pre_steps = pol.pre_smoothing_steps
            d.post_steps = pol.post_smoothing_steps
            d.damping = pol.smoother_damping
            d.chebyshev_degree = pol.chebyshev_degree
            d.chebyshev_range = pol.chebyshev_range
            d.eigen_iterations = pol.eigen_iterations
            d.coarse_direct_limit = pol.coarse_direct_dof_limit
            d.coarse_rtol = pol.coarse_relative_tolerance
            d.coarse_atol = pol.coarse_absolute_tolerance
            d.coarse_maxiter = pol.coarse_maxiter
            d.graph = int(pol.graph)
            d.precision = 32 if pol.precision == "fp32" else 64
            d.smoother_restart = pol.smoother_restart
            d.spd = int(isinstance(self.problem.operator_class, SPD))
            # Levels whose reduced map is not a selection (affine Dirichlet
            # relations, mean rows) get their reduced diagonal from the host.
            if pol.cell_blocks is not None:
                bds = (_devabi.BlockDesc * len(self.plans))()
                self.block_colors = [0] * len(self.plans)
                for lvl in range(1, len(self.plans)):
                    blocks = (cut_cell_blocks(self.plans[lvl]) if pol.cell_blocks == "cut" else
                              cell_blocks_of(self.plans[lvl]))
                    if not blocks:
                        continue
                    off, dofs, inv, ncol = block_inverses(self.plans[lvl], blocks)
                    keep += [off, dofs, inv]
                    bds[lvl].count = len(blocks)
                    bds[lvl].offsets = _native.ptr(off, _native.c_int64)
                    bds[lvl].dofs = _native.ptr(dofs, _native.c_int32)
                    bds[lvl].inverse = _native.ptr(inv, _native.c_double)
                    self.block_colors[lvl] = ncol
                keep.append(bds)
                d.level_blocks = bds
            probe = [getattr(p, "selection_map", True) and _device_diagonal_unsupported(p) for p in self.plans]
            if not all(getattr(p, "selection_map", True) for p in self.plans) or any(probe):
                from .solve import reduced_diagonal_general
                diags = [np.ascontiguousarray(probed_diagonal(p)) if pr_ else
                         None if getattr(p, "selection_map", True) else
                         np.ascontiguousarray(reduced_diagonal_general(p), dtype=np.float64)
                         for p, pr_ in zip(self.plans, probe)]
                arr = (_native.P_f64 * len(diags))(*[_native.ptr(a, _native.c_double) if a is not None else
                                                     _native.P_f64() for a in diags])
                keep += [diags, arr]
                d.level_diagonals = arr
            handle = ctypes.c_void_p()
            _devabi.check(lib.gmf_precond_create_mg(ctypes.byref(d), ctypes.byref(handle)))
            self._device = DevicePreconditioner(self.plans[-1], handle, keep=(keep, self))
        return self._device

    def level_lambda_max(self):
        from . import _devabi
        lib = _native.device()
        M = self.device_preconditioner()
        out = []
        for lvl in range(self.level_count):
            n, lam = ctypes.c_int64(), ctypes.c_double()
            _devabi.check(lib.gmf_mg_level_info(M.handle, lvl, ctypes.byref(n), ctypes.byref(lam)))
            out.append(lam.value)
        return out


# ---- cell blocks of the block smoother --------------------------------------------

def cut_cell_blocks(plan):
    """Reduced dofs of every cell that carries its own rule (the plan's
    point-cloud cell batches), one sorted int32 array per cell."""
    if not getattr(plan, "selection_map", True):
        raise ValueError("cell blocks need a reduced map that selects dofs")
    out = []
    for cb in getattr(plan, "cloud_buckets", []):
        red = plan.reduced_index[cb.term_dof]
        keep = red >= 0
        cells, dofs = cb.term_cell[keep], red[keep]
        order = np.lexsort((dofs, cells))
        cells, dofs = cells[order], dofs[order]
        bounds = np.searchsorted(cells, np.arange(cb.cell_count + 1))
        for c in range(cb.cell_count):
            d = np.unique(dofs[bounds[c]:bounds[c + 1]])
            if len(d):
                out.append(d.astype(np.int32))
    return out


def cell_blocks_of(plan, max_dofs=512):
    """Reduced dofs of every cell of the plan's tensor and point-cloud cell
    batches with at most max_dofs of them, one sorted int32 array per cell."""
    if not getattr(plan, "selection_map", True):
        raise ValueError("cell blocks need a reduced map that selects dofs")
    out = []
    for bk in list(plan.buckets) + list(getattr(plan, "cloud_buckets", [])):
        red = plan.reduced_index[bk.term_dof]
        keep = red >= 0
        cells, dofs = bk.term_cell[keep], red[keep]
        order = np.lexsort((dofs, cells))
        cells, dofs = cells[order], dofs[order]
        bounds = np.searchsorted(cells, np.arange(bk.cell_count + 1))
        for c in range(bk.cell_count):
            d = np.unique(dofs[bounds[c]:bounds[c + 1]])
            if 0 < len(d) <= max_dofs:
                out.append(d.astype(np.int32))
    return out


def _item_dofs(plan):
    """(item_ptr, item_dofs): the reduced dofs every item of the plan (cell,
    surface piece, face) touches."""
    ptr, dofs = [np.zeros(1, dtype=np.int64)], []
    base = 0
    for bk in plan.all_buckets:
        if bk.cell_count == 0:
            continue
        red = plan.reduced_index[bk.term_dof]
        keep = red >= 0
        cells, d = bk.term_cell[keep], red[keep]
        order = np.argsort(cells, kind="stable")
        counts = np.bincount(cells, minlength=bk.cell_count)
        ptr.append(base + np.cumsum(counts))
        dofs.append(d[order])
        base += int(counts.sum())
    return np.ascontiguousarray(np.concatenate(ptr)), np.ascontiguousarray(np.concatenate(dofs).astype(np.int64))


def block_colors(plan, blocks):
    """Colors of the blocks such that no item of the plan touches dofs of two
    blocks of one color (gh_conflict_coloring): the blocks of one color can be
    probed with one apply per local position."""
    lib = _native.host()
    lib.gh_conflict_coloring.argtypes = [ctypes.c_int64, _native.P_i64, _native.P_i64, ctypes.c_int64,
                                         _native.P_i64, _native.P_i64, ctypes.c_int64, _native.P_i32,
                                         ctypes.POINTER(ctypes.c_int32)]
    iptr, idofs = _item_dofs(plan)
    lens = np.array([len(b) for b in blocks], dtype=np.int64)
    bptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum(lens)]).astype(np.int64))
    bdofs = np.ascontiguousarray(np.concatenate(blocks).astype(np.int64)) if len(blocks) else np.zeros(0, np.int64)
    colors = np.zeros(len(blocks), dtype=np.int32)
    count = ctypes.c_int32()
    _native.check_host(lib.gh_conflict_coloring(len(blocks), _native.ptr(bptr, _native.c_int64),
                                                _native.ptr(bdofs, _native.c_int64), len(iptr) - 1,
                                                _native.ptr(iptr, _native.c_int64), _native.ptr(idofs, _native.c_int64),
                                                plan.reduced_dof_count, _native.ptr(colors, _native.c_int32),
                                                ctypes.byref(count)))
    return colors


def _device_diagonal_unsupported(plan):
    """True when the device's exact reduced diagonal refuses the plan: point-
    cloud items whose modes expand into several dofs (hanging faces of cut
    cells on coarse finite-cell levels) or repeat a dof."""
    for cb in list(getattr(plan, "cloud_buckets", [])) + list(getattr(plan, "surface_buckets", [])):
        if len(cb.multi_offsets) > 1:
            return True
        li = cb.local_index
        for c in range(cb.cell_count):
            row = li[c][li[c] >= 0]
            if len(np.unique(row)) != len(row):
                return True
    return False


def probed_diagonal(plan):
    """Reduced operator diagonal by colored probing (gmf_reduced_blocks on
    one-dof blocks): one apply per color of dofs no item couples; for plans
    whose local diagonal is not exact and that the device diagonal does not
    cover (point-cloud items with multi-term expansions). The reference
    probes one unit vector per dof in that case (src/solve.jl:722-738)."""
    from . import _devabi, ops
    lib = _native.device()
    _devabi.declare_mg(lib)
    n = plan.reduced_dof_count
    blocks = [np.array([j], dtype=np.int32) for j in range(n)]
    colors = block_colors(plan, blocks)
    off = np.arange(n + 1, dtype=np.int64)
    dofs = np.arange(n, dtype=np.int32)
    desc = _devabi.BlockDesc()
    desc.count = n
    desc.offsets = _native.ptr(off, _native.c_int64)
    desc.dofs = _native.ptr(dofs, _native.c_int32)
    out = np.zeros(n)
    _devabi.check(lib.gmf_reduced_blocks(ops._ws(plan).ws, ctypes.byref(desc), _native.ptr(colors, _native.c_int32),
                                         int(colors.max()) + 1, _native.ptr(out, _native.c_double)))
    return out


def block_inverses(plan, blocks):
    """Inverses of B_b = R_b A R_b^T (row-major, one after another) from the
    device operator by colored probing (gmf_reduced_blocks); each block must be
    symmetric positive definite."""
    from . import _devabi, ops
    lib = _native.device()
    _devabi.declare_mg(lib)
    lens = np.array([len(b) for b in blocks], dtype=np.int64)
    off = np.zeros(len(blocks) + 1, dtype=np.int64)
    off[1:] = np.cumsum(lens)
    dofs = np.ascontiguousarray(np.concatenate(blocks).astype(np.int32))
    colors = block_colors(plan, blocks)
    desc = _devabi.BlockDesc()
    desc.count = len(blocks)
    desc.offsets = _native.ptr(off, _native.c_int64)
    desc.dofs = _native.ptr(dofs, _native.c_int32)
    vals = np.zeros(int((lens ** 2).sum()))
    _devabi.check(lib.gmf_reduced_blocks(ops._ws(plan).ws, ctypes.byref(desc),
                                         _native.ptr(colors, _native.c_int32), int(colors.max()) + 1,
                                         _native.ptr(vals, _native.c_double)))
    voff = np.zeros(len(blocks) + 1, dtype=np.int64)
    voff[1:] = np.cumsum(lens ** 2)
    inv = np.empty_like(vals)
    for m in np.unique(lens):
        idx = np.nonzero(lens == m)[0]
        B = np.stack([vals[voff[b]:voff[b + 1]].reshape(m, m) for b in idx])
        B = 0.5 * (B + B.transpose(0, 2, 1))
        try:
            L = np.linalg.cholesky(B)
        except np.linalg.LinAlgError:
            raise ValueError("a cell block of the multigrid smoother is not positive definite") from None
        Li = np.linalg.inv(L)
        Binv = Li.transpose(0, 2, 1) @ Li
        for k, b in enumerate(idx):
            inv[voff[b]:voff[b + 1]] = Binv[k].reshape(-1)
    return off, dofs, inv, int(colors.max()) + 1


def compile_geometric_multigrid(problem, policy=None, fine_plan=None):
    return MultigridHierarchy(problem, policy or GeometricMultigridPreconditioner(), fine_plan)


def check_cg_policy(problem, policy):
    """src/multigrid.jl:247-257: CG needs an SPD problem and a symmetric V-cycle."""
    if not isinstance(problem.operator_class, SPD):
        raise ValueError("CG with GeometricMultigridPreconditioner requires an SPD affine problem; use FGMRESSolver")
    if policy.pre_smoothing_steps != policy.post_smoothing_steps:
        raise ValueError("CG with GeometricMultigridPreconditioner requires equal pre- and post-smoothing steps so "
                         "the V-cycle remains symmetric")
    if policy.resolved_smoother(problem.operator_class) not in ("jacobi", "chebyshev"):
        raise ValueError("CG with GeometricMultigridPreconditioner requires a symmetric smoother")
