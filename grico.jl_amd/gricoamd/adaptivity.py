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
applied, so every target leaf lies inside exactly one source leaf. The
projection runs as a device solve: the target mass operator (value channel
only, no constraints) is compiled like any other cell accumulator, the load
u_old(x_q) is evaluated on the host at the target quadrature points (located
in the source grid by descending its refinement tree) and the mass system is
solved by Jacobi-CG through the device library. The target quadrature takes
max(p_source, p_target) + 1 Gauss points per axis, which integrates the mass
matrix and the load exactly, so a target space that contains the source space
reproduces the old field up to the solve tolerance.

Not built: h-derefinement (a target cell covering several source leaves needs
its load integrated leaf by leaf, src/transfer.jl:22-25) and the automatic
indicator-driven planning of src/indicators.jl.
"""

import numpy as np

from .mesh import Domain
from .problem import AffineProblem, ScalarField, SPD, TestChannels
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
        degrees (src/adaptivity.jl:889-940). Returns the first child."""
        leaf = self._checked_leaf(leaf)
        axis = self._checked_axis(axis)
        level = int(self.target_domain.grid.levels()[leaf, axis])
        if level + 1 > self.limits.max_h_level[axis]:
            raise ValueError(f"leaf {leaf} reached max_h_level[{axis}]")
        deg = self.degrees.pop(leaf)
        first = self.target_domain.refine(leaf, axis)
        self.degrees[first] = deg
        self.degrees[first + 1] = deg
        return first

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
        active = self.target_domain.active_leaves()
        return active[np.isin(active, np.fromiter(self.degrees, dtype=np.int64))]


class SpaceTransition:
    """Source space, target space and the source leaf holding each target
    leaf (src/transition.jl:349-387, source_leaves 143)."""

    def __init__(self, source_space, target_space, source_leaf_of_target):
        self.source_space = source_space
        self.target_space = target_space
        self.source_leaf_of_target = source_leaf_of_target

    def source_leaves(self, target_leaf):
        return [int(self.source_leaf_of_target[int(target_leaf)])]


def transition(plan):
    """Compile the target space of an AdaptivityPlan with the source space's
    basis and continuity (src/transition.jl:349-387)."""
    src = plan.source_space
    leaves = plan.target_leaves()
    degrees = np.asarray([plan.degrees[int(l)] for l in leaves], dtype=np.int32).reshape(len(leaves), plan.dim)
    options = SpaceOptions(basis=src.options.basis, degree=src.options.degree,
                           quadrature=src.options.quadrature, continuity=src.options.continuity)
    target = HpSpace(plan.target_domain, options, leaves=leaves, degrees=degrees)
    # Parent walk in the target grid until a source leaf is reached.
    parent = plan.target_domain.grid.parent()
    source_set = set(int(l) for l in src.leaves)
    owner = {}
    for l in leaves:
        c = int(l)
        while c not in source_set:
            c = int(parent[c])
            if c < 0:
                raise RuntimeError(f"target leaf {int(l)} has no source leaf")
        owner[int(l)] = c
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


def locate(domain, points):
    """Active leaf of `domain` containing each point (tuple of D arrays of one
    shape), by descending the refinement tree from the root cell."""
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
    while True:
        inner = np.nonzero(fc[cell] >= 0)[0]
        if inner.size == 0:
            return cell
        c = cell[inner]
        a = split[c]
        k = np.arange(c.size)
        mid = domain.cell_lowers(c)[k, a] + 0.5 * domain.cell_sizes(c)[k, a]
        coord = np.stack(pts)[a, inner]
        first = fc[c]
        # The child on the upper side of the split holds the larger coordinate.
        upper_child = np.where(crd[first, a] > crd[first + 1, a], first, first + 1)
        lower_child = np.where(crd[first, a] > crd[first + 1, a], first + 1, first)
        cell[inner] = np.where(coord >= mid, upper_child, lower_child)


def evaluate(state, points):
    """u_h at arbitrary points of the state's domain (tuple of D arrays of one
    shape); returns an array of that shape."""
    plan = state.plan
    space = plan.space
    dom = space.domain
    D = dom.dim
    shape = np.shape(points[0])
    pts = [np.asarray(p, dtype=np.float64).reshape(-1) for p in points]
    leaf = locate(dom, pts)
    out = np.zeros(pts[0].size)
    row_of = np.full(dom.grid.cell_count(), -1, dtype=np.int64)
    for b in plan.buckets:
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
    if getattr(plan, "cloud_buckets", None):
        raise NotImplementedError("evaluation of point-cloud cells is not supported")
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
    for l in tgt.leaves:
        s = transition.source_leaf_of_target[int(l)]
        drop = np.asarray(src.degrees[src.leaf_index[s]]) - np.asarray(tgt.degrees[tgt.leaf_index[int(l)]])
        extra = max(extra, 1 + int(drop.max()))
    options = SpaceOptions(basis=tgt.options.basis, degree=tgt.options.degree,
                           quadrature=DegreePlusQuadrature(extra), continuity=tgt.options.continuity)
    space = HpSpace(tgt.domain, options, leaves=tgt.leaves, degrees=tgt.degrees)
    u = ScalarField(space, "u")
    problem = AffineProblem(u, SPD())
    problem.add_cell_accumulator(u, u, _MassOperator())
    problem.add_cell_accumulator(u, _TransferLoad(state))
    return problem


def transfer_state(transition, state, relative_tolerance=1e-13, maxiter=None):
    """L2 projection of `state` onto the transition's target space, solved on
    the device by Jacobi-CG (src/transfer.jl:892-1000). Returns a State of the
    projection plan (full layout on the target space)."""
    from .plan import compile as compile_plan
    from .solve import CGSolver, JacobiPreconditioner, solve

    plan = compile_plan(projection_problem(transition, state))
    return solve(plan, solver=CGSolver(preconditioner=JacobiPreconditioner()),
                 relative_tolerance=relative_tolerance, maxiter=maxiter)
