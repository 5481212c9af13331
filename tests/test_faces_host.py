"""Boundary-face and interface accumulators: host plan and oracle, pinned by
the reference's goldens and by closed forms (CPU only).

Pins: the boundary-mass golden of test/matrix_free.jl:837-846, the 1D jump
penalty worked out by hand, constants in the kernel of a cell-size-scaled jump
penalty (test/matrix_free.jl:906-920), interface loads integrated against the
constant function (total interface measure), and polynomial exactness of a
symmetric interior penalty DG solve (the DG space holds the quadratic exact
solution, so consistency makes the discrete solution exact up to rounding;
a wrong face weight, point, normal or side breaks it).
"""

import numpy as np
import pytest

import gricoamd as g
import pyoracle as O

from problems import (_refined_grid, GradientJumpPenalty, InterfaceAffine, OneSidedAffine, quadratic_exact, sipg_poisson)


class _BoundaryMass:
    def boundary_accumulate(self, q, trial, test_component):
        return 3.0 * trial.value


class _Jump:
    def __init__(self, scale):
        self.scale = scale

    def interface_accumulate(self, q, trial, test_component):
        c = self.scale(q) * g.jump(trial.value)
        z = tuple(0.0 * c for _ in q.normal)
        return g.TraceTestChannels(g.TestChannels(-c, z), g.TestChannels(c, z))


class _PlusLoad:
    def __init__(self, c):
        self.c = c

    def interface_rhs_accumulate(self, q, test_component):
        z = tuple(0.0 * x for x in q.point)
        return g.TraceTestChannels(g.TestChannels(0.0 * q.point[0], z), g.TestChannels(self.c + 0.0 * q.point[0], z))


def _dg(dim, cells, degree, refine=()):
    domain = _refined_grid(dim, cells, refine)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree),
                                             continuity="dg"))
    return g.ScalarField(space, "u")


def _constant_one(space):
    """Coefficients of the constant 1: every vertex mode (indices in {0, 1}) is 1."""
    x = np.zeros(space.dof_count)
    for leaf in space.leaves:
        for m in space.local_modes(leaf):
            if all(k <= 1 for k in m):
                for dof, coef in space.mode_terms(leaf, m):
                    x[dof] += coef
    return x


def test_boundary_mass_golden():
    u = _dg(1, 1, 1)
    problem = g.AffineProblem(u, g.SPD())
    problem.add_boundary_accumulator(g.BoundaryFace(0, g.UPPER), u, u, _BoundaryMass())
    plan = g.compile(problem)
    np.testing.assert_allclose(O.apply(plan, np.array([0.25, -0.5])), [0.0, -1.5], atol=1e-15)


def test_interface_jump_closed_form():
    """Two 1D cells, p = 1: the minus trace is the minus cell's right vertex
    mode, the plus trace the plus cell's left one; y = 2 [u] (-e_minus + e_plus)."""
    u = _dg(1, 2, 1)
    problem = g.AffineProblem(u, g.SPD())
    problem.add_interface_accumulator(u, u, _Jump(lambda q: 2.0))
    plan = g.compile(problem)
    sp = u.space
    minus_right = sp.mode_terms(sp.leaves[int(sp.interface_minus[0])], (1,))[0][0]
    plus_left = sp.mode_terms(sp.leaves[int(sp.interface_plus[0])], (0,))[0][0]
    x = np.array([0.25, -0.5, 0.75, -0.125])
    jump = x[plus_left] - x[minus_right]
    expect = np.zeros(4)
    expect[minus_right] = -2.0 * jump
    expect[plus_left] = 2.0 * jump
    np.testing.assert_allclose(O.apply(plan, x), expect, atol=1e-15)
    d = O.diagonal_physical(plan)
    np.testing.assert_allclose(d, np.diag(O.reduced_matrix(plan)), atol=1e-14)


@pytest.mark.parametrize("dim,refine", [(2, ()), (2, ((0, 0), (-1, 1))), (3, ((0, 2),))])
def test_jump_penalty_annihilates_constants(dim, refine):
    u = _dg(dim, 2, 2, refine)
    problem = g.AffineProblem(u, g.SPD())
    problem.add_interface_accumulator(u, u, _Jump(lambda q: np.minimum.reduce([c for c in q.cell_size])))
    plan = g.compile(problem)
    one = _constant_one(u.space)
    np.testing.assert_allclose(O.apply(plan, one), 0.0, atol=1e-13)
    assert np.abs(O.apply(plan, np.random.default_rng(1).standard_normal(len(one)))).max() > 1e-3


@pytest.mark.parametrize("dim,cells,measure", [(2, 3, 4.0), (3, 2, 3.0)])
def test_interface_load_against_constant(dim, cells, measure):
    """sum_i b_i 1_i = c * (total interior interface measure)."""
    u = _dg(dim, cells, 2)
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, _Dummy())
    problem.add_interface_accumulator(u, _PlusLoad(0.7))
    plan = g.compile(problem)
    b = plan.rhs_physical()
    assert abs(b @ _constant_one(u.space) - 0.7 * measure) < 1e-12


class _Dummy:
    def cell_accumulate(self, q, trial, test_component):
        return 0.0 * trial.value


@pytest.mark.parametrize("dim,cells,refine,degree", [
    (2, 3, (), 2), (2, 3, ((4, 0), (-1, 1)), 2), (2, 2, ((0, 0),), 3), (3, 2, ((0, 1),), 2)])
def test_sipg_polynomial_exactness(dim, cells, refine, degree):
    problem = sipg_poisson(dim=dim, cells=cells, degree=degree, refine=refine)
    plan = g.compile(problem)
    A = O.reduced_matrix(plan)
    np.testing.assert_allclose(A, A.T, atol=1e-11 * np.abs(A).max())
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    x = np.linalg.solve(A, plan.reduced_rhs())
    full = np.zeros(plan.full_dof_count)
    full[plan.solve_dofs] = x
    err = g.l2_error(g.State(plan, full), quadratic_exact)
    assert err < 1e-10, err


def test_sipg_p_mismatch_exactness():
    """Leaves of degree 2 and 3 side by side: the quadratic is still in the space."""
    problem = sipg_poisson(dim=2, cells=3, degrees=lambda domain, leaf: 3 if leaf % 2 else 2)
    plan = g.compile(problem)
    assert len({(b.degrees, b.plus_degrees) for b in plan.interface_buckets}) > 1
    A = O.reduced_matrix(plan)
    x = np.linalg.solve(A, plan.reduced_rhs())
    full = np.zeros(plan.full_dof_count)
    full[plan.solve_dofs] = x
    assert g.l2_error(g.State(plan, full), quadratic_exact) < 1e-10


def test_nonsymmetric_face_forms_diagonal():
    """Diagonal of the reference's nonsymmetric one-sided and interface test
    forms equals the diagonal of the assembled matrix (test/matrix_free.jl:514-535)."""
    u = _dg(2, 2, 2, ((0, 0),))
    problem = g.AffineProblem(u, g.NonsymmetricOperator())
    problem.add_boundary_accumulator(g.BoundaryFace(0, g.UPPER), u, u, OneSidedAffine(1.1, 0.2, 0.3, 0.4))
    problem.add_interface_accumulator(u, u, InterfaceAffine(1.4, -0.3, 0.2))
    plan = g.compile(problem)
    A = O.reduced_matrix(plan)
    np.testing.assert_allclose(O.diagonal_physical(plan), np.diag(A), atol=1e-13)
    assert np.abs(A - A.T).max() > 1e-3


def test_gradient_jump_penalty_on_cg_space_kills_global_quadratics():
    """[d_n u] vanishes for a global polynomial on a CG space (interface
    operator on shared trace dofs)."""
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (3, 3))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_interface_accumulator(u, u, GradientJumpPenalty(0.1))
    plan = g.compile(problem)
    # interpolate a quadratic by an L2 fit on random points of each leaf: the
    # least-squares problem over the whole space reproduces it exactly
    from gricoamd.quadrature import fe_basis_values_and_derivatives
    rng = np.random.default_rng(3)
    rows, rhs = [], []
    for li, leaf in enumerate(space.leaves):
        c, h = domain.cell_centers_and_half_sizes([leaf])
        xi = rng.uniform(-1, 1, size=(12, 2))
        pts = c + xi * h
        tv = [fe_basis_values_and_derivatives(xi[:, a], 2)[0] for a in range(2)]
        for k in range(12):
            row = np.zeros(space.dof_count)
            for m in space.local_modes(leaf):
                phi = tv[0][k, m[0]] * tv[1][k, m[1]]
                for dof, coef in space.mode_terms(leaf, m):
                    row[dof] += coef * phi
            rows.append(row)
            rhs.append(quadratic_exact((pts[k, 0], pts[k, 1])))
    x = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
    np.testing.assert_allclose(O.apply(plan, x), 0.0, atol=1e-11)
    assert np.abs(O.apply(plan, rng.standard_normal(space.dof_count))).max() > 1e-4
