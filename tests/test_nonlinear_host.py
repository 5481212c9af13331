"""Residual problems and Newton: host plan and oracle (CPU), pinned by the
reference's finite-difference tangent check (test/accumulators.jl:194-200,
305-312), consistency with the affine path for a linear residual, and the
Bratu example's L2 golden (test/examples.jl:30-35) with Newton's quadratic
convergence."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle as O
from gricoamd.nonlinear import ResidualPlan, ResidualProblem

import problems as P


def _fd_tangent(rplan, u, du, h=1e-7):
    return (O.residual(rplan, u + h * du) - O.residual(rplan, u)) / h


class _Exp:
    """test/accumulators.jl:122-128: residual (exp u, grad u), tangent (exp u du, grad du)."""

    def cell_residual_accumulate(self, q, state, test_component):
        return g.TestChannels(np.exp(state.value), tuple(state.gradient))

    def cell_tangent_accumulate(self, q, state, increment, test_component):
        return g.TestChannels(np.exp(state.value) * increment.value, tuple(increment.gradient))


def _dg_square():
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (1, 1))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2), continuity="dg"))
    return g.ScalarField(space, "u")


def test_tangent_matches_finite_difference_reference_case():
    u = _dg_square()
    pr = ResidualProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, _Exp())
    rplan = ResidualPlan(pr)
    n = rplan.full_dof_count
    x = np.array([0.1 * np.sin(0.19 * i) for i in range(1, n + 1)])
    dx = np.array([np.cos(0.29 * i) for i in range(1, n + 1)])
    np.testing.assert_allclose(O.tangent_apply(rplan, x, dx), _fd_tangent(rplan, x, dx), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("refine", [(), ((5, 0), (-1, 1))])
def test_tangent_matches_finite_difference_bratu(refine):
    rplan = ResidualPlan(P.bratu_problem((3, 3), 3, refine=refine))
    rng = np.random.default_rng(2)
    x = 0.3 * rng.standard_normal(rplan.full_dof_count)
    x[rplan.plan.fixed_dofs] = 0.0
    dx = rng.standard_normal(rplan.full_dof_count)
    dx[rplan.plan.fixed_dofs] = 0.0
    np.testing.assert_allclose(O.tangent_apply(rplan, x, dx), _fd_tangent(rplan, x, dx), rtol=2e-6, atol=2e-6)


class _LinearResidual:
    def __init__(self):
        self.src = P.SineLoad(2)

    def cell_residual_accumulate(self, q, state, test_component):
        return g.TestChannels(-self.src.cell_rhs_accumulate(q, 0), tuple(state.gradient))

    def cell_tangent_accumulate(self, q, state, increment, test_component):
        return g.TestChannels(0.0 * increment.value, tuple(increment.gradient))


def test_linear_residual_is_affine_apply_minus_rhs():
    affine = g.compile(P.poisson(2, 3, 3))
    u = affine.problem.field
    pr = ResidualProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, _LinearResidual())
    for con in affine.problem.constraints:
        pr.add_constraint(con)
    rplan = ResidualPlan(pr)
    x = np.random.default_rng(4).standard_normal(affine.full_dof_count)
    free = affine.solve_dofs
    expect = O.apply(affine, x, physical=True) - affine.rhs_physical()
    np.testing.assert_allclose(O.residual(rplan, x)[free], expect[free], atol=1e-12)


def test_bratu_example_golden_and_quadratic_convergence():
    rplan = ResidualPlan(P.bratu_problem((2, 2), 2, lam=0.5, amplitude=0.1))
    u, norms = O.newton_dense(rplan)
    err = g.l2_error(g.State(rplan.plan, u), P.bratu_exact(0.1))
    assert err <= 2.0e-3, err
    assert len(norms) <= 6
    # quadratic: each contraction at least squares the previous relative residual
    rel = [n / norms[0] for n in norms]
    assert rel[-1] < 1e-10 and rel[2] <= 10 * rel[1] ** 2 + 1e-14
