"""Device Krylov solvers against the oracle and the manufactured solution."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
import problems as P

pytestmark = pytest.mark.gpu


def test_c1_jacobi_cg(gpu):
    """Configuration 1: 2D 16x16, p=2, Jacobi-CG to 1e-10 (reference CPU path)."""
    plan = g.compile(P.poisson(2, 16, 2))
    state = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-10)
    b = plan.reduced_rhs()
    inv = 1.0 / pyoracle.diagonal_physical(plan)[plan.solve_dofs]
    x_ref, it = pyoracle.cg(plan, b, inv, rtol=1e-10)
    assert it > 0
    x_dev = state.coefficients[plan.solve_dofs]
    assert np.max(np.abs(x_dev - x_ref)) <= 1e-8 * np.max(np.abs(x_ref))
    l2 = g.l2_error(state, P.sine_exact)
    full_ref = np.zeros(plan.full_dof_count)
    full_ref[plan.solve_dofs] = x_ref
    l2_ref = g.l2_error(g.State(plan, full_ref), P.sine_exact)
    assert abs(l2 - l2_ref) <= 0.01 * l2_ref
    assert state.info.converged == 1
    assert state.info.final_residual_norm <= 1e-10 * state.info.rhs_norm


@pytest.mark.parametrize("dim,cells,degree", [(2, 8, 3), (3, 3, 3)])
def test_fgmres_matches_cg(gpu, dim, cells, degree):
    plan = g.compile(P.poisson(dim, cells, degree, kappa=P.kappa_c2 if dim == 3 else None))
    cg = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-11)
    fg = g.solve(plan, solver=g.FGMRESSolver(preconditioner=g.JacobiPreconditioner(), restart=20),
                 relative_tolerance=1e-11)
    assert np.max(np.abs(cg.coefficients - fg.coefficients)) <= 1e-8 * np.max(np.abs(cg.coefficients))


def test_unpreconditioned_cg_matches_dense(gpu):
    plan = g.compile(P.poisson(2, 3, 2))
    n = plan.reduced_dof_count
    A = np.column_stack([pyoracle.reduced_apply(plan, np.eye(n)[j]) for j in range(n)])
    b = plan.reduced_rhs()
    x, info = g.solve_reduced(plan, g.CGSolver(), b, relative_tolerance=1e-13)
    assert np.allclose(x.cpu().numpy(), np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)


def test_cg_nonconvergence_raises(gpu):
    plan = g.compile(P.poisson(2, 8, 3))
    with pytest.raises(g.ConvergenceError):
        g.solve(plan, solver=g.CGSolver(), relative_tolerance=1e-14, maxiter=2)


def test_manufactured_sine_3d(gpu):
    plan = g.compile(P.poisson(3, 4, 3))
    state = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-10)
    assert g.l2_error(state, P.sine_exact) < 2e-3
