"""Log-normal coefficient workload (p = 3, seeded Gaussian random field) on the
device against the oracle: full-layout apply, the Chebyshev V-cycle of its
hp-multigrid hierarchy, and GMG-CG solutions against the oracle's dense solve
and its CG solve."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd import ops
from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid

import problems as P

pytestmark = pytest.mark.gpu


def _problem(depth):
    return P.lognormal_poisson(depth=depth, degree=3, seed=2024)


@pytest.mark.parametrize("depth", [2, 3])
def test_apply_matches_oracle(gpu, depth):
    plan = g.compile(_problem(depth))
    x = P.rng_vector(plan.full_dof_count, 5)
    y = g.apply(plan, x)
    ref = pyoracle.apply(plan, x)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    xr = P.rng_vector(plan.reduced_dof_count, 6)
    yr = ops.reduced_apply(plan, xr)
    assert np.abs(yr - pyoracle.reduced_apply(plan, xr)).max() <= 1e-12 * np.abs(yr).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 5e-5)])
def test_vcycle_matches_oracle(gpu, precision, tol):
    problem = _problem(2)
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", precision=precision)
    h = compile_geometric_multigrid(problem, pol)
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 11)
    z = h.device_preconditioner().apply(r).cpu().numpy()
    ref = pyoracle.vcycle(mats, Ps, r, smoother="chebyshev", pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, lambdas=h.level_lambda_max(),
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    assert np.abs(z - ref).max() <= tol * np.abs(ref).max()


def _reduced(plan, full):
    return np.asarray(full)[plan.solve_dofs]


def _full(plan, x_red):
    out = np.zeros(plan.full_dof_count)
    out[plan.solve_dofs] = x_red
    out[plan.fixed_dofs] = plan.fixed_values
    return out


def _gmg_solve(problem):
    return g.solve(problem, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner(smoother="chebyshev")),
                   relative_tolerance=1e-12)


def test_gmg_cg_matches_dense_solve(gpu):
    problem = _problem(2)
    plan = g.compile(problem)
    st = _gmg_solve(problem)
    x_ref = np.linalg.solve(pyoracle.reduced_matrix(plan), plan.reduced_rhs())
    assert st.info.converged == 1
    assert np.abs(_reduced(plan, st.coefficients) - x_ref).max() <= 1e-8 * np.abs(x_ref).max()


def test_gmg_cg_matches_oracle_cg_8cubed(gpu):
    """8^3 cells (12,167 reduced dofs): device GMG-CG against the oracle's
    Jacobi-CG on the same reduced system, both to 1e-12; the solutions' L2
    norms agree to 1% (in fact to the solver tolerance), and GMG-CG needs far
    fewer iterations."""
    problem = _problem(3)
    plan = g.compile(problem)
    inv = 1.0 / _reduced(plan, pyoracle.diagonal_physical(plan))
    x_ref, it_ref = pyoracle.cg(plan, plan.reduced_rhs(), inv_diag=inv, rtol=1e-12, maxiter=20000)
    st = _gmg_solve(problem)
    assert np.abs(_reduced(plan, st.coefficients) - x_ref).max() <= 1e-7 * np.abs(x_ref).max()
    zero = lambda p: 0.0 * p[0]  # noqa: E731
    n_dev = g.l2_error(st, zero)
    n_ref = g.l2_error(g.State(plan, _full(plan, x_ref)), zero)
    assert n_dev == pytest.approx(n_ref, rel=1e-2)
    assert st.info.iterations < it_ref / 3
