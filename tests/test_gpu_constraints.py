"""Mean-value rows, affine Dirichlet relations and the general reduced map on
the device against the oracle: full-layout apply with constraint rows,
reduced apply through the coefficient CSR map, the Jacobi diagonal of a
non-selection map, solves against the oracle's dense solve, and the
reference's goldens (test/matrix_free.jl:1033-1053)."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd import ops
from gricoamd.solve import reduced_diagonal_general

import problems as P

pytestmark = pytest.mark.gpu


def _neumann(dim, cells, degree, target=0.0):
    pr = P.poisson(dim, cells, degree, dirichlet=False, rhs=False)
    pr.add_cell_accumulator(pr.field, P.CosLoad(dim))
    pr.add_constraint(g.MeanValue(pr.field, target))
    return pr


def _with_relations(plan):
    """Relations among free dofs (one chained through another), set before the
    device plan is built: u_a - 0.5 u_b + 0.25 u_c = 0.3, u_d - 2 u_a = -1."""
    free = plan.solve_dofs
    a, b, c, d = (int(v) for v in free[[3, 8, 13, 21]])
    plan.affine_rows = [(a, np.array([b, c]), np.array([0.5, -0.25]), 0.3), (d, np.array([a]), np.array([2.0]), -1.0)]
    plan._compile_reduced_map()
    return plan


CASES = {
    "neumann_2d_p3": lambda: g.compile(_neumann(2, 5, 3)),
    "neumann_3d_p3_bricks": lambda: g.compile(_neumann(3, 4, 3, target=0.2)),
    "dirichlet_relations_2d": lambda: _with_relations(g.compile(P.poisson(2, 4, 2, kappa=lambda p: 1.0 + 0.5 * p[0] * p[1]))),
}


@pytest.mark.parametrize("name", list(CASES))
def test_apply_and_reduced_apply_match_oracle(gpu, name):
    plan = CASES[name]()
    assert not plan.selection_map
    x = P.rng_vector(plan.full_dof_count, 3)
    y = g.apply(plan, x)
    ref = pyoracle.apply(plan, x)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    xr = P.rng_vector(plan.reduced_dof_count, 4)
    yr = ops.reduced_apply(plan, xr)
    rr = pyoracle.reduced_apply(plan, xr)
    assert np.abs(yr - rr).max() <= 1e-12 * np.abs(rr).max()


@pytest.mark.parametrize("name", ["neumann_2d_p3", "dirichlet_relations_2d"])
def test_diagonal_and_solve_match_dense(gpu, name):
    plan = CASES[name]()
    A = pyoracle.reduced_matrix(plan)
    d = reduced_diagonal_general(plan)
    assert np.abs(d - np.diag(A)).max() <= 1e-12 * np.abs(np.diag(A)).max()
    b = plan.reduced_rhs()
    x_ref = np.linalg.solve(A, b)
    st = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-12)
    assert st.info.converged == 1
    u_ref = pyoracle.expand(plan, x_ref)
    assert np.abs(st.coefficients - u_ref).max() <= 1e-9 * np.abs(u_ref).max()
    for (p, idx, cf, rhs) in plan.affine_rows:
        assert st.coefficients[p] - cf @ st.coefficients[idx] == pytest.approx(rhs, abs=1e-10)
    for (p, idx, cf, rhs) in plan.mean_rows:
        assert cf @ st.coefficients[idx] == pytest.approx(rhs, abs=1e-10)


def test_neumann_solution_is_the_zero_mean_solution(gpu):
    """Pure Neumann Poisson with f = 2 pi^2 cos cos and mean 0: the exact
    solution is cos(pi x) cos(pi y); the device solution's L2 error matches the
    oracle dense solution's to 1% and is small."""
    plan = g.compile(_neumann(2, 6, 3))
    st = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-12)
    e_dev = g.l2_error(st, P.cos_exact)
    x_ref = np.linalg.solve(pyoracle.reduced_matrix(plan), plan.reduced_rhs())
    e_ref = g.l2_error(g.State(plan, pyoracle.expand(plan, x_ref)), P.cos_exact)
    assert e_dev == pytest.approx(e_ref, rel=1e-2)
    assert e_dev < 1e-3


def test_mean_value_golden(gpu):
    """solve(mean_problem) = [2.0] (test/matrix_free.jl:1047-1053)."""
    st = g.solve(P.mean_golden_problem())
    assert st.coefficients.tolist() == pytest.approx([2.0])


def test_dirichlet_one_cell_closed_form(gpu):
    """The Dirichlet golden's setting (test/matrix_free.jl:1033-1045: one p = 1
    cell on [0, 1], u = 2 at x = 0) with a mass operator and unit load instead
    of the reference's local-identity kernel, which a lowered point action
    cannot express: M = [[1/3, 1/6], [1/6, 1/3]], b = [1/2, 1/2] gives u1 = 1/2."""
    domain = g.Domain((0.0,), (1.0,), (1,))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(1)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())

    class One:
        def cell_rhs_accumulate(self, q, test_component):
            return 1.0 + 0.0 * q.point[0]

    pr.add_cell_accumulator(u, One())
    pr.add_constraint(g.Dirichlet(u, g.BoundaryFace(0, g.LOWER), 2.0))
    st = g.solve(pr, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-14)
    assert st.coefficients.tolist() == pytest.approx([2.0, 0.5], rel=1e-12)


def test_gmg_on_mean_constrained_problem(gpu):
    """hp-multigrid where every level carries the mean row (src/multigrid.jl:
    596-612): levels with a non-selection map take their Jacobi diagonals from
    the host. The Chebyshev V-cycle equals the oracle's dense V-cycle and
    GMG-CG reaches the dense solution."""
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid
    pr = P.poisson(2, None, 3, refine_depth=3, dirichlet=False, rhs=False)
    pr.add_cell_accumulator(pr.field, P.CosLoad(2))
    pr.add_constraint(g.MeanValue(pr.field, 0.0))
    pol = GeometricMultigridPreconditioner(smoother="chebyshev")
    h = compile_geometric_multigrid(pr, pol)
    assert all(not p.selection_map for p in h.plans if p.reduced_dof_count)
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 11)
    z = h.device_preconditioner().apply(r).cpu().numpy()
    ref = pyoracle.vcycle(mats, Ps, r, smoother="chebyshev", pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, lambdas=h.level_lambda_max(),
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()
    st = g.solve(pr, solver=g.CGSolver(preconditioner=pol), relative_tolerance=1e-12)
    plan = h.plans[-1]
    x_ref = np.linalg.solve(mats[-1], plan.reduced_rhs())
    u_ref = pyoracle.expand(plan, x_ref)
    assert np.abs(st.coefficients - u_ref).max() <= 1e-9 * np.abs(u_ref).max()
