"""Right-hand sides of point-cloud batches on the device (cut-cell loads,
embedded-surface loads): gmf_rhs_physical + gmf_rhs_faces against the host
projection of the same point loads (plan.rhs_physical, itself pinned by the
finite-cell solves against 1 - r^2 and the oracle's dense assembly)."""
import numpy as np
import pytest

import gricoamd as g
from gricoamd import ops

import problems as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def _with_surface_data(pr, degree):
    field = pr.cell_operators[0][0]
    pr.add_surface_accumulator(field, P.NitscheData(10.0 * (degree + 1) ** 2))
    return pr


@pytest.mark.parametrize("case", ["explicit", "physical"])
def test_device_cloud_rhs_matches_host(gpu, case):
    if case == "explicit":
        pr = P.fcm_ball(n=4, degree=2, depth=2, n_theta=16, n_phi=32, alpha=1e-3)
    else:
        pr = P.fcm_ball_physical(n=4, degree=2, depth=2, n_theta=16, n_phi=32)
    _with_surface_data(pr, 2)
    plan = g.compile(pr)
    assert plan.cloud_buckets and plan.surface_buckets and plan.device_rhs
    dev = ops.rhs_physical(plan)
    host = plan.rhs_physical()
    assert np.abs(host).max() > 0
    assert rel(dev, host) <= 1e-12, case
    # reduced right-hand side (with the Dirichlet-free ball: P^T b)
    assert rel(ops.reduced_rhs(plan), plan.reduced_rhs()) <= 1e-12


def test_device_cloud_rhs_split_work_units(gpu, monkeypatch):
    # 256-point work units split the heavier cut cells over several
    # workgroups; their partial projections meet in the atomic scatter.
    monkeypatch.setenv("GMF_CLOUD_UNIT_POINTS", "256")
    pr = P.fcm_ball(n=4, degree=2, depth=3, n_theta=16, n_phi=32, alpha=1e-3)
    _with_surface_data(pr, 2)
    plan = g.compile(pr)
    assert max(int(np.diff(bk.point_offsets).max()) for bk in plan.cloud_buckets) > 256
    dev = ops.rhs_physical(plan)
    assert rel(dev, plan.rhs_physical()) <= 1e-12


def test_solve_uses_device_cloud_rhs(gpu):
    # solve() assembles b on the device for a finite-cell problem (extension
    # alpha = 0.1 keeps Jacobi-CG short); the solution equals the one from the
    # host right-hand side and solves the oracle's dense system.
    from gricoamd.solve import solve_reduced
    import pyoracle
    pr = P.fcm_ball(n=4, degree=2, depth=2, n_theta=16, n_phi=32, alpha=0.1)
    _with_surface_data(pr, 2)
    plan = g.compile(pr)
    assert plan.device_rhs
    solver = g.CGSolver(preconditioner=g.JacobiPreconditioner())
    st = g.solve(plan, solver=solver, relative_tolerance=1e-10, maxiter=20000)
    got = st.coefficients[plan.solve_dofs]
    b = plan.reduced_rhs()
    xh, _ = solve_reduced(plan, solver, b, 1e-10, 0.0, 20000, None)
    xh = xh.cpu().numpy() if hasattr(xh, "cpu") else np.asarray(xh)
    # both stop at ||r|| <= 1e-10 ||b||: iterates agree to the CG tolerance
    # times the condition number (about 1e2 here)
    assert np.abs(got - xh).max() <= 1e-6 * np.abs(xh).max()
    A = pyoracle.reduced_matrix(plan)
    assert np.linalg.norm(A @ got - b) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("case", ["explicit", "physical", "split"])
def test_device_l2_on_cut_cells(gpu, case, monkeypatch):
    # gmf_l2_error over tensor and cut cells against the host verification
    # sums (src/verification.jl:201-230) on a random coefficient vector.
    from gricoamd.verification import device_relative_l2_error, relative_l2_error, device_l2_error, l2_error
    if case == "split":
        monkeypatch.setenv("GMF_CLOUD_UNIT_POINTS", "256")
        pr = P.fcm_ball(n=4, degree=2, depth=3, n_theta=16, n_phi=32, alpha=1e-3)
    elif case == "explicit":
        pr = P.fcm_ball(n=4, degree=2, depth=2, n_theta=16, n_phi=32, alpha=1e-3)
    else:
        pr = P.fcm_ball_physical(n=4, degree=2, depth=2, n_theta=16, n_phi=32)
    plan = g.compile(pr)
    st = g.State(plan, np.random.default_rng(4).standard_normal(plan.full_dof_count))
    assert abs(device_relative_l2_error(st, P.ball_exact) - relative_l2_error(st, P.ball_exact)) <= \
        1e-12 * relative_l2_error(st, P.ball_exact)
    assert abs(device_l2_error(st, P.ball_exact) - l2_error(st, P.ball_exact)) <= 1e-12 * l2_error(st, P.ball_exact)
