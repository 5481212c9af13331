"""Device residual, tangent and Newton (gmf_interpolate_channels,
gmf_rhs_physical, gmf_plan_set_point_data, device CG) against the oracle's
residual! / tangent_apply! restatement and its dense Newton."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle as O
from gricoamd import nonlinear as NL

import problems as P

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = {
    "2d_p3": lambda: P.bratu_problem((4, 4), 3),
    "2d_hanging_p2": lambda: P.bratu_problem((3, 3), 2, refine=((4, 0), (-1, 1))),
    "3d_p2": lambda: P.bratu_problem((3, 3, 3), 2, lam=0.7, amplitude=0.3),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, gpu):
    print("case", request.param, flush=True)
    return request.param, NL.ResidualPlan(CASES[request.param]())


def _state(rplan, seed):
    rng = np.random.default_rng(seed)
    x = 0.4 * rng.standard_normal(rplan.full_dof_count)
    x[rplan.plan.fixed_dofs] = rplan.plan.fixed_values
    return x


def test_residual(case):
    name, rplan = case
    x = _state(rplan, 1)
    assert rel(NL.residual(rplan, x), O.residual(rplan, x)) <= TOL, name


def test_tangent_apply(case):
    name, rplan = case
    x = _state(rplan, 2)
    dx = np.random.default_rng(3).standard_normal(rplan.full_dof_count)
    assert rel(NL.tangent_apply(rplan, x, dx), O.tangent_apply(rplan, x, dx)) <= TOL, name
    # a second state re-lowers the same plan
    x2 = _state(rplan, 5)
    assert rel(NL.tangent_apply(rplan, x2, dx), O.tangent_apply(rplan, x2, dx)) <= TOL, name


def test_newton_matches_oracle_newton(case):
    name, rplan = case
    state = NL.newton_solve(rplan, relative_tolerance=1e-10, absolute_tolerance=1e-10,
                            linear_relative_tolerance=1e-12, maxiter=10)
    if rplan.reduced_dof_count <= 800:
        u, norms = O.newton_dense(rplan)
        assert np.abs(state.coefficients - u).max() <= 1e-8 * max(1.0, np.abs(u).max()), name
    h = state.info["residual_norms"]
    assert h[-1] <= 1e-10 * h[0]
    assert state.info["iterations"] <= 6


def test_bratu_example(gpu):
    """examples/nonlinear_bratu/driver.jl defaults (4 x 4 roots, p = 3,
    lambda = 1, A = 0.2) and the reference test configuration
    (test/examples.jl:30-35: L2 error <= 2e-3)."""
    rplan = NL.ResidualPlan(P.bratu_problem((2, 2), 2, lam=0.5, amplitude=0.1))
    st = NL.newton_solve(rplan, maxiter=10, relative_tolerance=1e-10, absolute_tolerance=1e-10,
                         linear_relative_tolerance=1e-10)
    assert g.l2_error(st, P.bratu_exact(0.1)) <= 2e-3
    rplan = NL.ResidualPlan(P.bratu_problem((4, 4), 3))
    st = NL.newton_solve(rplan, maxiter=10, relative_tolerance=1e-10, absolute_tolerance=1e-10,
                         linear_relative_tolerance=1e-10)
    coarse = g.l2_error(st, P.bratu_exact(0.2))
    rplan = NL.ResidualPlan(P.bratu_problem((8, 8), 3))
    st = NL.newton_solve(rplan, maxiter=10, relative_tolerance=1e-10, absolute_tolerance=1e-10,
                         linear_relative_tolerance=1e-10)
    fine = g.l2_error(st, P.bratu_exact(0.2))
    assert fine < coarse / 10  # O(h^4) at p = 3: a factor ~16 per halving


def test_newton_rejects_preconditioner(gpu):
    rplan = NL.ResidualPlan(P.bratu_problem((2, 2), 2))
    with pytest.raises(ValueError, match="preconditioning"):
        NL.newton_solve(rplan, preconditioner=g.JacobiPreconditioner())


def test_entry_point_errors(gpu):
    """Invalid arguments of the residual-support entry points come back as
    GMF_EINVAL / GMF_EUNSUPPORTED (the reference's ArgumentError class)."""
    import ctypes

    import torch
    from gricoamd import _devabi, _native, ops

    lib = _native.device()
    rplan = NL.ResidualPlan(P.bratu_problem((2, 2), 2))
    dev = ops._ws(rplan.plan)
    u = torch.zeros(rplan.full_dof_count, dtype=torch.float64, device="cuda")
    npts = sum(b.cell_count * b.n_points for b in rplan.plan.buckets)
    out = torch.empty(npts * 3, dtype=torch.float64, device="cuda")
    assert lib.gmf_interpolate_channels(dev.ws, u.data_ptr(), 2, out.data_ptr()) == -1  # GMF_EINVAL
    data = np.ascontiguousarray(rplan.plan.buckets[0].point_data)
    ptr = data.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.gmf_plan_set_point_data(dev.handle, 7, ptr) == -1  # GMF_EINVAL
    bad = data.copy()
    bad.reshape(-1)[3] = np.nan
    assert lib.gmf_plan_set_point_data(dev.handle, 0, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) \
        == -1
    assert lib.gmf_plan_set_point_data(dev.handle, 0, ptr) == 0
    # point-cloud items are outside the state interpolation
    from test_faces_host import _BoundaryMass, _dg
    f = _dg(1, 2, 1)
    pr = g.AffineProblem(f, g.SPD())
    pr.add_boundary_accumulator(g.BoundaryFace(0, g.UPPER), f, f, _BoundaryMass())
    plan = g.compile(pr)
    fd = ops._ws(plan)
    v = torch.zeros(plan.full_dof_count, dtype=torch.float64, device="cuda")
    assert lib.gmf_interpolate_channels(fd.ws, v.data_ptr(), 1, out.data_ptr()) == -8  # GMF_EUNSUPPORTED
