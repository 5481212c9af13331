"""Device boundary-face and interface items (one- and two-sided point-cloud
kernels) against the oracle's restatement of face_apply! / interface_apply!
(src/accumulators.jl:1133-1161, 1202-1214, 1576-1664), plus a device SIPG
solve that must reproduce a quadratic exactly."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle

import problems as P
from test_faces_host import _BoundaryMass, _dg

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cases():
    out = {}
    out["2d_sipg"] = P.sipg_poisson(dim=2, cells=4, degree=2)
    out["2d_sipg_hanging"] = P.sipg_poisson(dim=2, cells=3, degree=3, refine=((4, 0), (-1, 1), (0, 1)))
    out["2d_sipg_p_mismatch"] = P.sipg_poisson(dim=2, cells=3, degrees=lambda d, leaf: 2 + leaf % 3)
    out["3d_sipg"] = P.sipg_poisson(dim=3, cells=3, degree=3, refine=((13, 0),))
    out["3d_sipg_p7_small_chunks"] = P.sipg_poisson(dim=3, cells=2, degree=7)
    u = _dg(2, 3, 2, ((0, 0),))
    pr = g.AffineProblem(u, g.NonsymmetricOperator())
    pr.add_cell_accumulator(u, u, P.MassDiffusion(0.5, 1.0, b=(0.3, -0.2)))
    pr.add_boundary_accumulator(g.BoundaryFace(0, g.UPPER), u, u, P.OneSidedAffine(1.1, 0.2, 0.3, 0.4))
    pr.add_boundary_accumulator(g.BoundaryFace(1, g.LOWER), u, u, P.OneSidedAffine(0.9, -0.15, 0.25, 0.2))
    pr.add_interface_accumulator(u, u, P.InterfaceAffine(1.4, -0.3, 0.2))
    out["2d_nonsymmetric_forms"] = pr
    pr = P.poisson(2, 4, 3, rhs=False)
    pr.add_interface_accumulator(pr.field, pr.field, P.GradientJumpPenalty(0.05))
    out["2d_cg_gradient_jump_dirichlet"] = pr
    return out


CASES = None
NAMES = ["2d_sipg", "2d_sipg_hanging", "2d_sipg_p_mismatch", "3d_sipg", "3d_sipg_p7_small_chunks",
         "2d_nonsymmetric_forms", "2d_cg_gradient_jump_dirichlet"]


@pytest.fixture(scope="module", params=NAMES)
def case(request, gpu):
    global CASES
    if CASES is None:
        CASES = {k: g.compile(v) for k, v in cases().items()}
    print("case", request.param, flush=True)
    return request.param, CASES[request.param]


def test_apply(case):
    name, plan = case
    assert plan.interface_buckets
    x = P.rng_vector(plan.full_dof_count, 3)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) <= TOL, name
    assert rel(g.apply_physical(plan, x), pyoracle.apply(plan, x, physical=True)) <= TOL, name


def test_reduced_apply(case):
    name, plan = case
    xr = P.rng_vector(plan.reduced_dof_count, 4)
    assert rel(g.reduced_apply(plan, xr), pyoracle.reduced_apply(plan, xr)) <= TOL, name


def test_diagonal(case):
    name, plan = case
    if name.startswith("2d_cg"):
        # trace dofs shared by both sides of an interface: the local diagonal is
        # not the operator diagonal, and the exact reduced diagonal refuses
        with pytest.raises(Exception, match="point-cloud"):
            g.reduced_diagonal(plan)
        return
    assert rel(g.diagonal_physical(plan), pyoracle.diagonal_physical(plan)) <= TOL, name
    if plan.reduced_dof_count <= 700:
        assert rel(g.reduced_diagonal(plan), pyoracle.reduced_matrix(plan).diagonal()) <= TOL, name


def test_boundary_mass_golden(gpu):
    u = _dg(1, 1, 1)
    pr = g.AffineProblem(u, g.SPD())
    pr.add_boundary_accumulator(g.BoundaryFace(0, g.UPPER), u, u, _BoundaryMass())
    plan = g.compile(pr)
    assert np.allclose(g.apply(plan, np.array([0.25, -0.5])), [0.0, -1.5], atol=1e-15)


@pytest.mark.parametrize("dim,cells,refine", [(2, 6, ((7, 0), (-1, 1))), (3, 3, ((0, 0),))])
def test_sipg_solve_reproduces_quadratic(gpu, dim, cells, refine):
    """Device CG + Jacobi on the SIPG system: the quadratic exact solution lies
    in the DG space, so the solve reproduces it to the solver tolerance."""
    problem = P.sipg_poisson(dim=dim, cells=cells, degree=2, refine=refine)
    plan = g.compile(problem)
    state = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-13)
    assert state.info.converged
    assert g.relative_l2_error(state, P.quadratic_exact) < 1e-9


def _describe(plan):
    from gricoamd import ops
    return ops._ws(plan).describe()


@pytest.mark.parametrize("name", ["2d_sipg", "3d_sipg", "3d_sipg_p7_small_chunks", "2d_sipg_p_mismatch",
                                  "2d_sipg_hanging", "2d_nonsymmetric_forms"])
def test_faces_take_the_tensor_path(gpu, name):
    """Every interface batch, hanging ones included (batched by sub-face
    position), and every boundary-face batch runs the sum-factorized kernel."""
    global CASES
    if CASES is None:
        CASES = {k: g.compile(v) for k, v in cases().items()}
    lines = [l for l in _describe(CASES[name]).splitlines() if "interface points=" in l or "cloud points=" in l]
    assert lines and all("face-tensor" in l for l in lines), "\n".join(lines)


@pytest.mark.parametrize("name", ["2d_sipg_hanging", "3d_sipg", "2d_nonsymmetric_forms"])
def test_tensor_path_matches_point_cloud_kernel(gpu, monkeypatch, name):
    """The same plan with GMF_FACE_TENSOR=0 (every face batch on the
    point-cloud kernel) gives the same action."""
    fast = g.compile(cases()[name])
    monkeypatch.setenv("GMF_FACE_TENSOR", "0")
    slow = g.compile(cases()[name])
    x = P.rng_vector(fast.full_dof_count, 9)
    ys = g.apply_physical(slow, x)
    monkeypatch.delenv("GMF_FACE_TENSOR")
    assert "face-tensor" not in _describe(slow)
    yf = g.apply_physical(fast, x)
    assert "face-tensor" in _describe(fast)
    assert rel(yf, ys) <= 1e-13, name


@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
@pytest.mark.parametrize("dim,depth,degree", [(2, 2, 2), (3, 1, 2)])
def test_sipg_vcycle_matches_oracle(gpu, smoother, dim, depth, degree):
    """hp-multigrid on a DG space with interface and boundary-face operators
    rediscretized on every level (src/multigrid.jl:596-612) against the
    oracle's dense V-cycle."""
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid
    problem = P.sipg_poisson(dim=dim, degree=degree, refine_depth=depth, rhs=False)
    pol = GeometricMultigridPreconditioner(smoother=smoother)
    h = compile_geometric_multigrid(problem, pol)
    assert h.level_count >= 3
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    lambdas = h.level_lambda_max() if smoother == "chebyshev" else None
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 11)
    z = h.device_preconditioner().apply(r).cpu().numpy()
    ref = pyoracle.vcycle(mats, Ps, r, smoother=smoother, pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, omega=pol.smoother_damping, lambdas=lambdas,
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()


def test_sipg_gmg_cg_reproduces_quadratic(gpu):
    """Chebyshev-Jacobi smoothing: the interior penalty puts lambda_max(D^-1 A)
    above 2 / 0.8, where the default damped Jacobi smoother (the reference's
    :auto choice for SPD operators) diverges and the V-cycle stops being a
    positive definite preconditioner (the oracle's dense PCG stalls at 0.38
    relative residual after 300 iterations on this problem)."""
    from gricoamd.multigrid import GeometricMultigridPreconditioner
    problem = P.sipg_poisson(dim=2, degree=2, refine_depth=4)
    gmg = g.solve(problem, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner(smoother="chebyshev")),
                  relative_tolerance=1e-12)
    jac = g.solve(g.compile(problem), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-12)
    assert g.relative_l2_error(gmg, P.quadratic_exact) < 1e-9
    assert gmg.info.iterations < jac.info.iterations / 3


def _rhs_cases():
    from test_faces_host import _PlusLoad
    out = {
        "2d_sipg": P.sipg_poisson(dim=2, cells=4, degree=2),
        "2d_sipg_hanging": P.sipg_poisson(dim=2, cells=3, degree=3, refine=((4, 0), (-1, 1), (0, 1))),
        "3d_sipg_hanging": P.sipg_poisson(dim=3, cells=3, degree=2, refine=((13, 0),)),
    }
    pr = P.sipg_poisson(dim=2, cells=3, degree=2, refine=((4, 0),))
    pr.add_interface_accumulator(pr.field, _PlusLoad(0.7))
    out["2d_interface_loads"] = pr
    return out


@pytest.mark.parametrize("name", ["2d_sipg", "2d_sipg_hanging", "3d_sipg_hanging", "2d_interface_loads"])
def test_device_face_rhs_matches_host(gpu, name):
    """gmf_rhs_physical + gmf_rhs_faces against the host projection of the same
    loads (itself pinned by the SIPG exactness and interface-load checks)."""
    from gricoamd import ops
    plan = g.compile(_rhs_cases()[name])
    assert plan.device_rhs
    dev = ops.rhs_physical(plan)
    host = plan.rhs_physical()
    assert rel(dev, host) <= 1e-12, name


@pytest.mark.parametrize("name", ["2d_sipg_hanging", "3d_sipg_hanging", "2d_interface_loads"])
def test_device_face_rhs_on_point_cloud_kernel(gpu, monkeypatch, name):
    """With GMF_FACE_TENSOR=0 every face and interface batch (one- and
    two-sided) runs the point-cloud kernel, whose load projection
    (cell_rhs_cloud) then gives the host projection too."""
    from gricoamd import ops
    monkeypatch.setenv("GMF_FACE_TENSOR", "0")
    plan = g.compile(_rhs_cases()[name])
    assert plan.device_rhs
    dev = ops.rhs_physical(plan)
    assert "face-tensor" not in _describe(plan)
    assert rel(dev, plan.rhs_physical()) <= 1e-12, name


def test_device_rhs_falls_back_for_load_only_faces(gpu):
    """A boundary side with a load but no operator is not in the device plan:
    the host path takes over, and solve() still reproduces the solution."""
    u = _dg(1, 3, 2)
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.MassDiffusion(1.0, 1.0))

    class _Load:
        def boundary_rhs_accumulate(self, q, test_component):
            return 2.0 + 0.0 * q.point[0]

    pr.add_boundary_accumulator(g.BoundaryFace(0, g.UPPER), u, _Load())
    plan = g.compile(pr)
    assert not plan.device_rhs
    st = g.solve(plan, solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()), relative_tolerance=1e-12)
    A = pyoracle.reduced_matrix(plan)
    x = np.linalg.solve(A, plan.reduced_rhs())
    assert np.abs(st.coefficients[plan.solve_dofs] - x).max() <= 1e-9 * np.abs(x).max()
