"""Device hp-multigrid V-cycle against the oracle's dense V-cycle, and GMG-CG
solves against Jacobi-CG and dense solves (test/multigrid.jl:91-101, 129-149)."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid

import problems as P

pytestmark = pytest.mark.gpu

CASES = {
    "p_2d": lambda: P.poisson(2, 3, 4, rhs=False),
    "h_2d": lambda: P.poisson(2, None, 1, refine_depth=2, rhs=False),
    "hp_3d": lambda: P.poisson(3, None, 2, refine_depth=1, kappa=P.kappa_c2, rhs=False),
    "hp_3d_p4": lambda: P.poisson(3, None, 4, refine_depth=1, kappa=P.kappa_c2, rhs=False),
    "corner": lambda: P.corner_refined(2, base=2, levels=1, degree=2),
    "dg": lambda: P.poisson(2, None, 2, refine_depth=1, continuity="dg", dirichlet=False, rhs=False,
                            operator=P.MassDiffusion(1.0, 1.0)),
}


def _device_vcycle(h, r):
    M = h.device_preconditioner()
    return M.apply(r).cpu().numpy()


@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
@pytest.mark.parametrize("name", list(CASES))
def test_vcycle_matches_oracle(gpu, name, smoother):
    problem = CASES[name]()
    pol = GeometricMultigridPreconditioner(smoother=smoother)
    h = compile_geometric_multigrid(problem, pol)
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    lambdas = h.level_lambda_max() if smoother == "chebyshev" else None
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 11)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r, smoother=smoother, pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, omega=pol.smoother_damping, lambdas=lambdas,
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()
    # a V-cycle with equal pre/post smoothing is symmetric: r1.M r2 == r2.M r1
    r2 = P.rng_vector(len(r), 12)
    z2 = _device_vcycle(h, r2)
    assert float(r2 @ z) == pytest.approx(float(r @ z2), rel=1e-10)


@pytest.mark.parametrize("name", ["p_2d", "hp_3d", "hp_3d_p4"])
def test_chebyshev_bounds(gpu, name):
    """1.2 x the Lanczos estimate brackets lambda_max(D^-1 A) of every smoothed level."""
    h = compile_geometric_multigrid(CASES[name](), GeometricMultigridPreconditioner(smoother="chebyshev"))
    lam = h.level_lambda_max()
    for l in range(1, h.level_count):
        A = pyoracle.reduced_matrix(h.plans[l])
        if A.size == 0:
            continue
        d = np.diag(A)
        true = np.linalg.eigvalsh(A / np.sqrt(np.outer(d, d))).max()
        assert 0.9 * true <= lam[l] / 1.2 <= true * (1 + 1e-8)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_chebyshev_equals_unfused(gpu, monkeypatch, precision):
    """Chebyshev steps fused into the brick levels' applies (the default) give
    the V-cycle of the separate apply + vector kernels: the fused store performs
    the same operations on the same values. Not bit for bit, because a brick's
    eight cells add into its shared lattice points with LDS atomics in no fixed
    order, so two applies of one vector may differ in the last bit; a wrong
    coefficient, vector or sign changes the result at O(1). The p = 4, 2, 1
    levels of 2^3 cells run on bricks (forced in conftest); the h-levels of
    4 and 2 cells do not and take the unfused path in both runs."""
    from gricoamd import ops
    problem = CASES["hp_3d_p4"]()
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", precision=precision)
    h = compile_geometric_multigrid(problem, pol)
    assert " bricks=1 " in ops._ws(h.plans[-1]).describe()
    rs = [P.rng_vector(h.plans[-1].reduced_dof_count, seed) for seed in (3, 4)]
    fused = [_device_vcycle(h, r) for r in rs]
    monkeypatch.setenv("GMF_MG_FUSED", "0")
    plain = h.with_policy(pol)
    tol = 1e-12 if precision == "fp64" else 1e-5  # last-bit differences of FP64 / FP32 levels
    for r, z in zip(rs, fused):
        assert np.abs(_device_vcycle(plain, r) - z).max() <= tol * np.abs(z).max()


@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
def test_merged_h_steps_vcycle_matches_oracle(gpu, smoother):
    """h_steps_per_level = 3 (one h-level per full 2x2x2 coarsening) on a
    4^3-cell p = 2 mesh: the device V-cycle over composed transfers against
    the oracle's dense V-cycle with the same level matrices and transfers."""
    problem = P.poisson(3, None, 2, refine_depth=2, kappa=P.kappa_c2, rhs=False)
    pol = GeometricMultigridPreconditioner(smoother=smoother, h_steps_per_level=3)
    h = compile_geometric_multigrid(problem, pol)
    assert [len(sp.leaves) for sp in h.spaces] == [1, 8, 64, 64]
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    lambdas = h.level_lambda_max() if smoother == "chebyshev" else None
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 5)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r, smoother=smoother, pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, omega=pol.smoother_damping, lambdas=lambdas,
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()


def test_gmg_cg_matches_jacobi_cg_mass(gpu):
    """DG p=3 mass on two cells (test/multigrid.jl:91-101)."""
    d = g.Domain((0.0,), (1.0,), (2,))
    sp = g.HpSpace(d, g.SpaceOptions(degree=g.UniformDegree(3), continuity="dg"))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())

    class One:
        def cell_rhs_accumulate(self, q, test_component):
            return 1.0 + 0.0 * q.point[0]

    pr.add_cell_accumulator(u, One())
    gmg = g.solve(pr, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner()))
    jac = g.solve(g.compile(pr), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-12)
    assert np.allclose(gmg.coefficients, jac.coefficients, atol=1e-8)
    # the L2 projection of 1 is 1: coefficient 1 on the vertex modes, 0 elsewhere
    assert g.l2_error(gmg, lambda p: 1.0 + 0.0 * p[0]) < 1e-7


@pytest.mark.parametrize("smoother,degree", [("jacobi", 2), ("chebyshev", 4)])
def test_gmg_cg_poisson_3d(gpu, smoother, degree):
    """GMG-CG reaches the Jacobi-CG solution. The reference's damped Jacobi
    smoother (omega = 0.8) is unstable where lambda_max(D^-1 A) > 2 / omega
    (3D modal bases reach 4-5), so only the Chebyshev smoother, whose interval
    comes from the Lanczos estimate, is held to an iteration bound."""
    pr = P.poisson(3, None, degree, refine_depth=2, kappa=P.kappa_c2)
    gmg = g.solve(pr, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner(smoother=smoother)),
                  relative_tolerance=1e-10)
    jac = g.solve(g.compile(pr), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-12)
    scale = np.abs(jac.coefficients).max()
    assert np.abs(gmg.coefficients - jac.coefficients).max() <= 1e-8 * scale
    assert gmg.info.converged == 1
    if smoother == "chebyshev":
        assert gmg.info.iterations < jac.info.iterations / 3


def test_default_solver_uses_gmg(gpu):
    """solve(problem) defaults to GMG-CG on SPD problems with a coarsening level
    (src/multigrid.jl:199-266)."""

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            x, y = q.point
            return 1.0 + x + np.exp(y) * np.cos(3.0 * x)

    pr = P.poisson(2, None, 2, refine_depth=3, rhs=False, kappa=lambda p: 1.0 + 0.5 * p[0])
    pr.add_cell_accumulator(pr.field, Load())
    st = g.solve(pr, relative_tolerance=1e-10)
    ref = g.solve(g.compile(pr), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-12)
    assert st.info.iterations < ref.info.iterations
    assert np.abs(st.coefficients - ref.coefficients).max() <= 1e-8 * np.abs(ref.coefficients).max()


def test_gmg_requires_problem(gpu):
    pr = P.poisson(2, None, 2, refine_depth=2)
    with pytest.raises(ValueError):
        g.solve(g.compile(pr), solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner()))
    with pytest.raises(ValueError):
        g.solve(pr, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner(pre_smoothing_steps=1,
                                                                                      post_smoothing_steps=2)))


def test_krylov_coarse_solve(gpu):
    """22x22 p=2 mass: the coarse level (529 dofs) is solved by Jacobi-CG."""
    d = g.Domain((0.0, 0.0), (1.0, 1.0), (22, 22))
    sp = g.HpSpace(d, g.SpaceOptions(degree=g.UniformDegree(2)))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())
    pr.add_cell_accumulator(u, P.SineLoad(2))
    h = compile_geometric_multigrid(pr)
    assert h.coarse_solver_kind() == "krylov"
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 3)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r)
    # the coarse CG stops at 1e-12 relative: agreement to that order
    assert np.abs(z - ref).max() <= 1e-9 * np.abs(ref).max()
    st = g.solve(pr, solver=g.CGSolver(preconditioner=GeometricMultigridPreconditioner()))
    assert st.info.converged == 1


def test_graph_replays_match_eager(gpu):
    """The captured V-cycle replayed on several inputs, with eager device work
    between the replays, agrees with the eager V-cycle every time."""
    make = lambda: P.poisson(3, None, 4, refine_depth=1, kappa=P.kappa_c2, rhs=False)
    hg = compile_geometric_multigrid(make(), GeometricMultigridPreconditioner(smoother="chebyshev", graph=True))
    he = compile_geometric_multigrid(make(), GeometricMultigridPreconditioner(smoother="chebyshev", graph=False))
    n = hg.plans[-1].reduced_dof_count
    for seed in range(4):
        r = P.rng_vector(n, 100 + seed)
        zg = _device_vcycle(hg, r)
        g.apply(hg.plans[-1], P.rng_vector(hg.plans[-1].full_dof_count, seed))  # eager work in between
        ze = _device_vcycle(he, r)
        gpu.cuda.synchronize()
        assert np.abs(zg - ze).max() <= 1e-13 * np.abs(ze).max()


# ---- FP32 V-cycle inside the FP64 Krylov method ---------------------------------

@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
@pytest.mark.parametrize("name", list(CASES))
def test_fp32_vcycle_matches_oracle(gpu, name, smoother):
    """The FP32 V-cycle (FP32 level vectors, smoothers, transfers and level
    operators; FP64 coarse solve) against the oracle's FP64 V-cycle with the
    same eigenvalue bounds. Tolerance: 5e-5 of the largest output entry, a
    few hundred FP32 roundings (unit 6e-8) through the smoothing steps of all
    levels; the error must also exceed 1e-12 of it, so that the FP32 path is
    the one that ran."""
    problem = CASES[name]()
    pol = GeometricMultigridPreconditioner(smoother=smoother, precision="fp32")
    h = compile_geometric_multigrid(problem, pol)
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    lambdas = h.level_lambda_max() if smoother == "chebyshev" else None
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 11)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r, smoother=smoother, pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, omega=pol.smoother_damping, lambdas=lambdas,
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    err = np.abs(z - ref).max()
    scale = np.abs(ref).max()
    assert err <= 5e-5 * scale
    assert err > 1e-12 * scale


@pytest.mark.parametrize("name", ["hp_3d", "hp_3d_p4"])
def test_fp32_levels_use_fp32_kernels(gpu, name):
    """Uniform 3D levels run FP32 cell kernels: the FP32 level operator agrees
    with the FP64 one to FP32 accuracy, not to FP64 accuracy."""
    h = compile_geometric_multigrid(CASES[name](), GeometricMultigridPreconditioner(smoother="chebyshev",
                                                                                    precision="fp32"))
    h.device_preconditioner()
    plan = h.plans[-1]
    x = P.rng_vector(plan.reduced_dof_count, 5)
    y64 = g.reduced_apply(plan, x)
    y32 = g.ops.reduced_apply_f32(plan, x)
    scale = np.abs(y64).max()
    assert np.abs(y32 - y64).max() <= 2e-6 * scale
    assert np.abs(y32 - y64).max() > 1e-12 * scale


@pytest.mark.parametrize("degree", [2, 4])
def test_fp32_gmg_cg_poisson_3d(gpu, degree):
    """FP64 CG preconditioned by the FP32 Chebyshev V-cycle reaches 1e-10 and
    the Jacobi-CG solution; it needs at most two iterations more than the FP64
    V-cycle."""
    pr = P.poisson(3, None, degree, refine_depth=2, kappa=P.kappa_c2)
    pol64 = GeometricMultigridPreconditioner(smoother="chebyshev")
    pol32 = GeometricMultigridPreconditioner(smoother="chebyshev", precision="fp32")
    s64 = g.solve(pr, solver=g.CGSolver(preconditioner=pol64), relative_tolerance=1e-10)
    s32 = g.solve(pr, solver=g.CGSolver(preconditioner=pol32), relative_tolerance=1e-10)
    jac = g.solve(g.compile(pr), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-12)
    scale = np.abs(jac.coefficients).max()
    assert s32.info.converged == 1
    assert np.abs(s32.coefficients - jac.coefficients).max() <= 1e-8 * scale
    assert s32.info.iterations <= s64.info.iterations + 2


def test_fp32_graph_replays_match_eager(gpu):
    """Replays of the captured FP32 V-cycle agree with the eager one to FP32
    rounding (the brick kernels add shared lattice points with LDS atomics,
    whose order varies from run to run)."""
    make = lambda: P.poisson(3, None, 4, refine_depth=1, kappa=P.kappa_c2, rhs=False)
    hg = compile_geometric_multigrid(make(), GeometricMultigridPreconditioner(smoother="chebyshev", graph=True,
                                                                              precision="fp32"))
    he = compile_geometric_multigrid(make(), GeometricMultigridPreconditioner(smoother="chebyshev", graph=False,
                                                                              precision="fp32"))
    n = hg.plans[-1].reduced_dof_count
    for seed in range(4):
        r = P.rng_vector(n, 200 + seed)
        zg = _device_vcycle(hg, r)
        g.apply(hg.plans[-1], P.rng_vector(hg.plans[-1].full_dof_count, seed))
        ze = _device_vcycle(he, r)
        gpu.cuda.synchronize()
        assert np.abs(zg - ze).max() <= 1e-5 * np.abs(ze).max()


def test_precision_policy_errors():
    with pytest.raises(ValueError):
        GeometricMultigridPreconditioner(precision="fp16")


# ---- nonsymmetric operators: GMRES smoother, LU / FGMRES coarse solves -----------

def _advection(cells, degree, b=(3.0, 1.0)):
    d = g.Domain((0.0, 0.0), (1.0, 1.0), (1, 1))
    g.uniform_refine(d, int(round(np.log2(cells))))
    sp = g.HpSpace(d, g.SpaceOptions(degree=g.UniformDegree(degree)))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.NonsymmetricOperator())
    pr.add_cell_accumulator(u, u, P.MassDiffusion(1.0, 1.0, b=b))
    pr.add_cell_accumulator(u, P.SineLoad(2))
    for axis in range(2):
        for side in (g.LOWER, g.UPPER):
            pr.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    return pr


@pytest.mark.parametrize("cells,degree,limit", [(4, 2, 512), (8, 3, 512), (8, 2, 10)])
def test_gmres_vcycle_matches_oracle(gpu, cells, degree, limit):
    """Device V-cycle with the reference's GMRES smoother against the oracle's
    (least-squares restatement); coarse solve dense LU, or FGMRES when the
    coarse level exceeds `limit` (agreement then to the coarse tolerance)."""
    pol = GeometricMultigridPreconditioner(coarse_direct_dof_limit=limit)
    h = compile_geometric_multigrid(_advection(cells, degree), pol)
    assert h.smoother == "gmres"
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 21)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r, smoother="gmres", pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, omega=pol.smoother_damping, restart=pol.smoother_restart)
    tol = 1e-10 if h.coarse_solver_kind() == "dense_cholesky" else 1e-8
    assert np.abs(z - ref).max() <= tol * np.abs(ref).max()


def test_fgmres_gmg_nonsymmetric_solve(gpu):
    """FGMRES with the GMRES-smoothed V-cycle reaches the dense oracle solution
    in fewer iterations than FGMRES with Jacobi."""
    pr = _advection(16, 2)
    st = g.solve(pr, solver=g.FGMRESSolver(preconditioner=GeometricMultigridPreconditioner()),
                 relative_tolerance=1e-10)
    plan = st.plan
    A = pyoracle.reduced_matrix(plan)
    x = np.linalg.solve(A, plan.rhs_physical()[plan.solve_dofs])
    ref = np.zeros(plan.full_dof_count)
    ref[plan.solve_dofs] = x
    assert np.abs(st.coefficients - ref).max() <= 1e-8 * np.abs(ref).max()
    jac = g.solve(g.compile(pr), solver=g.FGMRESSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-10)
    assert st.info.iterations < jac.info.iterations


def test_device_dense_coarse_inverse_spd(gpu):
    """A coarse level above the reference's 512-dof direct limit is inverted on
    the device (Cholesky factorization and inverse): V-cycle against the
    oracle's dense V-cycle (exact coarse solve)."""
    problem = P.poisson(2, None, 2, refine_depth=5, rhs=False)
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", max_levels=2, coarse_direct_dof_limit=2000)
    h = compile_geometric_multigrid(problem, pol)
    assert h.reduced_dof_counts()[0] > 512 and h.coarse_solver_kind() == "dense_cholesky"
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 31)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r, smoother="chebyshev", pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, lambdas=h.level_lambda_max(),
                          cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()


def test_device_dense_coarse_inverse_lu(gpu):
    """Nonsymmetric coarse level above 512 dofs: device LU inverse."""
    pol = GeometricMultigridPreconditioner(max_levels=2, coarse_direct_dof_limit=2000)
    h = compile_geometric_multigrid(_advection(32, 2), pol)
    assert h.reduced_dof_counts()[0] > 512
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 32)
    z = _device_vcycle(h, r)
    ref = pyoracle.vcycle(mats, Ps, r, smoother="gmres", pre=pol.pre_smoothing_steps,
                          post=pol.post_smoothing_steps, omega=pol.smoother_damping, restart=pol.smoother_restart)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()
