"""Multigrid on finite-cell plans: the cell-block smoother's V-cycle against
the oracle's dense V-cycle with its own block inverses, and GMG-CG on a
PhysicalDomain ball (rediscretized levels with their own cut-cell rules and
sphere pieces) against the oracle's dense solve and the exact solution."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.multigrid import GeometricMultigridPreconditioner, MultigridHierarchy, compile_geometric_multigrid
from gricoamd.problem import PointQuadrature

import problems as P

pytestmark = pytest.mark.gpu


def _cut_problem(seed=5, cells=3):
    """cells^3 root cells at p = 2 (p-levels only), every other cell with its
    own random positive rule, mass + diffusion."""
    rng = np.random.default_rng(seed)
    pr = P.poisson(3, cells, 2, operator=P.MassDiffusion(0.5, 1.0), rhs=False, dirichlet=False)
    for leaf in pr.field.space.leaves[::2]:
        n = int(rng.integers(20, 120))
        pr.add_cell_quadrature(int(leaf), PointQuadrature(rng.uniform(-1, 1, (n, 3)), rng.uniform(0.01, 0.1, n)))
    return pr


def _oracle_vcycle(h, pol, r, use_blocks=True):
    mats = [pyoracle.assembled_reduced_matrix(p).toarray() for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    blocks = [pyoracle.cell_block_dofs(p) for p in h.plans] if use_blocks else None
    lambdas = h.level_lambda_max() if pol.smoother == "chebyshev" else None
    return pyoracle.vcycle(mats, Ps, r, smoother=pol.smoother, pre=pol.pre_smoothing_steps,
                           post=pol.post_smoothing_steps, omega=pol.smoother_damping, lambdas=lambdas,
                           cheb_degree=pol.chebyshev_degree, cheb_range=pol.chebyshev_range, blocks=blocks)


@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_block_vcycle_matches_oracle(gpu, smoother, precision):
    pol = GeometricMultigridPreconditioner(smoother=smoother, cell_blocks="cut", precision=precision,
                                           pre_smoothing_steps=1, post_smoothing_steps=1)
    h = compile_geometric_multigrid(_cut_problem(), pol)
    assert h.level_count == 2
    M = h.device_preconditioner()
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 3)
    z = M.apply(r).cpu().numpy()
    ref = _oracle_vcycle(h, pol, r)
    tol = 1e-10 if precision == "fp64" else 5e-5
    assert np.abs(z - ref).max() <= tol * np.abs(ref).max()
    if precision == "fp64":
        # the block smoother changes the V-cycle (the blocks were used)
        plain = _oracle_vcycle(h, pol, r, use_blocks=False) if smoother == "jacobi" else None
        if plain is not None:
            assert np.abs(plain - ref).max() > 1e-6 * np.abs(ref).max()
        r2 = P.rng_vector(len(r), 4)
        z2 = M.apply(r2).cpu().numpy()
        assert float(r2 @ z) == pytest.approx(float(r @ z2), rel=1e-10)


def test_block_probing_matches_oracle_blocks(gpu):
    from gricoamd.multigrid import block_colors, cut_cell_blocks
    from gricoamd import _devabi, _native, ops
    import ctypes
    plan = g.compile(_cut_problem(7, cells=6))
    blocks = cut_cell_blocks(plan)
    A = pyoracle.assembled_reduced_matrix(plan).toarray()
    lib = _native.device()
    _devabi.declare_mg(lib)
    lens = np.array([len(b) for b in blocks], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    dofs = np.concatenate(blocks).astype(np.int32)
    colors = block_colors(plan, blocks)
    desc = _devabi.BlockDesc()
    desc.count, desc.offsets, desc.dofs = len(blocks), _native.ptr(off, _native.c_int64), _native.ptr(dofs, _native.c_int32)
    vals = np.zeros(int((lens ** 2).sum()))
    _devabi.check(lib.gmf_reduced_blocks(ops._ws(plan).ws, ctypes.byref(desc), _native.ptr(colors, _native.c_int32),
                                         int(colors.max()) + 1, _native.ptr(vals, _native.c_double)))
    pos = 0
    for b, m in zip(blocks, lens):
        B = vals[pos:pos + m * m].reshape(m, m)
        ref = A[np.ix_(b, b)]
        assert np.abs(B - ref).max() <= 1e-12 * np.abs(ref).max()
        pos += m * m
    assert int(colors.max()) + 1 < len(blocks)  # probing shares applies between blocks


def test_fcm_ball_gmg_cg(gpu):
    pr = P.fcm_ball_physical(n=8, degree=2, depth=2)
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", cell_blocks="cut", h_steps_per_level=3,
                                           chebyshev_degree=3)
    st = g.solve(pr, solver=g.CGSolver(preconditioner=pol), relative_tolerance=1e-8)
    assert st.info.converged == 1
    plan = st.plan
    A = pyoracle.assembled_reduced_matrix(plan)
    b = plan.reduced_rhs()
    x = st.coefficients[plan.solve_dofs]
    # residual of the device solution in the oracle's operator, and the
    # oracle's sparse direct solution
    assert np.linalg.norm(b - A @ x) <= 2e-8 * np.linalg.norm(b)
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(A.tocsc(), b)
    assert np.abs(x - xd).max() <= 1e-5 * np.abs(xd).max()
    # discretization error of the finite-cell ball at 8^3 cells, p = 2,
    # subdivision depth 2, alpha = 1e-3 (not a solver tolerance)
    rel = g.relative_l2_error(st, P.ball_exact)
    assert rel < 0.03, rel
    jac = g.solve(g.compile(pr), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-8, maxiter=200000)
    assert st.info.iterations < jac.info.iterations / 5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_whole_cell_block_vcycle_matches_oracle(gpu, precision):
    # cell_blocks="cells": one block per cell (tensor cells of an adaptive hp
    # mesh with hanging nodes and degrees 1..3) on every smoothing level.
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", cell_blocks="cells", precision=precision,
                                           pre_smoothing_steps=1, post_smoothing_steps=1, chebyshev_degree=2)
    problem = P.adaptive_hp(base=4, levels=2, pmax=3)
    h = compile_geometric_multigrid(problem, pol)
    M = h.device_preconditioner()
    assert h.level_count >= 2 and max(h.block_colors) > 0
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 6)
    z = M.apply(r).cpu().numpy()
    mats = [pyoracle.assembled_reduced_matrix(p).toarray() for p in h.plans]
    Ps = [t.matrix() for t in h.transfers]
    blocks = [pyoracle.cell_block_dofs(p, all_cells=True) for p in h.plans]
    ref = pyoracle.vcycle(mats, Ps, r, smoother="chebyshev", pre=1, post=1, omega=pol.smoother_damping,
                          lambdas=h.level_lambda_max(), cheb_degree=2, cheb_range=pol.chebyshev_range, blocks=blocks)
    tol = 1e-10 if precision == "fp64" else 5e-5
    assert np.abs(z - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
@pytest.mark.parametrize("blocks", [None, "cut"])
@pytest.mark.parametrize("coarse", ["dense", "cg"])
def test_block_vcycle_matches_columns(gpu, smoother, blocks, coarse):
    # gmf_precond_apply_block: one V-cycle over 5 columns (all levels, a
    # dense coarse level as one matrix product) gives every column the
    # one-column V-cycle's result
    pol = GeometricMultigridPreconditioner(smoother=smoother, cell_blocks=blocks, pre_smoothing_steps=1,
                                           post_smoothing_steps=1,
                                           coarse_direct_dof_limit=4096 if coarse == "dense" else 0)
    h = compile_geometric_multigrid(_cut_problem(), pol)
    M = h.device_preconditioner()
    n = h.plans[-1].reduced_dof_count
    R = np.stack([P.rng_vector(n, 20 + j) for j in range(5)], axis=1)
    Z = M.apply_block(R).cpu().numpy()
    tol = 1e-12 if coarse == "dense" else 1e-9  # the coarse CG stops at its own tolerance
    for j in range(R.shape[1]):
        zj = M.apply(R[:, j]).cpu().numpy()
        assert np.abs(Z[:, j] - zj).max() <= tol * np.abs(zj).max(), j


def test_block_vcycle_physical_hierarchy_and_fp32_fallback(gpu):
    # h- and p-levels of a finite-cell ball with cut-cell blocks (FP64: one
    # block V-cycle); the FP32 V-cycle applies column by column
    pr = P.fcm_ball_physical(n=4, degree=2, depth=2, n_theta=16, n_phi=32)
    for precision in ("fp64", "fp32"):
        pol = GeometricMultigridPreconditioner(smoother="chebyshev", cell_blocks="cut", h_steps_per_level=3,
                                               chebyshev_degree=3, precision=precision,
                                               coarse_direct_dof_limit=40000)
        h = compile_geometric_multigrid(pr, pol)
        M = h.device_preconditioner()
        n = h.plans[-1].reduced_dof_count
        R = np.stack([P.rng_vector(n, 40 + j) for j in range(3)], axis=1)
        Z = M.apply_block(R).cpu().numpy()
        # FP32 V-cycles sum some contributions with atomics: two applies of
        # one column agree to FP32 rounding, not bit for bit
        tol = 1e-12 if precision == "fp64" else 1e-5
        for j in range(R.shape[1]):
            zj = M.apply(R[:, j]).cpu().numpy()
            assert np.abs(Z[:, j] - zj).max() <= tol * np.abs(zj).max(), (precision, j)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vcycle_on_grouped_bricks_matches_oracle(gpu, monkeypatch, precision):
    # 3^3 cells form eight partial bricks (host grouping with empty positions);
    # bricks forced on every level, FP64 and FP32 instantiations in the V-cycle
    monkeypatch.setenv("GMF_BRICK_MIN_CELLS", "8")
    monkeypatch.setenv("GMF_KEEP_BRICKS", "1")
    pr = P.poisson(3, 3, 2, kappa=P.kappa_c2, rhs=False)
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", precision=precision, pre_smoothing_steps=1,
                                           post_smoothing_steps=1)
    h = compile_geometric_multigrid(pr, pol)
    assert h.level_count == 2
    for p in h.plans:
        assert p.buckets[0].brick_cells is not None and (p.buckets[0].brick_cells < 0).any()
    from gricoamd import ops
    assert ops._ws(h.plans[-1]).tuning()[0]["kernel"] == "brick"
    M = h.device_preconditioner()
    r = P.rng_vector(h.plans[-1].reduced_dof_count, 8)
    z = M.apply(r).cpu().numpy()
    ref = _oracle_vcycle(h, pol, r, use_blocks=False)
    tol = 1e-10 if precision == "fp64" else 5e-5
    assert np.abs(z - ref).max() <= tol * np.abs(ref).max()
