"""hp-multigrid hierarchy and transfers on the host (CPU).

Level structure goldens from the reference's own tests (test/multigrid.jl:
98-104 two-level DG, 153-160 [2, 4, 6], 162-171 [2, 4, 8], 173-178 hanging),
the 64^3-by-refinement hierarchy structure (p 4 -> 2 -> 1, then binary
merges; src/multigrid.jl:443-469), transfer adjointness (test/multigrid.jl:
49-57) and the Galerkin identity P^T A_fine P = A_coarse, which holds exactly
because every coarse space embeds in the next finer one and the Gauss rules
integrate both levels' products exactly: a wrong 1D embedding matrix, a lost
hanging term or a wrong reduced index breaks it.
"""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid

import problems as P


def dg_mass(degree, cells, refine=None):
    d = g.Domain((0.0,), (1.0,), cells)
    if refine:
        refine(d)
    sp = g.HpSpace(d, g.SpaceOptions(degree=g.UniformDegree(degree), continuity="dg"))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())
    return pr


def test_level_goldens():
    def two_level_split(d):
        first = d.refine(0, 0)
        d.refine(first + 1, 0)

    h = compile_geometric_multigrid(dg_mass(1, (1,), two_level_split))
    assert h.reduced_dof_counts() == [2, 4, 6]
    assert h.coarse_solver_kind() == "dense_cholesky"
    h = compile_geometric_multigrid(dg_mass(3, (1,), lambda d: d.refine(0, 0)))
    assert h.reduced_dof_counts() == [2, 4, 8]
    h = compile_geometric_multigrid(dg_mass(3, (2,)))
    assert h.level_count == 2


def test_hanging_two_levels():
    d = g.Domain((0.0, 0.0), (1.0, 1.0), (2, 1))
    d.refine(0, 0)
    sp = g.HpSpace(d, g.SpaceOptions(degree=g.UniformDegree(1)))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())
    h = compile_geometric_multigrid(pr)
    assert h.level_count == 2


def test_coarse_solver_threshold():
    """22x22 p=2 mass: the p=1 coarse level has > 512 dofs -> Krylov coarse
    solve, unless the dense limit is raised (test/multigrid.jl:119-129)."""
    d = g.Domain((0.0, 0.0), (1.0, 1.0), (22, 22))
    sp = g.HpSpace(d, g.SpaceOptions(degree=g.UniformDegree(2)))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())
    h = compile_geometric_multigrid(pr)
    assert h.plans[0].reduced_dof_count > 512
    assert h.coarse_solver_kind() == "krylov"
    h = compile_geometric_multigrid(pr, GeometricMultigridPreconditioner(coarse_direct_dof_limit=4096))
    assert h.coarse_solver_kind() == "dense_cholesky"


def test_refined_cube_hierarchy_structure():
    """Depth-3 uniform refinement of one root, p=4: three p-levels on the fine
    mesh, then one binary merge per level down to the root cell."""
    h = compile_geometric_multigrid(P.poisson(3, None, 4, refine_depth=3, rhs=False))
    leaves = [len(sp.leaves) for sp in h.spaces]
    degrees = [int(sp.degrees[0][0]) for sp in h.spaces]
    assert leaves == [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 512, 512]
    assert degrees == [1] * 10 + [2, 4]
    assert h.reduced_dof_counts()[-3:] == [343, 15 ** 3, 31 ** 3]
    # with max_levels = 16 the 64^3 construction keeps 13 merges: coarsest 4x4x2 cells
    h16 = GeometricMultigridPreconditioner()
    assert h16.max_levels == 16


def test_policy_validation():
    with pytest.raises(ValueError):
        GeometricMultigridPreconditioner(smoother_damping=float("inf"))
    with pytest.raises(ValueError):
        GeometricMultigridPreconditioner(coarse_relative_tolerance=float("inf"))
    with pytest.raises(ValueError):
        GeometricMultigridPreconditioner(coarse_absolute_tolerance=float("nan"))
    with pytest.raises(ValueError):
        GeometricMultigridPreconditioner(p_sequence="halve")


CASES = {
    "p_2d": lambda: P.poisson(2, 3, 4, rhs=False),
    "h_2d": lambda: P.poisson(2, None, 1, refine_depth=2, rhs=False),
    "hp_3d": lambda: P.poisson(3, None, 2, refine_depth=1, kappa=P.kappa_c2, rhs=False),
    "corner": lambda: P.corner_refined(2, base=2, levels=1, degree=2),
    "dg": lambda: P.poisson(2, None, 2, refine_depth=1, continuity="dg", dirichlet=False, rhs=False,
                            operator=P.MassDiffusion(1.0, 1.0)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_transfer_galerkin_identity(name):
    h = compile_geometric_multigrid(CASES[name]())
    mats = [pyoracle.reduced_matrix(p) for p in h.plans]
    rng = np.random.default_rng(4)
    for l, t in enumerate(h.transfers):
        Pm = t.matrix()
        Ac, Af = mats[l], mats[l + 1]
        if Ac.size == 0:
            continue
        G = Pm.T @ Af @ Pm
        assert np.abs(G - Ac).max() <= 1e-11 * np.abs(Ac).max()
        # adjointness of the host prolongate/restrict pair
        c = np.sin(0.2 * np.arange(1, t.coarse_plan.reduced_dof_count + 1))
        f = np.cos(0.3 * np.arange(1, t.fine_plan.reduced_dof_count + 1))
        pf = t.prolongate_add(np.zeros_like(f), c)
        rc = t.restrict(f)
        assert float(pf @ f) == pytest.approx(float(c @ rc), rel=1e-13, abs=1e-13)
        assert np.all(np.diff(t.fine_indices) >= 0)
        x = rng.standard_normal(Pm.shape[1])
        assert np.allclose(Pm @ x, t.prolongate_add(np.zeros(Pm.shape[0]), x), atol=1e-14)


def test_oracle_vcycle_is_a_contraction():
    """The oracle V-cycle (dense restatement) reduces the error of a Poisson
    problem by a mesh-independent factor (damped Jacobi V(2,2) at p = 2:
    about 0.62 per cycle as a stationary iteration)."""
    rates = []
    for depth in (2, 3):
        h = compile_geometric_multigrid(P.poisson(2, None, 2, refine_depth=depth, rhs=False))
        mats = [pyoracle.reduced_matrix(p) for p in h.plans]
        Ps = [t.matrix() for t in h.transfers]
        A = mats[-1]
        b = np.random.default_rng(0).standard_normal(A.shape[0])
        x = np.zeros_like(b)
        for _ in range(6):
            x = x + pyoracle.vcycle(mats, Ps, b - A @ x)
        err = np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(np.linalg.solve(A, b))
        rates.append(err ** (1 / 6))
    assert max(rates) < 0.7
    assert abs(rates[0] - rates[1]) < 0.05


# ---- GMRES smoother (nonsymmetric operators) --------------------------------------

def _nonsym_matrix(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * 0.3 + np.diag(4.0 + rng.random(n))
    return A


def test_oracle_gmres_smoother_full_krylov_space_solves():
    """With n Arnoldi steps and damping 1 the GMRES smoother minimizes the
    residual over the whole space: x = A^-1 b."""
    A = _nonsym_matrix(12, 1)
    b = np.random.default_rng(2).standard_normal(12)
    x0 = np.random.default_rng(3).standard_normal(12)
    x = pyoracle.gmres_smooth(A, 1.0 / np.diag(A), x0, b, 12, 12, 1.0)
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-10)


def test_oracle_gmres_smoother_one_step_is_minimal_residual():
    """One Arnoldi step: x += omega (r . Az) / (Az . Az) z with z = D^-1 r."""
    A = _nonsym_matrix(9, 4)
    b = np.random.default_rng(5).standard_normal(9)
    x0 = np.random.default_rng(6).standard_normal(9)
    dinv = 1.0 / np.diag(A)
    r = b - A @ x0
    z = dinv * r
    w = A @ z
    expect = x0 + 0.8 * (r @ w) / (w @ w) * z
    assert np.allclose(pyoracle.gmres_smooth(A, dinv, x0, b, 1, 8, 0.8), expect, atol=1e-13)


def _advection_problem(cells=4, degree=2, b=(3.0, 1.0)):
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


def test_gmres_smoother_policy_checks():
    pr = _advection_problem()
    with pytest.raises(ValueError):
        compile_geometric_multigrid(pr, GeometricMultigridPreconditioner(pre_smoothing_steps=9,
                                                                          post_smoothing_steps=9))
    with pytest.raises(ValueError):
        compile_geometric_multigrid(pr, GeometricMultigridPreconditioner(precision="fp32"))
    h = compile_geometric_multigrid(pr, GeometricMultigridPreconditioner())
    assert h.smoother == "gmres"


@pytest.mark.parametrize("steps", [2, 3])
@pytest.mark.parametrize("name", ["cube_p2", "corner_3d"])
def test_merged_h_steps_galerkin_identity(name, steps):
    """h_steps_per_level = k keeps every k-th h-level; the kept levels'
    prolongation (a product of k single-merge transfers) must still satisfy
    P^T A_fine P = A_coarse exactly, and the kept spaces are the single-step
    hierarchy's spaces at every k-th h-level."""
    problem = {"cube_p2": lambda: P.poisson(3, None, 2, refine_depth=2, rhs=False),
               "corner_3d": lambda: P.corner_refined(3, base=2, levels=1, degree=2)}[name]()
    h1 = compile_geometric_multigrid(problem)
    hk = compile_geometric_multigrid(problem, GeometricMultigridPreconditioner(h_steps_per_level=steps))
    leaves1 = [len(sp.leaves) for sp in h1.spaces]
    p_levels = sum(1 for sp in h1.spaces if len(sp.leaves) == leaves1[-1]) - 1
    n_h = len(leaves1) - 1 - p_levels
    picked = sorted({0} | set(range(n_h % steps, n_h, steps)))
    assert [len(sp.leaves) for sp in hk.spaces] == [leaves1[i] for i in picked] + leaves1[n_h:]
    assert hk.level_count == len(picked) + p_levels + 1
    mats = [pyoracle.reduced_matrix(p) for p in hk.plans]
    for l, t in enumerate(hk.transfers):
        Ac, Af = mats[l], mats[l + 1]
        assert t.coarse_plan is hk.plans[l] and t.fine_plan is hk.plans[l + 1]
        order = np.lexsort((t.coarse_indices, t.fine_indices))
        assert np.array_equal(order, np.arange(t.nnz)), "entries sorted by (fine, coarse)"
        if Ac.size == 0:
            continue
        Pm = t.matrix()
        assert np.abs(Pm.T @ Af @ Pm - Ac).max() <= 1e-11 * np.abs(Ac).max()
    with pytest.raises(ValueError):
        h1.with_policy(GeometricMultigridPreconditioner(h_steps_per_level=steps))


def test_dense_coarse_level_size_is_checked():
    """A dense coarse level above 46340 dofs (n^2 >= 2^31, beyond rocSOLVER's
    32-bit sizes) is rejected before anything is built on the device."""
    from gricoamd import multigrid as mgm
    h = compile_geometric_multigrid(P.poisson(2, None, 2, refine_depth=2, rhs=False))
    h = h.with_policy(GeometricMultigridPreconditioner(max_levels=2))
    n = h.plans[0].reduced_dof_count
    assert n > 0 and h.coarse_solver_kind() == "dense_cholesky"
    assert mgm.DENSE_COARSE_MAX ** 2 < 2 ** 31 <= (mgm.DENSE_COARSE_MAX + 1) ** 2
    h._check_dense_coarse()  # a small dense coarse level passes
    saved = mgm.DENSE_COARSE_MAX
    try:
        mgm.DENSE_COARSE_MAX = n - 1
        with pytest.raises(ValueError, match="dense coarse multigrid level"):
            h._check_dense_coarse()
    finally:
        mgm.DENSE_COARSE_MAX = saved


# ---- block smoother (oracle) and physical-domain hierarchies ----------------------

def _spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return M @ M.T + n * np.eye(n)


def test_oracle_block_smoother_closed_forms():
    A = _spd(12, 1)
    b = np.random.default_rng(2).standard_normal(12)
    # one block holding every dof: S = A^-1
    S = pyoracle.block_smoother(A, [np.arange(12)])
    assert np.allclose(S(b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    # single-dof blocks: S = D^-1
    S = pyoracle.block_smoother(A, [[i] for i in range(12)])
    assert np.allclose(S(b), b / np.diag(A), rtol=1e-14)
    # no blocks: point Jacobi
    S = pyoracle.block_smoother(A, [])
    assert np.allclose(S(b), b / np.diag(A), rtol=1e-14)


def test_oracle_block_smoother_is_spd_with_overlap():
    A = _spd(15, 3)
    blocks = [np.arange(0, 6), np.arange(4, 10), np.array([9, 11, 12])]
    S = pyoracle.block_smoother(A, blocks)
    M = np.stack([S(e) for e in np.eye(15)], axis=1)
    assert np.allclose(M, M.T, atol=1e-13)
    assert np.linalg.eigvalsh(M).min() > 0.0
    # a dof no block covers keeps its Jacobi entry
    assert np.isclose(M[14, 14], 1.0 / A[14, 14]) and np.allclose(M[14, :14], 0.0)


def test_oracle_chebyshev_with_point_blocks_equals_jacobi_chebyshev():
    A = _spd(10, 4)
    b = np.random.default_rng(5).standard_normal(10)
    dinv = 1.0 / np.diag(A)
    S = pyoracle.block_smoother(A, [[i] for i in range(10)])
    x1 = pyoracle._chebyshev(A, dinv, np.zeros(10), b, True, 3.0, 4, 20.0)
    x2 = pyoracle._chebyshev(A, S, np.zeros(10), b, True, 3.0, 4, 20.0)
    assert np.allclose(x1, x2, rtol=1e-14, atol=1e-14)


def test_physical_domain_rules_equal_explicit_rules():
    """compile() on a PhysicalDomain generates the rules and surface pieces
    that the explicit attachments give: the same operator."""
    auto = g.compile(P.fcm_ball_physical(n=4, degree=2, depth=2))
    expl = g.compile(P.fcm_ball_physical(n=4, degree=2, depth=2, explicit=True))
    assert auto.full_dof_count == expl.full_dof_count
    assert len(auto.cloud_buckets) and len(auto.surface_buckets)
    x = P.rng_vector(auto.full_dof_count, 11)
    ya, ye = pyoracle.apply(auto, x), pyoracle.apply(expl, x)
    assert np.abs(ya - ye).max() <= 1e-14 * np.abs(ye).max()


def test_physical_domain_hierarchy_levels():
    from gricoamd.embedded import OUTSIDE
    from gricoamd.multigrid import MultigridHierarchy, block_colors, cut_cell_blocks
    pr = P.fcm_ball_physical(n=8, degree=2, depth=1)
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", h_steps_per_level=3, cell_blocks="cut")
    h = MultigridHierarchy(pr, pol)
    dom = pr.field.space.domain
    counts = h.reduced_dof_counts()
    assert counts == sorted(counts) and len(counts) >= 3
    for plan in h.plans:
        cls = dom.classify(plan.space.leaves)
        assert not np.any(cls == OUTSIDE)
        # every cut leaf of the level carries its own rule, and only those
        ruled = sorted(int(c) for cb in plan.cloud_buckets for c in cb.cells)
        assert ruled == sorted(int(l) for l in plan.space.leaves[cls == 2])
        assert plan.surface_buckets
        blocks = cut_cell_blocks(plan)
        colors = block_colors(plan, blocks)
        # no item of the plan touches dofs of two blocks of one color
        for bk in plan.all_buckets:
            red = plan.reduced_index[bk.term_dof]
            for c in range(bk.cell_count):
                touched = set(red[(bk.term_cell == c) & (red >= 0)].tolist())
                seen = {}
                for i, bl in enumerate(blocks):
                    if touched & set(bl.tolist()):
                        assert seen.setdefault(colors[i], i) == i, (c, i)


def test_explicit_rules_on_h_levels_are_rejected():
    pr = P.fcm_ball_physical(n=4, degree=2, depth=1, explicit=True)
    from gricoamd.multigrid import MultigridHierarchy
    with pytest.raises(ValueError, match="not an active leaf"):
        MultigridHierarchy(pr, GeometricMultigridPreconditioner(h_steps_per_level=1))
