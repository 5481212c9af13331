"""CPU checks that pin the oracle (the reference algorithm restated in C under
oracle/) and the host plan compiler against the reference's own known-answer
tests and against what the mathematics fixes.

Each check is chosen so that a plausible slip in the restatement (a dropped
quadrature weight, a wrong Jacobian power, a transposed table, a lost
constraint term, a wrong sign in a Krylov update) fails it:

* polynomial KATs from test/core.jl:152-167 and the closed form
  psi_n = (P_n - P_{n-2}) / sqrt(4n - 2) against numpy's Legendre series;
* Gauss rules against numpy.polynomial.legendre.leggauss and exactness to
  degree 2n - 1 (test/core.jl:289-351);
* apply(mass_plan, ones(2)) == [0.5, 0.5] (test/matrix_free.jl:564);
* energy of projected polynomials (closed-form integrals) on uniform,
  hanging and 3D meshes -- exercises the weights, J^-1 scaling, the
  derivative tables, the term maps and the hanging-face coefficients;
* the tensor (sum-factorized) path equals the full-table path
  (src/accumulators.jl:1022-1054 vs 1955-2013) and the lowered path;
* symmetry of SPD accumulators, constants in the kernel of pure diffusion;
* local diagonal == probed operator diagonal (test/matrix_free.jl:514-535);
* oracle CG == dense solve; quadratic manufactured solution with nonzero
  Dirichlet data exact (test/verification.jl:87-109).
"""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.quadrature import basis_tables, fe_basis_values_and_derivatives, gauss_legendre_rule

import problems as P


# ---------------------------------------------------------------- polynomials

def test_integrated_legendre_kats():
    v, d = fe_basis_values_and_derivatives([-1.0, 1.0], 4)
    assert np.allclose(v[0], [1, 0, 0, 0, 0], atol=1e-14)
    assert np.allclose(v[1], [0, 1, 0, 0, 0], atol=1e-14)
    _, d = fe_basis_values_and_derivatives([0.0], 1)
    assert np.allclose(d[0], [-0.5, 0.5], atol=1e-14)
    _, d = fe_basis_values_and_derivatives([0.0], 2)
    assert np.allclose(d[0], [-0.5, 0.5, 0.0], atol=1e-14)
    v, _ = fe_basis_values_and_derivatives([0.0], 2)
    assert v[0][2] == pytest.approx(-0.6123724356957945, abs=1e-15)  # -sqrt(6)/4


def test_integrated_legendre_closed_form():
    from numpy.polynomial import legendre as L
    x = np.linspace(-1, 1, 37)
    p = 9
    v, d = fe_basis_values_and_derivatives(x, p)
    assert np.allclose(v[:, 0], (1 - x) / 2, atol=1e-15)
    assert np.allclose(v[:, 1], (1 + x) / 2, atol=1e-15)
    for n in range(2, p + 1):
        cn = np.zeros(n + 1)
        cn[n] = 1.0
        cm = np.zeros(n + 1)
        cm[n - 2] = 1.0
        ref = (L.legval(x, cn) - L.legval(x, cm)) / np.sqrt(4 * n - 2)
        assert np.allclose(v[:, n], ref, atol=1e-13)
        dref = np.sqrt((2 * n - 1) / 2.0) * L.legval(x, np.eye(n)[n - 1])
        assert np.allclose(d[:, n], dref, atol=1e-12)
    v0, d0 = fe_basis_values_and_derivatives(x, 0)
    assert np.all(v0 == 1.0) and np.all(d0 == 0.0)


@pytest.mark.parametrize("n", list(range(1, 21)))
def test_gauss_legendre_vs_numpy(n):
    pts, wts = gauss_legendre_rule(n)
    rp, rw = np.polynomial.legendre.leggauss(n)
    assert np.allclose(pts, rp, atol=1e-14)
    assert np.allclose(wts, rw, atol=1e-14)
    assert np.all(np.diff(pts) > 0)
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert float(np.dot(wts, pts ** k)) == pytest.approx(exact, abs=1e-13)


def test_basis_tables_layout():
    B, G = basis_tables(3, 4)
    pts, _ = gauss_legendre_rule(4)
    v, d = fe_basis_values_and_derivatives(pts, 3)
    assert B.shape == (4, 4) and np.array_equal(B, v.T) and np.array_equal(G, d.T)


# ------------------------------------------------------------ dense helpers

def dense_physical(plan, fn=pyoracle.apply):
    n = plan.full_dof_count
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = fn(plan, e, physical=True)
    return A


def projected(dim, exact, domain_builder, degree=2, basis=None):
    """Coefficients of the L2 projection of `exact` onto the space (mass solve
    with the oracle's mass matrix) plus the space's stiffness matrix."""

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            return exact(q.point)

    domain = domain_builder()
    space = g.HpSpace(domain, g.SpaceOptions(basis=basis or g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    mp = g.AffineProblem(u, g.SPD())
    mp.add_cell_accumulator(u, u, P.Mass())
    mp.add_cell_accumulator(u, Load())
    mplan = g.compile(mp)
    M = dense_physical(mplan)
    c = np.linalg.solve(M, mplan.rhs_physical())
    kp = g.AffineProblem(u, g.SPD())
    kp.add_cell_accumulator(u, u, P.Diffusion())
    K = dense_physical(g.compile(kp))
    return c, M, K


def _unit_square():
    return g.Domain((0.0, 0.0), (1.0, 1.0), (2, 2))


def _hanging():
    d = g.Domain((0.0, 0.0), (2.0, 1.0), (2, 1))
    d.refine(1, 1)
    return d


def _hanging_deep():
    d = g.Domain((0.0, 0.0), (2.0, 1.0), (2, 1))
    f = d.refine(1, 0)
    f2 = d.refine(f, 1)
    d.refine(f2, 0)
    return d


def _unit_cube():
    return g.Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (2, 2, 2))


# ----------------------------------------------------------------- goldens

def test_mass_golden_dg_p1():
    """apply(mass_plan, ones(2)) == [0.5, 0.5] (test/matrix_free.jl:538-564)."""
    domain = g.Domain((0.0,), (1.0,), (1,))
    space = g.HpSpace(domain, g.SpaceOptions(degree=g.UniformDegree(1), continuity="dg"))
    u = g.ScalarField(space, "u")
    p = g.AffineProblem(u, g.SPD())
    p.add_cell_accumulator(u, u, P.Mass())
    plan = g.compile(p)
    assert np.allclose(pyoracle.apply(plan, np.ones(2)), [0.5, 0.5], atol=1e-15)
    assert np.allclose(pyoracle.apply_lowered(plan, np.ones(2)), [0.5, 0.5], atol=1e-15)
    # exact local mass [[1/3, 1/6], [1/6, 1/3]] on [0, 1]
    assert np.allclose(dense_physical(plan), [[1 / 3, 1 / 6], [1 / 6, 1 / 3]], atol=1e-15)


@pytest.mark.parametrize("case", ["square", "hanging", "hanging_deep", "cube"])
def test_projected_polynomial_energy(case):
    """c^T K c equals the closed-form energy of a polynomial inside the space;
    c^T M 1 equals the domain measure."""
    if case == "square":
        dim, build, u, energy, vol = 2, _unit_square, lambda p: p[0] ** 2 + p[0] * p[1], 3.0, 1.0
    elif case == "hanging":
        dim, build, u, energy, vol = 2, _hanging, lambda p: p[0] ** 2 + p[0] * p[1], 18.0, 2.0
    elif case == "hanging_deep":
        dim, build, u, energy, vol = 2, _hanging_deep, lambda p: p[0] ** 2 + p[0] * p[1], 18.0, 2.0
    else:
        dim, build, u, energy, vol = 3, _unit_cube, lambda p: p[0] ** 2 + p[1] * p[2], 2.0, 1.0
    c, M, K = projected(dim, u, build)
    assert float(c @ K @ c) == pytest.approx(energy, rel=1e-11)
    c1, _, _ = projected(dim, lambda p: np.ones_like(p[0]), build)
    assert float(c1 @ M @ c1) == pytest.approx(vol, rel=1e-12)
    assert np.abs(K @ c1).max() < 1e-11  # constants in the kernel of diffusion


# --------------------------------------------------- path equivalences

def _plans():
    return {
        "2d": g.compile(P.poisson(2, 3, 2, kappa=lambda p: 1.0 + p[0] * p[1])),
        "2d_trunk": g.compile(P.poisson(2, 3, 3, basis=g.TrunkBasis())),
        "2d_general": g.compile(P.poisson(2, 2, 2, operator=P.MassDiffusion(0.7, 1.3, b=(0.2, -0.4)))),
        "hanging": g.compile(P.hanging_problem(2)),
        "hanging3d": g.compile(P.hanging_problem(2, dim=3)),
        "corner": g.compile(P.corner_refined(2, base=2, levels=2, degree=2)),
        "3d": g.compile(P.poisson(3, 2, 3, kappa=P.kappa_c2)),
        "dg": g.compile(P.poisson(2, 2, 2, continuity="dg", dirichlet=False)),
    }


PLANS = None


def plans():
    global PLANS
    if PLANS is None:
        PLANS = _plans()
    return PLANS


@pytest.mark.parametrize("name", ["2d", "2d_trunk", "2d_general", "hanging", "hanging3d", "corner", "3d", "dg"])
def test_tensor_equals_full_table_and_lowered(name):
    plan = plans()[name]
    x = P.rng_vector(plan.full_dof_count, 3)
    y = pyoracle.apply(plan, x)
    yb = pyoracle.apply_basis(plan, x)
    yl = pyoracle.apply_lowered(plan, x, threads=2)
    scale = np.abs(y).max()
    assert np.abs(y - yb).max() <= 1e-12 * scale
    assert np.abs(y - yl).max() <= 1e-12 * scale


@pytest.mark.parametrize("name", ["2d", "hanging", "hanging3d", "corner", "3d"])
def test_spd_symmetry(name):
    plan = plans()[name]
    A = dense_physical(plan, pyoracle.apply_lowered)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    red = A[np.ix_(plan.solve_dofs, plan.solve_dofs)]
    assert np.linalg.eigvalsh(red).min() > 0


@pytest.mark.parametrize("name", ["2d", "2d_trunk", "2d_general", "3d", "dg"])
def test_local_diagonal_equals_probe(name):
    plan = plans()[name]
    A = dense_physical(plan, pyoracle.apply_lowered)
    d = pyoracle.diagonal_physical(plan)
    assert np.allclose(d, np.diag(A), rtol=1e-12, atol=1e-14 * np.abs(A).max())


def test_fixed_rows_identity():
    plan = plans()["2d"]
    x = P.rng_vector(plan.full_dof_count, 5)
    y = pyoracle.apply(plan, x)
    assert np.array_equal(y[plan.fixed_dofs], x[plan.fixed_dofs])
    # the gather reads every coefficient; only the scatter skips fixed rows
    # (src/assembly.jl:540-564)
    phys = pyoracle.apply(plan, x, physical=True)
    assert np.allclose(y[plan.solve_dofs], phys[plan.solve_dofs], rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ solves

@pytest.mark.parametrize("name", ["2d", "hanging3d", "corner", "3d"])
def test_oracle_cg_equals_dense(name):
    plan = plans()[name]
    if plan.reduced_dof_count == plan.full_dof_count:
        pytest.skip("pure Neumann problem")
    A = dense_physical(plan, pyoracle.apply_lowered)[np.ix_(plan.solve_dofs, plan.solve_dofs)]
    b = P.rng_vector(plan.reduced_dof_count, 9)
    x, it = pyoracle.cg(plan, b, inv_diag=1.0 / np.diag(A), rtol=1e-13)
    assert it > 0
    ref = np.linalg.solve(A, b)
    assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()


def test_quadratic_manufactured_nonzero_dirichlet():
    """-u'' = 2, u(0) = u(1) = 1 on two p=2 cells: u = x(1-x) + 1 exactly
    (test/verification.jl:87-109)."""

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            return 2.0 + 0.0 * q.point[0]

    domain = g.Domain((0.0,), (1.0,), (2,))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    p = g.AffineProblem(u, g.SPD())
    p.add_cell_accumulator(u, u, P.Diffusion())
    p.add_cell_accumulator(u, Load())
    p.add_constraint(g.Dirichlet(u, g.BoundaryFace(0, g.LOWER), 1.0))
    p.add_constraint(g.Dirichlet(u, g.BoundaryFace(0, g.UPPER), 1.0))
    plan = g.compile(p)
    assert np.allclose(plan.fixed_values, 1.0, atol=1e-14)
    b = (plan.rhs_physical() - pyoracle.apply(plan, plan.shift, physical=True))[plan.solve_dofs]
    x, _ = pyoracle.cg(plan, b, rtol=1e-14)
    full = plan.shift.copy()
    full[plan.solve_dofs] = x
    state = g.State(plan, full)
    exact = lambda pt: pt[0] * (1.0 - pt[0]) + 1.0
    assert g.l2_error(state, exact) <= 1e-12
    assert g.relative_l2_error(state, exact) <= 1e-12


def test_l2_error_constant_goldens():
    """l2_error of the zero state against 1 on [0,1] is 1; against (x) is sqrt(1/3)."""
    domain = g.Domain((0.0,), (1.0,), (1,))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(1)))
    u = g.ScalarField(space, "u")
    p = g.AffineProblem(u, g.SPD())
    p.add_cell_accumulator(u, u, P.Mass())
    plan = g.compile(p)
    state = g.State(plan, np.zeros(plan.full_dof_count))
    assert g.l2_error(state, lambda pt: 1.0 + 0.0 * pt[0]) == pytest.approx(1.0, abs=1e-14)
    assert g.l2_error(state, lambda pt: pt[0]) == pytest.approx(np.sqrt(1.0 / 3.0), abs=1e-14)


# ---- oracle right-hand side and verification sums, pinned by closed forms ----

class _Load:
    def __init__(self, fn):
        self.fn = fn

    def cell_rhs_accumulate(self, q, test_component):
        return self.fn(q.point)


def _mass_problem(load, degree=2):
    d = g.Domain((0.0,) * 3, (1.0,) * 3, (2, 1, 2))
    sp = g.HpSpace(d, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(sp, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())
    pr.add_cell_accumulator(u, _Load(load))
    return pr


def _mass_projection(load):
    plan = g.compile(_mass_problem(load))
    r = pyoracle.rhs_from_loads(plan, *plan.lowered_loads())
    M = pyoracle.reduced_matrix(plan)  # no constraints: reduced = full
    return plan, r, np.linalg.solve(M, r)


def test_oracle_rhs_and_l2_constant():
    """Load 1: the mass projection c of 1 is exact, c.r = volume, u_h - 1 = 0."""
    plan, r, c = _mass_projection(lambda p: 1.0 + 0.0 * p[0])
    assert float(c @ r) == pytest.approx(1.0, rel=1e-12)
    w = plan.point_weights()
    s0, s1 = pyoracle.l2_sums(plan, c, w, plan.point_values(lambda p: 1.0 + 0.0 * p[0]))
    assert s0 < 1e-24 and s1 == pytest.approx(1.0, rel=1e-12)


def test_oracle_rhs_gradient_channel_and_l2_moment():
    """Gradient load (1, 0, 0) against the exact interpolant of x:
    integral of d x / d x = 1; integral of x^2 = 1/3."""
    plan_x, _, cx = _mass_projection(lambda p: p[0])

    class GradLoad:
        def cell_rhs_accumulate(self, q, test_component):
            z = 0.0 * q.point[0]
            return g.TestChannels(z, (1.0 + z, z, z))

    d = plan_x.problem
    pr = g.AffineProblem(d.field, g.SPD())
    pr.add_cell_accumulator(d.field, d.field, P.Mass())
    pr.add_cell_accumulator(d.field, GradLoad())
    plan_g = g.compile(pr)
    loads, channels = plan_g.lowered_loads()
    assert channels == 4
    rg = pyoracle.rhs_from_loads(plan_g, loads, channels)
    assert float(cx @ rg) == pytest.approx(1.0, rel=1e-12)
    w = plan_x.point_weights()
    s0, _ = pyoracle.l2_sums(plan_x, cx, w, plan_x.point_values(lambda p: p[0]))
    assert s0 < 1e-24
    s0, _ = pyoracle.l2_sums(plan_x, cx, w, plan_x.point_values(lambda p: 0.0 * p[0]))
    assert s0 == pytest.approx(1.0 / 3.0, rel=1e-12)


def test_assembled_matrix_equals_reduced_apply():
    """The oracle's element-matrix assembly against its own apply: tensor
    cells with a general point action, cut cells with random rules and
    surface pieces, and the finite-cell ball."""
    from test_cloud_host import Nitsche
    from gricoamd.problem import PointQuadrature, SurfaceQuadrature
    rng = np.random.default_rng(4)
    pr = P.poisson(3, 2, 2, operator=P.MassDiffusion(0.7, 1.3, b=(0.2, -0.1, 0.3)), rhs=False, dirichlet=False)
    for leaf in pr.field.space.leaves[::3]:
        n = int(rng.integers(3, 40))
        pr.add_cell_quadrature(int(leaf), PointQuadrature(rng.uniform(-1, 1, (n, 3)), rng.uniform(0.01, 0.2, n)))
    for leaf in pr.field.space.leaves[::2]:
        pr.add_surface_quadrature(SurfaceQuadrature(int(leaf), PointQuadrature(rng.uniform(-1, 1, (5, 3)),
                                                                               rng.uniform(0.05, 0.3, 5)),
                                                    rng.standard_normal((5, 3))))
    pr.add_surface_accumulator(pr.field, pr.field, Nitsche(6.0))
    for problem in (pr, P.fcm_ball(n=3, degree=2, depth=1)):
        plan = g.compile(problem)
        A = pyoracle.assembled_reduced_matrix(plan)
        x = P.rng_vector(plan.reduced_dof_count, 8)
        y = pyoracle.reduced_apply(plan, x)
        assert np.abs(A @ x - y).max() <= 1e-13 * np.abs(y).max()
