"""Point-quadrature cells and embedded-surface pieces on the host: the oracle's
basis path (src/accumulators.jl:1022-1054, 1133-1161) pinned by the
reference's surface golden (test/matrix_free.jl:969-984), by reproducing the
tensor path when a cell's custom rule is its own Gauss rule, and by closed-form
integrals over half cells and embedded lines."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.problem import PointQuadrature, SurfaceQuadrature
from gricoamd.quadrature import fe_basis_values_and_derivatives, gauss_legendre_rule

import problems as P


class SurfaceMass:
    def __init__(self, s):
        self.s = s

    def surface_accumulate(self, q, trial, test_component):
        return self.s * trial.value


class Nitsche:
    """Symmetric Nitsche terms for -div(grad u) with u = 0 on the surface:
    -(du/dn) v - u (dv/dn) + gamma / h u v."""

    def __init__(self, gamma):
        self.gamma = gamma

    def surface_accumulate(self, q, trial, test_component):
        n = q.normal
        dn = sum(nc * gc for nc, gc in zip(n, trial.gradient))
        h = q.cell_size[0]
        return g.TestChannels(-dn + self.gamma / h * trial.value, tuple(-trial.value * nc for nc in n))


def tensor_rule(shape):
    rules = [gauss_legendre_rule(n) for n in shape]
    grids = np.meshgrid(*[r[0] for r in rules][::-1], indexing="ij")
    pts = np.stack([x.reshape(-1) for x in grids[::-1]], axis=1)
    wg = np.meshgrid(*[r[1] for r in rules][::-1], indexing="ij")
    w = np.ones_like(wg[0])
    for x in wg:
        w = w * x
    return PointQuadrature(pts, w.reshape(-1))


def test_surface_mass_golden():
    """One DG p=1 cell on [0, 1], a surface point at x = 0.5 with weight 1:
    5 u(0.5) phi_i(0.5) (test/matrix_free.jl:969-984)."""
    domain = g.Domain((0.0,), (1.0,), (1,))
    space = g.HpSpace(domain, g.SpaceOptions(degree=g.UniformDegree(1), continuity="dg"))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.NonsymmetricOperator())
    pr.add_surface_quadrature(SurfaceQuadrature(0, PointQuadrature([[0.0]], [1.0]), [[1.0]]))
    pr.add_surface_accumulator(u, u, SurfaceMass(5.0))
    plan = g.compile(pr)
    y = pyoracle.apply(plan, np.array([0.25, -0.5]))
    assert np.allclose(y, [-0.3125, -0.3125], atol=1e-15)


@pytest.mark.parametrize("dim,degree", [(2, 3), (3, 2)])
def test_own_gauss_rule_reproduces_tensor_path(dim, degree):
    base = P.poisson(dim, 2, degree, kappa=lambda p: 1.0 + 0.3 * p[0], rhs=False)
    custom = P.poisson(dim, 2, degree, kappa=lambda p: 1.0 + 0.3 * p[0], rhs=False)
    rule = tensor_rule((degree + 1,) * dim)
    for leaf in custom.field.space.leaves[::2]:
        custom.add_cell_quadrature(int(leaf), rule)
    pa, pb = g.compile(base), g.compile(custom)
    assert len(pb.cloud_buckets) == 1 and pb.cloud_buckets[0].cell_count == len(custom.field.space.leaves[::2])
    x = P.rng_vector(pa.full_dof_count, 2)
    ya, yb = pyoracle.apply(pa, x), pyoracle.apply(pb, x)
    assert np.abs(ya - yb).max() <= 1e-13 * np.abs(ya).max()
    da, db = pyoracle.diagonal_physical(pa), pyoracle.diagonal_physical(pb)
    assert np.abs(da - db).max() <= 1e-13 * np.abs(da).max()


def _fit(space, leaf, fn, degree, dim):
    """Local coefficients (box order) of a polynomial in the cell's modal basis."""
    rng = np.random.default_rng(0)
    xi = rng.uniform(-1, 1, size=(200, dim))
    cols = []
    vals = [fe_basis_values_and_derivatives(xi[:, a], degree)[0] for a in range(dim)]
    phi = np.ones((200, 1))
    for a in range(dim):
        phi = (vals[a][:, :, None] * phi[:, None, :]).reshape(200, -1)
    lo, size = space.domain.cell_lowers([leaf])[0], space.domain.cell_sizes([leaf])[0]
    x = [lo[a] + 0.5 * (xi[:, a] + 1) * size[a] for a in range(dim)]
    c, *_ = np.linalg.lstsq(phi, fn(x), rcond=None)
    return c


def test_half_cell_rule_and_embedded_line():
    """One p=2 cell on [0,1]^2: a rule on the left half integrates u = 1 to
    0.5; a line x = 0.25 with unit normal (1, 0): int u v = 1 for u = v = 1,
    and the Nitsche consistency term -int du/dn v = -0.5 for u = x^2, v = 1."""
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (1, 1))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    # half-cell volume rule: Gauss 3x3 on [-1, 0] x [-1, 1]
    r = tensor_rule((3, 3))
    pts = r.points.copy()
    pts[:, 0] = 0.5 * (pts[:, 0] - 1.0)
    half = PointQuadrature(pts, 0.5 * r.weights)
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Mass())
    pr.add_cell_quadrature(0, half)
    plan = g.compile(pr)
    one = np.zeros(plan.full_dof_count)
    for dof, c in [space.mode_terms(0, m)[0] for m in [(0, 0), (1, 0), (0, 1), (1, 1)]]:
        one[dof] = 1.0
    assert float(one @ pyoracle.apply(plan, one, physical=True)) == pytest.approx(0.5, abs=1e-14)
    # embedded line x = 0.25 <=> xi_x = -0.5; Gauss points along y, reference weights sum 2
    gp, gw = gauss_legendre_rule(4)
    line = PointQuadrature(np.stack([np.full(4, -0.5), gp], axis=1), gw)
    sp = g.AffineProblem(u, g.NonsymmetricOperator())
    sp.add_surface_quadrature(SurfaceQuadrature(0, line, [[1.0, 0.0]]))
    sp.add_surface_accumulator(u, u, SurfaceMass(1.0))
    splan = g.compile(sp)
    assert float(one @ pyoracle.apply(splan, one, physical=True)) == pytest.approx(1.0, abs=1e-14)
    # -int du/dn v over the line for u = x^2 (du/dx = 0.5 there), v = 1
    cu = _fit(space, 0, lambda x: x[0] ** 2, 2, 2)
    xu = np.zeros(plan.full_dof_count)
    for box, m in enumerate([(i, j) for j in range(3) for i in range(3)]):
        for dof, c in space.mode_terms(0, m):
            xu[dof] += cu[box] / c

    class NormalDerivative:
        def surface_accumulate(self, q, trial, test_component):
            return -sum(nc * gc for nc, gc in zip(q.normal, trial.gradient))

    np_ = g.AffineProblem(u, g.NonsymmetricOperator())
    np_.add_surface_quadrature(SurfaceQuadrature(0, line, [[1.0, 0.0]]))
    np_.add_surface_accumulator(u, u, NormalDerivative())
    nplan = g.compile(np_)
    assert float(one @ pyoracle.apply(nplan, xu, physical=True)) == pytest.approx(-0.5, abs=1e-13)


def test_nitsche_symmetric():
    """The symmetric Nitsche form is symmetric on the oracle."""
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (2, 2))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, P.Diffusion())
    gp, gw = gauss_legendre_rule(3)
    rng = np.random.default_rng(1)
    for leaf in space.leaves:
        n = rng.standard_normal(2)
        pts = np.stack([rng.uniform(-0.9, 0.9, 3), gp * 0.8], axis=1)
        pr.add_surface_quadrature(SurfaceQuadrature(int(leaf), PointQuadrature(pts, gw), [n]))
    pr.add_surface_accumulator(u, u, Nitsche(10.0))
    plan = g.compile(pr)
    A = np.stack([pyoracle.apply(plan, e, physical=True) for e in np.eye(plan.full_dof_count)], axis=1)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
