"""Device point-cloud kernels (cut-cell rules, embedded-surface pieces) against
the oracle's basis path."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.problem import PointQuadrature, SurfaceQuadrature
from gricoamd.quadrature import gauss_legendre_rule

import problems as P
from test_cloud_host import Nitsche, SurfaceMass

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def random_rules(pr, rng, every=2, lo=1, hi=90):
    D = pr.field.space.dim
    for leaf in pr.field.space.leaves[::every]:
        n = int(rng.integers(lo, hi))
        pr.add_cell_quadrature(int(leaf), PointQuadrature(rng.uniform(-1, 1, size=(n, D)), rng.uniform(0.01, 0.2, n)))


def random_surfaces(pr, rng, every=1, n=7):
    D = pr.field.space.dim
    for leaf in pr.field.space.leaves[::every]:
        pts = rng.uniform(-1, 1, size=(n, D))
        pr.add_surface_quadrature(SurfaceQuadrature(int(leaf), PointQuadrature(pts, rng.uniform(0.05, 0.3, n)),
                                                    rng.standard_normal((n, D))))


def cases():
    rng = np.random.default_rng(7)
    out = {}
    pr = P.poisson(3, 2, 2, operator=P.MassDiffusion(0.7, 1.3, b=(0.2, -0.1, 0.3)), rhs=False)
    random_rules(pr, rng)
    out["3d_cut_general"] = pr
    pr = P.poisson(3, 2, 3, kappa=P.kappa_c2, rhs=False)
    random_rules(pr, rng, every=3, hi=150)
    random_surfaces(pr, rng)
    pr.add_surface_accumulator(pr.field, pr.field, Nitsche(8.0))
    out["3d_cut_nitsche"] = pr
    pr = P.hanging_problem(2)
    random_rules(pr, rng)
    random_surfaces(pr, rng, n=4)
    pr.add_surface_accumulator(pr.field, pr.field, SurfaceMass(2.0))
    out["2d_hanging_cut"] = pr
    pr = P.poisson(2, 3, 5, rhs=False, dirichlet=False, continuity="dg", operator=P.MassDiffusion(1.0, 1.0))
    random_rules(pr, rng, every=1, lo=60, hi=140)
    out["2d_dg_p5_manypoints"] = pr
    return out


CASES = None


@pytest.fixture(scope="module", params=["3d_cut_general", "3d_cut_nitsche", "2d_hanging_cut", "2d_dg_p5_manypoints"])
def case(request, gpu):
    global CASES
    if CASES is None:
        CASES = {k: g.compile(v) for k, v in cases().items()}
    return request.param, CASES[request.param]


def test_apply(case):
    name, plan = case
    assert plan.cloud_buckets or plan.surface_buckets
    x = P.rng_vector(plan.full_dof_count, 3)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) <= TOL, name
    assert rel(g.apply_physical(plan, x), pyoracle.apply(plan, x, physical=True)) <= TOL, name


def test_reduced_apply(case):
    name, plan = case
    xr = P.rng_vector(plan.reduced_dof_count, 4)
    assert rel(g.reduced_apply(plan, xr), pyoracle.reduced_apply(plan, xr)) <= TOL, name


def test_diagonal(case):
    name, plan = case
    if any(b.multi_indices.size for b in plan.all_buckets):
        pytest.skip("multi-term expansions: the local diagonal is not the operator diagonal")
    assert rel(g.diagonal_physical(plan), pyoracle.diagonal_physical(plan)) <= TOL, name
    if plan.reduced_dof_count <= 700:
        assert rel(g.reduced_diagonal(plan), pyoracle.reduced_matrix(plan).diagonal()) <= TOL, name


def test_surface_golden(gpu):
    domain = g.Domain((0.0,), (1.0,), (1,))
    space = g.HpSpace(domain, g.SpaceOptions(degree=g.UniformDegree(1), continuity="dg"))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.NonsymmetricOperator())
    pr.add_surface_quadrature(SurfaceQuadrature(0, PointQuadrature([[0.0]], [1.0]), [[1.0]]))
    pr.add_surface_accumulator(u, u, SurfaceMass(5.0))
    plan = g.compile(pr)
    assert np.allclose(g.apply(plan, np.array([0.25, -0.5])), [-0.3125, -0.3125], atol=1e-15)
