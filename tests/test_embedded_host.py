"""Finite-cell geometry on the host: leaf classification, cut-cell candidate
rules (src/regions.jl:304-329, 477-517, 736-817) and the sphere surface rule,
checked against the closed-form volume and area of a ball."""

import numpy as np
import pytest

import gricoamd as g
from gricoamd.embedded import (CUT, INSIDE, OUTSIDE, ImplicitRegion, add_cut_cell_quadratures, classify_leaves,
                               cut_cell_rules, physical_space, sphere_surface_quadratures)


def ball(r=1.0):
    return ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - r * r, subdivision_depth=3)


def unit_ball_domain(n=12):
    return g.Domain((-1.2, -1.2, -1.2), (2.4, 2.4, 2.4), (n, n, n))


def test_classification_counts():
    dom = unit_ball_domain(12)
    cls = classify_leaves(dom, ball(), dom.active_leaves())
    assert set(np.unique(cls)) <= {INSIDE, OUTSIDE, CUT}
    # centre cell inside, corner cell outside
    lo = dom.cell_lowers(dom.active_leaves())
    centre = np.argmin(np.linalg.norm(lo + 0.1, axis=1))
    corner = np.argmin(np.linalg.norm(lo + 1.2, axis=1))
    assert cls[centre] == INSIDE and cls[corner] == OUTSIDE
    assert (cls == CUT).sum() > 0


@pytest.mark.parametrize("depth,tol", [(1, 3e-2), (3, 6e-3)])
def test_volume_converges(depth, tol):
    dom = unit_ball_domain(12)
    region = ImplicitRegion(ball().classifier, subdivision_depth=depth)
    leaves = dom.active_leaves()
    cls = classify_leaves(dom, region, leaves)
    h = 2.4 / 12
    vol = (cls == INSIDE).sum() * h ** 3
    rules = cut_cell_rules(dom, region, leaves[cls == CUT], (3, 3, 3))
    vol += sum(r.weights.sum() for r in rules.values()) * (h / 2) ** 3
    assert abs(vol - 4.0 / 3.0 * np.pi) <= tol * 4.0 / 3.0 * np.pi
    for r in rules.values():
        assert np.all(np.abs(r.points) <= 1.0) and np.all(r.weights > 0)


def test_sphere_surface_area_and_normals():
    dom = unit_ball_domain(12)
    space, cut = physical_space(dom, ball(), g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    for sq in sphere_surface_quadratures(space, (0, 0, 0), 1.0, n_theta=32, n_phi=64):
        pr.add_surface_quadrature(sq)
    plan = g.compile(pr)
    area = sum(b.weights.sum() for b in plan.surface_buckets)
    assert area == pytest.approx(4 * np.pi, rel=1e-12)
    for b in plan.surface_buckets:
        x = np.stack(b.phys_points, axis=1)
        assert np.allclose(np.linalg.norm(x, axis=1), 1.0, atol=1e-12)
        assert np.allclose(b.normals, x, atol=1e-12)


def test_cut_rules_attach():
    dom = unit_ball_domain(8)
    region = ImplicitRegion(ball().classifier, subdivision_depth=2)
    space, cut = physical_space(dom, region, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(2)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    n = add_cut_cell_quadratures(pr, region, cut)
    assert n > 0 and len(pr.cell_quadratures) == len(cut)
    plan = g.compile(pr)
    assert sum(b.cell_count for b in plan.cloud_buckets) == len(cut)
