"""Non-negative moment fitting of cut-cell quadratures (src/regions.jl:519-708).

The oracle (oracle/moment_oracle.py) restates the reference step for step; it
is pinned here by what the mathematics fixes:
- its Legendre values are numpy's Legendre series;
- its NNLS equals scipy's Lawson-Hanson solver wherever the problem has a
  unique solution (full column rank), including one with active bounds;
- a fitted rule uses a subset of the candidate points with non-negative
  weights and reproduces the candidate's moments to the tolerance;
- on tiny 1D and 2D cuts, brute force over single removals confirms that the
  elimination stopped only where no removal keeps an exact non-negative fit.

The reference's own goldens (test/embedded_geometry.jl:49-58): the half cube
at depth 1 has weight 4, a sliver keeps at least one point of positive
weight. The product (gricoamd.embedded.moment_fit) must pick the oracle's
points with the oracle's weights on sphere cut cells.
"""

import numpy as np
import pytest
from scipy.optimize import nnls as scipy_nnls

import gricoamd as g
import moment_oracle as mo
from gricoamd import embedded as em

TOL = mo.tolerance()


def test_legendre_values_match_numpy():
    x = np.linspace(-1.0, 1.0, 9)
    for deg in range(7):
        ours = np.array([mo.legendre_values(v, deg) for v in x])
        ref = np.stack([np.polynomial.legendre.legval(x, np.eye(deg + 1)[k]) for k in range(deg + 1)], axis=1)
        assert np.abs(ours - ref).max() <= 1e-14
        assert np.abs(em._legendre_table(x, deg) - ours).max() == 0.0


@pytest.mark.parametrize("seed", range(4))
def test_oracle_nnls_matches_scipy_on_unique_problems(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((12, 7))
    b = rng.standard_normal(12)
    w, res = mo.nnls(A, b, TOL)
    ws, rs = scipy_nnls(A, b)
    assert np.abs(w - ws).max() <= 1e-10
    assert res == pytest.approx(rs / max(np.linalg.norm(b), 1.0), rel=1e-10)
    # closed form with active bounds: on an orthonormal A the solution is max(A^T b, 0)
    Q, _ = np.linalg.qr(rng.standard_normal((12, 7)))
    c = rng.standard_normal(12)
    w3, _ = mo.nnls(Q, c, TOL)
    assert np.abs(w3 - np.maximum(Q.T @ c, 0.0)).max() <= 1e-12
    # an exactly representable non-negative combination is recovered
    wt = np.abs(rng.standard_normal(7))
    w2, r2 = mo.nnls(A, A @ wt, TOL)
    assert np.abs(w2 - wt).max() <= 1e-10 and r2 <= 1e-13


def _check_fitted_rule(cand_points, cand_weights, points, weights, shape):
    # subset of the candidate points
    key = {tuple(p): i for i, p in enumerate(np.round(cand_points, 15))}
    assert all(tuple(p) in key for p in np.round(points, 15))
    assert np.all(weights >= 0.0)
    A = mo.moment_matrix(cand_points, shape)
    m = A @ cand_weights
    fitted = mo.moment_matrix(points, shape) @ weights
    assert np.linalg.norm(fitted - m) / max(np.linalg.norm(m), 1.0) <= TOL


def _exact_fit(A, b):
    """Exact non-negative fit on these columns (the elimination's acceptance test)."""
    w, res = mo.nnls(A, b, TOL)
    return res <= TOL


@pytest.mark.parametrize("dim,shape,depth", [(1, (2,), 3), (1, (3,), 4), (2, (2, 2), 2)])
def test_elimination_stops_where_no_removal_fits(dim, shape, depth):
    dom = g.Domain((0.0,) * dim, (1.0,) * dim, (1,) * dim)
    region = em.ImplicitRegion(lambda x: sum((c - 0.1) ** 2 for c in x) - 0.55, subdivision_depth=depth)
    rules = em.cut_cell_rules(dom, region, np.array([0]), shape)
    cand = rules[0]
    points, weights = mo.moment_fit(cand.points, cand.weights, shape)
    assert len(points) < cand.point_count(), "the cut must be compressible"
    _check_fitted_rule(cand.points, cand.weights, points, weights, shape)
    if len(points) > int(np.prod(shape)):
        A = mo.moment_matrix(points, shape)
        m = mo.moment_matrix(cand.points, shape) @ cand.weights
        for i in range(len(points)):
            assert not _exact_fit(np.delete(A, i, axis=1), m)


def test_reference_goldens():
    """test/embedded_geometry.jl:49-58 on the product path."""
    dom = g.Domain((0.0,), (1.0,), (1,))
    sliver = em.fitted_cut_cell_rules(dom, em.ImplicitRegion(lambda x: x[0] - 1.0e-6, subdivision_depth=0),
                                      np.array([0]), (1,), workers=1)[0]
    assert sliver.point_count() >= 1 and sliver.weights.sum() > 0.0
    dom3 = g.Domain((0.0,) * 3, (1.0,) * 3, (1, 1, 1))
    region = em.ImplicitRegion(lambda x: x[0] - 0.5, subdivision_depth=1)
    cand = em.cut_cell_rules(dom3, region, np.array([0]), (2, 2, 2))[0]
    half = em.fitted_cut_cell_rules(dom3, region, np.array([0]), (2, 2, 2), workers=1)[0]
    assert half.weights.sum() == pytest.approx(4.0, abs=1e-12)
    assert half.point_count() < cand.point_count()
    _check_fitted_rule(cand.points, cand.weights, half.points, half.weights, (2, 2, 2))


def _sphere_cut_rules(shape, count, seed):
    dom = g.Domain((-1.2,) * 3, (2.4,) * 3, (16, 16, 16))
    region = em.ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 1.0, subdivision_depth=2)
    leaves = dom.active_leaves()
    cut = leaves[em.classify_leaves(dom, region, leaves) == em.CUT]
    sample = np.random.default_rng(seed).choice(cut, size=count, replace=False)
    return em.cut_cell_rules(dom, region, sample, shape)


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 4, 4)])
def test_product_matches_oracle_on_sphere_cells(shape):
    rules = _sphere_cut_rules(shape, 4, 1)
    fitted = 0
    for leaf, cand in rules.items():
        p1, w1 = em.moment_fit(cand.points, cand.weights, shape)
        p2, w2 = mo.moment_fit(cand.points, cand.weights, shape)
        assert np.array_equal(p1, p2)
        assert np.abs(w1 - w2).max() <= 1e-12 * np.abs(w2).max()
        if len(w2) < cand.point_count():
            fitted += 1
            _check_fitted_rule(cand.points, cand.weights, p1, w1, shape)
    assert fitted >= 2


def test_fitted_rules_in_worker_processes_equal_serial():
    dom = g.Domain((-1.2,) * 3, (2.4,) * 3, (8, 8, 8))
    region = em.ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 1.0, subdivision_depth=1)
    leaves = dom.active_leaves()
    cut = leaves[em.classify_leaves(dom, region, leaves) == em.CUT][:70]
    serial = em.fitted_cut_cell_rules(dom, region, cut, (3, 3, 3), workers=1)
    pooled = em.fitted_cut_cell_rules(dom, region, cut, (3, 3, 3), workers=2)
    assert list(serial) == list(pooled)
    for k in serial:
        assert np.array_equal(serial[k].points, pooled[k].points)
        assert np.array_equal(serial[k].weights, pooled[k].weights)
