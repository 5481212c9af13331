"""Affine Dirichlet relations, mean-value rows and the general reduced map on
the host (src/plans.jl:750-805, 884-972; src/assembly.jl:112-314), pinned by
closed forms: rank-one and degenerate trace Gram systems with known answers,
consistency of random rank-deficient systems, the mean row of the constant
and linear functions, and the map reproducing every constraint row."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.plan import _dirichlet_relations

import problems as P


def test_relations_full_rank_fix_every_dof():
    rng = np.random.default_rng(0)
    B = rng.standard_normal((5, 5))
    G = B @ B.T + 5 * np.eye(5)
    r = rng.standard_normal(5)
    dofs = np.array([3, 7, 8, 11, 20])
    fd, fv, rows = _dirichlet_relations(dofs, G, r)
    assert rows == []
    order = np.argsort(fd)
    assert np.array_equal(fd[order], dofs)
    assert np.allclose(fv[order], np.linalg.solve(G, r), rtol=1e-12)


def test_relations_uncontrolled_dof_stays_free():
    """G = diag(1, 0): the second trace dof has no trace; only the first is fixed."""
    fd, fv, rows = _dirichlet_relations(np.array([4, 9]), np.diag([1.0, 0.0]), np.array([2.0, 0.0]))
    assert rows == [] and fd.tolist() == [4] and fv.tolist() == pytest.approx([2.0])


def test_relations_identical_traces_give_one_relation():
    """Two dofs with the same trace: G = [[1, 1], [1, 1]], data 2 -> u_a + u_b = 2."""
    fd, fv, rows = _dirichlet_relations(np.array([4, 9]), np.ones((2, 2)), np.array([2.0, 2.0]))
    assert len(fd) == 0 and len(rows) == 1
    pivot, idx, cf, rhs = rows[0]
    assert {pivot, int(idx[0])} == {4, 9}
    assert cf.tolist() == pytest.approx([-1.0]) and rhs == pytest.approx(2.0)


def test_relations_reject_inconsistent_data():
    with pytest.raises(ValueError, match="inconsistent"):
        _dirichlet_relations(np.array([0, 1]), np.diag([1.0, 0.0]), np.array([2.0, 1.0]))


def test_relations_solve_rank_deficient_systems():
    """Random G = B B^T of rank 4 in 7 dofs with consistent data: every vector
    satisfying the relations (free dofs arbitrary) solves G u = r."""
    rng = np.random.default_rng(1)
    B = rng.standard_normal((7, 4))
    G = B @ B.T
    r = G @ rng.standard_normal(7)
    dofs = np.arange(7) * 3
    fd, fv, rows = _dirichlet_relations(dofs, G, r)
    assert len(fd) + len(rows) == 4
    pos = {int(d): i for i, d in enumerate(dofs)}
    for trial in range(3):
        u = rng.standard_normal(7)
        for d, v in zip(fd.tolist(), fv.tolist()):
            u[pos[d]] = v
        for (p, idx, cf, rhs) in rows:
            u[pos[p]] = rhs + sum(c * u[pos[int(j)]] for j, c in zip(idx, cf))
        assert np.allclose(G @ u, r, atol=1e-10 * np.abs(r).max())


def _neumann(dim=2, cells=4, degree=2, target=0.0):
    pr = P.poisson(dim, cells, degree, dirichlet=False, rhs=False)
    pr.add_cell_accumulator(pr.field, P.CosLoad(dim))
    pr.add_constraint(g.MeanValue(pr.field, target))
    return pr


def test_mean_row_integrates_constant_and_linear():
    """The mean row is the integral of each basis function: against the
    coefficient vector of 1 (vertex dofs 1) it gives the measure, against that
    of x (vertex dofs at their x) the integral of x."""
    plan = g.compile(_neumann(2, 3, 3, target=0.7))
    (pivot, idx, cf, rhs), = plan.mean_rows
    row = np.zeros(plan.full_dof_count)
    row[idx] = cf
    b = plan.buckets[0]
    one = np.zeros(plan.full_dof_count)
    xs = np.zeros(plan.full_dof_count)
    vert = [(0, 0), (1, 0), (0, 1), (1, 1)]
    # vertex modes: box index with every axis index in {0, 1} (0: lower, 1: upper vertex)
    N = b.modes
    for c in range(b.cell_count):
        cx = np.array([p[c].mean() for p in b.phys_points])
        for (i, j) in vert:
            box = i + N[0] * j
            dof = b.local_index[c, box]
            one[dof] = 1.0
            xs[dof] = cx[0] + (0.5 if i else -0.5) * b.sizes[c, 0]
    assert row @ one == pytest.approx(1.0, rel=1e-13)
    assert row @ xs == pytest.approx(0.5, rel=1e-13)
    assert rhs == pytest.approx(0.7 * 1.0)
    assert pivot == min(idx.tolist())


def test_reduced_map_reproduces_constraint_rows():
    """shift + P x satisfies the mean row and injected affine relations for
    any reduced x; the oracle's expansion of the map agrees."""
    plan = g.compile(_neumann(2, 3, 2, target=0.25))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(plan.reduced_dof_count)
    u = plan.expand_host(x)
    (pivot, idx, cf, rhs), = plan.mean_rows
    assert cf @ u[idx] == pytest.approx(rhs, abs=1e-12)
    assert np.allclose(pyoracle.expand(plan, x), u, rtol=0, atol=1e-13)
    # relations among interior dofs, one of them on a dof of the mean row
    free = plan.solve_dofs
    a, b_, c, d = (int(v) for v in free[[5, 9, 14, 20]])
    plan.affine_rows = [(a, np.array([b_, c]), np.array([0.5, -0.25]), 0.3), (d, np.array([a]), np.array([2.0]), -1.0)]
    plan._compile_reduced_map()
    x = rng.standard_normal(plan.reduced_dof_count)
    u = plan.expand_host(x)
    assert u[a] - 0.5 * u[b_] + 0.25 * u[c] == pytest.approx(0.3, abs=1e-12)
    assert u[d] - 2.0 * u[a] == pytest.approx(-1.0, abs=1e-12)
    assert cf @ u[idx] == pytest.approx(rhs, abs=1e-12)
    # P^T is the transpose of P
    y = rng.standard_normal(plan.full_dof_count)
    Px = plan.expand_host(x, include_shift=False)
    assert Px @ y == pytest.approx(x @ plan.project_host(y), rel=1e-12)


def test_mean_value_golden_one_dof():
    """solve(mean_problem) = [2.0] (test/matrix_free.jl:1047-1053): one DG p = 0
    cell whose only dof is the mean pivot; no free dof remains."""
    pr = P.mean_golden_problem()
    plan = g.compile(pr)
    assert plan.reduced_dof_count == 0
    assert plan.expand_host(np.zeros(0)).tolist() == pytest.approx([2.0])
