"""Device operator parity against the oracle (FP64, 1e-12 relative).

Each case builds a plan on the host, applies it through libgrico_mf.so on the
HIP device and compares with the oracle's restatement of the reference
algorithm (oracle/oracle.c driven by oracle/pyoracle.py).
"""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


CASES = {
    "1d_p3": lambda: P.poisson(1, 4, 3),
    "2d_p2_c1": lambda: P.poisson(2, 16, 2),
    "2d_p5": lambda: P.poisson(2, 4, 5),
    "3d_p1": lambda: P.poisson(3, 3, 1),
    "3d_p2": lambda: P.poisson(3, 4, 2),
    "3d_p3": lambda: P.poisson(3, 3, 3),
    "3d_p4_kappa": lambda: P.poisson(3, 3, 4, kappa=P.kappa_c2),
    "3d_p4_tree": lambda: P.poisson(3, 1, 4, kappa=P.kappa_c2, refine_depth=2),
    "3d_p6": lambda: P.poisson(3, 2, 6),
    "3d_p8": lambda: P.poisson(3, 2, 8),
    "2d_general": lambda: P.poisson(2, 5, 3, operator=P.MassDiffusion(2.0, 0.5, (0.3, -0.7))),
    "3d_general": lambda: P.poisson(3, 3, 3, operator=P.MassDiffusion(1.5, 1.0, (0.2, 0.1, -0.4))),
    "3d_trunk": lambda: P.poisson(3, 3, 4, basis=g.TrunkBasis()),
    "2d_dg": lambda: P.poisson(2, 4, 2, dirichlet=False, continuity="dg"),
    "2d_hanging": lambda: P.hanging_problem(2),
    "3d_hanging": lambda: P.hanging_problem(3, dim=3),
    "2d_corner": lambda: P.corner_refined(2, 4, 2, 3),
    "3d_corner": lambda: P.corner_refined(3, 2, 1, 2),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, gpu):
    print(f"[case {request.param}]", flush=True)
    return request.param, g.compile(CASES[request.param]())


def test_apply_matches_oracle(case):
    name, plan = case
    x = P.rng_vector(plan.full_dof_count, seed=11)
    y = g.apply(plan, x)
    y_ref = pyoracle.apply(plan, x)
    assert rel(y, y_ref) <= TOL, name


def test_apply_deterministic_vector(case):
    """The reference's seed-free sample vector (src/diagnostics.jl:109-121)."""
    name, plan = case
    i = np.arange(1, plan.full_dof_count + 1, dtype=np.float64)
    x = np.sin(0.019 * i) + 0.1 * np.cos(0.043 * i)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) <= TOL, name


def test_apply_physical_matches_oracle(case):
    name, plan = case
    x = P.rng_vector(plan.full_dof_count, seed=12)
    assert rel(g.apply_physical(plan, x), pyoracle.apply(plan, x, physical=True)) <= TOL, name


def test_reduced_apply_matches_oracle(case):
    name, plan = case
    x = P.rng_vector(plan.reduced_dof_count, seed=13)
    full = np.zeros(plan.full_dof_count)
    full[plan.solve_dofs] = x
    ref = pyoracle.apply(plan, full, physical=True)[plan.solve_dofs]
    assert rel(g.reduced_apply(plan, x), ref) <= TOL, name


def test_expand_project_roundtrip(case):
    name, plan = case
    x = P.rng_vector(plan.reduced_dof_count, seed=14)
    full = g.expand_reduced(plan, x, include_shift=False)
    assert np.array_equal(full[plan.solve_dofs], x)
    assert np.array_equal(g.project_reduced(plan, full), x)


def test_diagonal_matches_oracle(case):
    name, plan = case
    if any(b.multi_indices.size for b in plan.buckets):
        pytest.skip("multi-term expansions: the local diagonal kernel is not the operator diagonal")
    d = g.diagonal_physical(plan)
    assert rel(d, pyoracle.diagonal_physical(plan)) <= TOL, name


def test_reduced_diagonal_equals_probing(gpu):
    """Diagonal equals the diagonal of the probed reduced operator
    (test/matrix_free.jl:514-535 _reference_reduced_diagonal)."""
    plan = g.compile(P.poisson(3, 2, 3, kappa=P.kappa_c2))
    n = plan.reduced_dof_count
    probe = np.zeros(n)
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        probe[j] = pyoracle.reduced_apply(plan, e)[j]
    assert rel(g.reduced_diagonal(plan), probe) <= TOL


def test_reduced_diagonal_equals_probing_all_cases(case):
    """Exact diagonal of P^T A P on every case, hanging / p-mismatch meshes
    included (the reference probes there, src/solve.jl:722-738)."""
    name, plan = case
    n = plan.reduced_dof_count
    if n > 700:
        pytest.skip("probing is quadratic; covered by the smaller cases")
    probe = pyoracle.reduced_matrix(plan).diagonal()
    assert rel(g.reduced_diagonal(plan), probe) <= TOL, name


def test_alias_rejected(gpu):
    plan = g.compile(P.poisson(2, 2, 2))
    x = gpu.zeros(plan.full_dof_count, dtype=gpu.float64, device="cuda")
    with pytest.raises(ValueError):
        g.apply_into(plan, x, x)


def test_small_batches_take_cell_kernel(gpu, monkeypatch):
    """Below the brick threshold the fast path runs the one-launch per-cell
    kernel; results are unchanged."""
    monkeypatch.setenv("GMF_BRICK_MIN_CELLS", "32768")
    plan = g.compile(P.poisson(3, 4, 4, kappa=P.kappa_c2))
    assert "path=fast-cell" in plan.device().describe()
    x = P.rng_vector(plan.full_dof_count, seed=21)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) <= TOL
