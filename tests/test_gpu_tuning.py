"""Plan-time autotuner of the bucketed kernels (gmf_plan_tuning): every bucket
reports the kernel family and launch shape it runs with, the kept shape is the
fastest one timed, and every shape the tuner can pick computes the oracle's
operator (each forced in turn)."""

import contextlib
import os

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd import ops
from gricoamd.problem import PointQuadrature

import problems as P

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@contextlib.contextmanager
def env(**kv):
    old = {k: os.environ.get(k) for k in kv}
    try:
        for k, v in kv.items():
            os.environ[k] = str(v)
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def device_plan(problem, **kv):
    """Compile and create the device plan (where the tuner runs) under env kv."""
    plan = g.compile(problem)
    with env(**kv):
        dev = ops._ws(plan)
    return plan, dev


def cut_problem(seed=3, big=2600):
    """3D p=2 cells where every other cell carries its own rule, a few of them
    with thousands of points (several work units at every candidate size)."""
    rng = np.random.default_rng(seed)
    pr = P.poisson(3, 3, 2, operator=P.MassDiffusion(0.7, 1.3), rhs=False)
    for i, leaf in enumerate(pr.field.space.leaves[::2]):
        n = big if i % 4 == 0 else int(rng.integers(5, 90))
        pr.add_cell_quadrature(int(leaf), PointQuadrature(rng.uniform(-1, 1, size=(n, 3)), rng.uniform(0.01, 0.2, n)))
    return pr


def check_report(dev, kinds):
    rep = dev.tuning()
    assert len(rep) == dev.bucket_count
    for t in rep:
        assert t["kernel"] in kinds, t
        if t["kernel"] in ("empty", "face-tensor"):
            continue
        assert t["candidates"] >= 1, t
        assert t["best_us"] > 0.0, t
        if t["default_us"] > 0.0:
            assert t["best_us"] <= t["default_us"] * (1.0 + 1e-6), t
    return rep


def test_report_of_every_bucket_family(gpu):
    _, dev = device_plan(P.adaptive_hp(base=2, levels=2, pmax=6))
    rep = check_report(dev, {"cell", "generic", "brick"})
    assert {"generic"} <= {t["kernel"] for t in rep}
    for t in rep:
        if t["kernel"] == "generic":
            assert t["param0"] in (64, 128, 256)
        if t["kernel"] == "cell":
            assert t["param0"] in (64, 128, 256)
    _, dev = device_plan(cut_problem(), GMF_CLOUD_MATRIX=0)
    rep = check_report(dev, {"cell", "generic", "brick", "point-cloud"})
    cloud = [t for t in rep if t["kernel"] == "point-cloud"]
    assert cloud and all(t["param0"] in (256, 512, 1024, 2048) and t["param1"] in (32, 64) for t in cloud)
    assert any(t["candidates"] >= 2 for t in cloud)
    # timed choice between the point shapes and local matrices
    _, dev = device_plan(cut_problem())
    rep = check_report(dev, {"cell", "generic", "brick", "point-cloud", "point-cloud-matrix"})
    for t in rep:
        if t["kernel"] == "point-cloud-matrix":
            assert t["param0"] == 27 and t["best_us"] <= t["default_us"] * (1.0 + 1e-6)


def test_brick_or_cell_choice(gpu):
    # 16^3 cells at p = 1 form 512 bricks; without GMF_KEEP_BRICKS the tuner
    # may trade them for the per-cell kernel. Either way the apply is the oracle's.
    pr = P.poisson(3, 16, 1, kappa=P.kappa_c2, rhs=False)
    plan, dev = device_plan(pr, GMF_KEEP_BRICKS=0, GMF_BRICK_MIN_CELLS=4096)
    rep = check_report(dev, {"cell", "brick"})
    assert rep[0]["candidates"] == 4  # bricks + three per-cell workgroup sizes
    x = P.rng_vector(plan.full_dof_count, seed=5)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) < TOL


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_forced_cell_workgroups_match_oracle(gpu, threads):
    pr = P.poisson(3, 3, 4, kappa=P.kappa_c2, rhs=False)
    plan, dev = device_plan(pr, GMF_CELL_THREADS=threads, GMF_BRICK_MIN_CELLS=1 << 30)
    rep = dev.tuning()
    assert rep[0]["kernel"] == "cell" and rep[0]["param0"] == threads
    x = P.rng_vector(plan.full_dof_count, seed=1)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) < TOL


@pytest.mark.parametrize("points", [256, 2048])
def test_forced_cloud_units_match_oracle(gpu, points):
    plan, dev = device_plan(cut_problem(), GMF_CLOUD_UNIT_POINTS=points, GMF_CLOUD_MATRIX=0)
    rep = dev.tuning()
    assert all(t["param0"] == points for t in rep if t["kernel"] == "point-cloud")
    x = P.rng_vector(plan.full_dof_count, seed=2)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) < TOL
    xr = P.rng_vector(plan.reduced_dof_count, seed=3)
    assert rel(g.reduced_apply(plan, xr), pyoracle.reduced_apply(plan, xr)) < TOL


def test_forced_cloud_matrices_match_oracle(gpu):
    # every point-cloud batch through its local matrices (built from the point
    # data by the point-cloud arithmetic): full, reduced and block applies
    plan, dev = device_plan(cut_problem(), GMF_CLOUD_MATRIX=1)
    rep = dev.tuning()
    assert any(t["kernel"] == "point-cloud-matrix" for t in rep)
    x = P.rng_vector(plan.full_dof_count, seed=6)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) < TOL
    xr = P.rng_vector(plan.reduced_dof_count, seed=7)
    yr = g.reduced_apply(plan, xr)
    assert rel(yr, pyoracle.reduced_apply(plan, xr)) < TOL
    X = np.stack([P.rng_vector(plan.reduced_dof_count, seed=10 + j) for j in range(11)], axis=1)
    Y = g.reduced_apply_block(plan, X).cpu().numpy()
    ref = np.stack([pyoracle.reduced_apply(plan, X[:, j]) for j in range(X.shape[1])], axis=1)
    assert rel(Y, ref) < TOL
    # the block apply gives each column the one-column result
    assert np.abs(Y[:, 0] - np.asarray(g.reduced_apply(plan, X[:, 0]))).max() <= 1e-14 * np.abs(Y).max()


def test_tuner_off_keeps_defaults(gpu):
    plan, dev = device_plan(cut_problem(), GMF_TUNE=0)
    for t in dev.tuning():
        assert t["candidates"] == 0
        if t["kernel"] == "point-cloud":
            assert (t["param0"], t["param1"]) == (512, 64)
    x = P.rng_vector(plan.full_dof_count, seed=4)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) < TOL


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_grouped_bricks_on_ball_match_oracle(gpu, degree):
    # the tensor cells of a finite-cell ball leave gaps in their 2x2x2 groups
    # (cut and outside cells are not in the batch): the host grouping with
    # empty positions puts them on the brick kernel; full, reduced and block
    # applies against the oracle
    pr = P.fcm_ball(n=8, degree=degree, depth=1, n_theta=12, n_phi=24)
    plan, dev = device_plan(pr, GMF_KEEP_BRICKS=1, GMF_BRICK_MIN_CELLS=8)
    tensor = [i for i, b in enumerate(plan.buckets) if not getattr(b, "cloud", False)]
    assert tensor and plan.buckets[tensor[0]].brick_cells is not None
    assert (plan.buckets[tensor[0]].brick_cells < 0).any()
    assert dev.tuning()[tensor[0]]["kernel"] == "brick"
    x = P.rng_vector(plan.full_dof_count, seed=21)
    assert rel(g.apply(plan, x), pyoracle.apply(plan, x)) < TOL
    xr = P.rng_vector(plan.reduced_dof_count, seed=22)
    assert rel(g.reduced_apply(plan, xr), pyoracle.reduced_apply(plan, xr)) < TOL
    X = np.stack([P.rng_vector(plan.reduced_dof_count, seed=30 + j) for j in range(5)], axis=1)
    Y = g.reduced_apply_block(plan, X).cpu().numpy()
    ref = np.stack([pyoracle.reduced_apply(plan, X[:, j]) for j in range(X.shape[1])], axis=1)
    assert rel(Y, ref) < TOL
