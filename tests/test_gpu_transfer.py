"""State transfer across an hp transition solved on the device.

Reference: src/transfer.jl:892-1000 (transfer_state with a linear_solve policy:
the global L2 projection of the old field on the target space, the
reference's variational path 431-437). The device Jacobi-CG solve of the projection is
compared with the oracle's dense FP64 solve of the same discrete system, and a
target space that contains the source space must reproduce the old field at
arbitrary points (nested spaces).
"""

import numpy as np
import pytest

import gricoamd as g
from gricoamd import adaptivity as ad

import pyoracle
from test_adaptivity_host import dense_projection, hanging_space, random_points, state_on

pytestmark = pytest.mark.gpu

RTOL = 1e-13
JACOBI_CG = g.CGSolver(preconditioner=g.JacobiPreconditioner())


def _bound(plan):
    """Coefficient error allowed by the CG stopping rule ||r|| <= RTOL ||b||:
    RTOL times the condition number of the Jacobi-scaled mass matrix (up to
    4.8e5 for the 3D hanging p = 2 space in the integrated Legendre basis)."""
    A = pyoracle.reduced_matrix(plan)
    d = 1.0 / np.sqrt(np.diag(A))
    ev = np.linalg.eigvalsh(d[:, None] * A * d[None, :])
    return RTOL * ev[-1] / ev[0]


def _refined_plan(space):
    plan = ad.AdaptivityPlan(space)
    leaves = [int(l) for l in space.leaves]
    plan.request_h_refinement(leaves[0], space.dim - 1)
    plan.request_h_refinement(leaves[-1], 0)
    plan.request_p_refinement(leaves[len(leaves) // 2], increment=1)
    return plan


@pytest.mark.parametrize("dim,degree,continuity", [(2, 2, "cg"), (2, 3, "dg"), (3, 2, "cg")])
def test_device_transfer_refinement(gpu, dim, degree, continuity):
    space = hanging_space(degree=degree, continuity=continuity, dim=dim)
    rng = np.random.default_rng(17 + dim)
    state = state_on(space, rng.standard_normal(space.scalar_dof_count()))
    t = ad.transition(_refined_plan(space))
    print(f"transfer refinement dim={dim} p={degree} {continuity}", flush=True)
    new = ad.transfer_state(t, state, linear_solve=JACOBI_CG, relative_tolerance=RTOL)
    ref = dense_projection(t, state)
    scale = np.max(np.abs(ref.coefficients))
    tol = _bound(ref.plan) * scale
    assert np.max(np.abs(new.coefficients - ref.coefficients)) <= tol
    pts = random_points(rng, 300, dim)
    np.testing.assert_allclose(ad.evaluate(new, pts), ad.evaluate(state, pts), rtol=0, atol=tol)


def test_device_transfer_p_derefinement(gpu):
    space = hanging_space(degree=3, continuity="cg", dim=2)
    rng = np.random.default_rng(23)
    state = state_on(space, rng.standard_normal(space.scalar_dof_count()))
    plan = ad.AdaptivityPlan(space)
    for l in space.leaves:
        plan.request_p_derefinement(int(l), decrement=2)
    t = ad.transition(plan)
    print("transfer p-derefinement 3 -> 1", flush=True)
    new = ad.transfer_state(t, state, linear_solve=JACOBI_CG, relative_tolerance=RTOL)
    ref = dense_projection(t, state)
    scale = np.max(np.abs(ref.coefficients))
    assert np.max(np.abs(new.coefficients - ref.coefficients)) <= _bound(ref.plan) * scale


@pytest.mark.parametrize("continuity", ["cg", "dg"])
def test_device_transfer_h_derefinement(gpu, continuity):
    """Merged cells carry composite rules and run on the point-cloud kernel."""
    from test_adaptivity_host import _derefined_plan

    space = hanging_space(degree=2, continuity=continuity, dim=2)
    rng = np.random.default_rng(41)
    state = state_on(space, rng.standard_normal(space.scalar_dof_count()))
    plan, _ = _derefined_plan(space)
    t = ad.transition(plan)
    print(f"transfer h-derefinement {continuity}", flush=True)
    new = ad.transfer_state(t, state, linear_solve=JACOBI_CG, relative_tolerance=RTOL)
    ref = dense_projection(t, state)
    assert new.plan.cloud_buckets, "merged cells must use composite rules"
    scale = np.max(np.abs(ref.coefficients))
    assert np.max(np.abs(new.coefficients - ref.coefficients)) <= _bound(ref.plan) * scale
