"""Adaptive hp corner problem (anisotropic dyadic refinement toward the origin,
hanging nodes, per-cell degree rising with the distance from the corner,
u = r^(2/3)): device operator, FP32-V-cycle GMG-CG and the verification L2
error against the oracle's dense FP64 solve of the same discrete system."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle
from gricoamd.multigrid import GeometricMultigridPreconditioner

import problems as P

pytestmark = pytest.mark.gpu


def _oracle_state(plan):
    """Dense FP64 solve of the reduced system on the CPU (oracle matrix)."""
    A = pyoracle.reduced_matrix(plan)
    shift_action = pyoracle.apply(plan, plan.shift, physical=True)
    b = (plan.rhs_physical() - shift_action)[plan.solve_dofs]
    x = np.linalg.solve(A, b)
    full = plan.shift.copy()
    full[plan.solve_dofs] = x
    return g.State(plan, full)


@pytest.mark.parametrize("base,levels,pmax", [(2, 2, 4), (4, 3, 5)])
def test_adaptive_apply_matches_oracle(gpu, base, levels, pmax):
    plan = g.compile(P.adaptive_hp(base, levels, pmax))
    degrees = sorted({b.degrees for b in plan.buckets})
    assert len(degrees) >= 3  # several uneven-degree buckets
    x = P.rng_vector(plan.full_dof_count, 4)
    y = g.apply(plan, x)
    ref = pyoracle.apply(plan, x)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_adaptive_fp32_gmg_cg_matches_oracle_solution(gpu):
    """FP64 CG with the FP32 Chebyshev V-cycle reaches the oracle's solution;
    the relative L2 error against u = r^(2/3) agrees with the oracle's to 1%
    (in fact to solver tolerance)."""
    problem = P.adaptive_hp(4, 3, 5)
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", precision="fp32")
    st = g.solve(problem, solver=g.CGSolver(preconditioner=pol), relative_tolerance=1e-10)
    assert st.info.converged == 1
    ref = _oracle_state(st.plan)
    scale = np.abs(ref.coefficients).max()
    assert np.abs(st.coefficients - ref.coefficients).max() <= 1e-7 * scale
    e = g.relative_l2_error(st, P.corner_exact)
    e_ref = g.relative_l2_error(ref, P.corner_exact)
    assert abs(e - e_ref) <= 0.01 * e_ref
    assert e < 0.05
    jac = g.solve(g.compile(problem), solver=g.CGSolver(preconditioner=g.JacobiPreconditioner()),
                  relative_tolerance=1e-10)
    assert st.info.iterations < jac.info.iterations


def test_adaptive_error_decreases_with_refinement(gpu):
    """More corner refinement and higher far-field degree lower the error of the
    singular solution (a sanity check of the hp discretization)."""
    errs = []
    for levels, pmax in ((1, 3), (3, 5)):
        st = g.solve(P.adaptive_hp(4, levels, pmax), relative_tolerance=1e-10)
        errs.append(g.relative_l2_error(st, P.corner_exact))
    assert errs[1] < errs[0]
