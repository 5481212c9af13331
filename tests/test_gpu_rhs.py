"""Device right-hand side, reduced right-hand side and verification sums
against the oracle's dense-basis restatement (src/assembly.jl:488-507,
917-933; src/verification.jl:201-230)."""

import numpy as np
import pytest

import gricoamd as g
import pyoracle

import problems as P

pytestmark = pytest.mark.gpu


class GradLoad:
    """Value and gradient channels: v = 1 + x y, g = (sin x, z, x - y)."""

    def cell_rhs_accumulate(self, q, test_component):
        x, y, z = q.point
        return g.TestChannels(1.0 + x * y, (np.sin(x), z, x - y))


class GradLoad2D:
    def cell_rhs_accumulate(self, q, test_component):
        x, y = q.point
        return g.TestChannels(np.cos(x) * y, (x * x, 1.0 + 0.0 * y))


def _grad_problem():
    pr = P.poisson(3, None, 3, refine_depth=1, kappa=P.kappa_c2, rhs=False)
    pr.add_cell_accumulator(pr.field, GradLoad())
    return pr


def _grad_problem_2d():
    pr = P.poisson(2, None, 4, refine_depth=2, rhs=False)
    pr.add_cell_accumulator(pr.field, GradLoad2D())
    return pr


CASES = {
    "3d_p4_sine": lambda: P.poisson(3, None, 4, refine_depth=2, kappa=P.kappa_c2),
    "3d_p3_gradient_load": _grad_problem,
    "2d_p4_gradient_load": _grad_problem_2d,
    "adaptive_hanging": lambda: P.adaptive_hp(4, 3, 5),
    "adaptive_p8": lambda: P.adaptive_hp(2, 2, 8),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_rhs_matches_oracle(gpu, name):
    plan = g.compile(CASES[name]())
    loads, channels = plan.lowered_loads()
    y = g.ops.rhs_physical(plan)
    ref = pyoracle.rhs_from_loads(plan, loads, channels)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    # and the host (sum-factorized) right-hand side
    assert np.abs(y - plan.rhs_physical()).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["adaptive_hanging", "3d_p3_gradient_load"])
def test_device_reduced_rhs_matches_oracle(gpu, name):
    """b_red = P^T (rhs - A_phys shift); the adaptive case has nonzero Dirichlet data."""
    plan = g.compile(CASES[name]())
    b = g.ops.reduced_rhs(plan)
    r = pyoracle.rhs_from_loads(plan, *plan.lowered_loads())
    ref = (r - pyoracle.apply(plan, plan.shift, physical=True))[plan.solve_dofs]
    assert np.abs(b - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["adaptive_hanging", "3d_p4_sine", "adaptive_p8"])
def test_device_l2_matches_oracle(gpu, name):
    plan = g.compile(CASES[name]())
    u = P.rng_vector(plan.full_dof_count, 9)
    exact = P.corner_exact
    s = g.ops.l2_sums(plan, u, exact)
    ref = pyoracle.l2_sums(plan, u, plan.point_weights(), plan.point_values(exact))
    assert s[0] == pytest.approx(ref[0], rel=1e-12)
    assert s[1] == pytest.approx(ref[1], rel=1e-12)


def test_device_resident_solve_and_error(gpu):
    """solve() on a tensor-product plan uses the device right-hand side; the
    device L2 error of the solution equals the host one."""
    pr = P.adaptive_hp(4, 3, 5)
    st = g.solve(pr, relative_tolerance=1e-10)
    e_dev = g.device_relative_l2_error(st, P.corner_exact)
    e_host = g.relative_l2_error(st, P.corner_exact)
    assert e_dev == pytest.approx(e_host, rel=1e-10)
    assert e_dev < 0.01
