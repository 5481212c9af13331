"""gmf_reduced_apply_noalloc (the allocation-free reduced apply) against
gmf_reduced_apply on the same plans: two calls in a row on brick plans (the
second smaller, sharing the device scratch of the first), then a 2D plan on
the generic kernel, each within 1e-13 of the workspace entry point (the brick
kernels add shared lattice points with LDS atomics, so the two calls may
differ in the last bits)."""

import ctypes

import numpy as np
import pytest

import gricoamd as g
from gricoamd import _devabi, _native, ops

import problems as P

pytestmark = pytest.mark.gpu


def _apply_noalloc(plan, x, torch):
    lib = _native.device()
    lib.gmf_reduced_apply_noalloc.argtypes = [ctypes.c_void_p] * 4
    lib.gmf_reduced_apply_noalloc.restype = ctypes.c_int
    dev = ops._ws(plan)
    y = torch.full_like(x, float("nan"))
    stream = torch.cuda.current_stream().cuda_stream
    _devabi.check(lib.gmf_reduced_apply_noalloc(dev.handle, ctypes.c_void_p(x.data_ptr()),
                                                ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_workspace_apply(gpu, order):
    plans = [g.compile(P.lognormal_poisson(depth=3, degree=3, seed=1)),   # 8^3 cells, bricks
             g.compile(P.lognormal_poisson(depth=2, degree=4, seed=2)),   # 4^3 cells, smaller input
             g.compile(P.poisson(2, 6, 3, kappa=lambda p: 1.0 + p[0] * p[1]))]  # 2D, generic kernel
    assert "path=brick" in plans[0].device().describe()
    assert "path=brick" in plans[1].device().describe()
    assert plans[1].reduced_dof_count < plans[0].reduced_dof_count
    for i, plan in enumerate(plans):
        x = gpu.as_tensor(P.rng_vector(plan.reduced_dof_count, 40 + i), device="cuda")
        ref = ops.reduced_apply(plan, x)
        y = _apply_noalloc(plan, x, gpu)
        scale = float(ref.abs().max())
        assert bool(gpu.isfinite(y).all())
        assert float((y - ref).abs().max()) <= 1e-13 * scale
