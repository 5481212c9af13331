"""Operator entry points on device vectors (torch tensors on the HIP device).

Each function calls one C-ABI entry point of libgrico_mf.so; numpy inputs are
copied to the device and results copied back. There is no CPU fallback: a
missing device library or a missing GPU raises.

Reference counterparts: apply / apply! (src/assembly.jl:540-564),
_apply_physical! (603-622), _reduced_apply! (905-915), _expand_reduced! /
_project_reduced! (851-891), _diagonal_physical! / _reduced_diagonal!
(624-675).
"""

import ctypes

import numpy as np

from . import _devabi, _native


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the matrix-free backend requires a HIP device (torch.cuda.is_available() is False)")
    return torch


def to_device(x, n=None):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x.to(device="cuda", dtype=torch.float64).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")
    if n is not None and t.numel() != n:
        raise ValueError(f"vector must have length {n}")
    return t


def _out_like(x, t):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return t
    return t.cpu().numpy()


def _ws(plan):
    dev = plan.device()
    torch = _torch()
    dev.set_stream(torch.cuda.current_stream().cuda_stream)
    return dev


def _call(name, *args):
    _devabi.check(getattr(_native.device(), name)(*args))


def apply(plan, x):
    """Full-layout action with constraint rows (apply(plan, x))."""
    dev = _ws(plan)
    xt = to_device(x, plan.full_dof_count)
    y = _torch().empty_like(xt)
    _call("gmf_apply", dev.ws, xt.data_ptr(), y.data_ptr())
    return _out_like(x, y)


def apply_into(plan, y, x):
    """apply!(y, plan, x) on device tensors (no copies)."""
    dev = _ws(plan)
    if y.data_ptr() == x.data_ptr():
        raise ValueError("result and coefficient vector must not alias")
    _call("gmf_apply", dev.ws, x.data_ptr(), y.data_ptr())
    return y


def apply_async(plan, y, x):
    dev = _ws(plan)
    _call("gmf_apply_async", dev.ws, x.data_ptr(), y.data_ptr())
    return y


def apply_physical(plan, x):
    dev = _ws(plan)
    xt = to_device(x, plan.full_dof_count)
    y = _torch().empty_like(xt)
    _call("gmf_apply_physical", dev.ws, xt.data_ptr(), y.data_ptr())
    return _out_like(x, y)


def reduced_apply(plan, x):
    dev = _ws(plan)
    xt = to_device(x, plan.reduced_dof_count)
    y = _torch().empty_like(xt)
    _call("gmf_reduced_apply", dev.ws, xt.data_ptr(), y.data_ptr())
    return _out_like(x, y)


def reduced_apply_f32(plan, x):
    """FP32 reduced action (the FP32 multigrid level operator); returns FP64
    values of the FP32 result."""
    torch = _torch()
    dev = _ws(plan)
    xt = to_device(x, plan.reduced_dof_count).to(torch.float32)
    y = torch.empty_like(xt)
    _call("gmf_reduced_apply_f32", dev.ws, xt.data_ptr(), y.data_ptr())
    return _out_like(x, y.to(torch.float64))


def rhs_physical(plan, as_numpy=True):
    """Physical right-hand side on the device from the plan's lowered point
    loads: tensor-product cells (src/assembly.jl:488-507), then the point-cloud
    batches' loads: cut cells, embedded surfaces, boundary faces and interfaces
    (src/accumulators.jl:1022-1054, 1230-1274, 1666-1698). Plans with loads on
    batches the device plan does not hold (plan.device_rhs is False) raise."""
    torch = _torch()
    if not plan.device_rhs:
        raise ValueError("this plan's loads are not all on the device path (cell loads without a cell operator "
                         "next to face or surface batches, or faces without an operator); use plan.rhs_physical()")
    dev = _ws(plan)
    loads, channels = plan.lowered_loads()
    y = torch.zeros(plan.full_dof_count, dtype=torch.float64, device="cuda")
    if loads.size:  # plans without tensor-product cells have no tensor loads
        lt = to_device(loads)
        _call("gmf_rhs_physical", dev.ws, lt.data_ptr(), int(channels), y.data_ptr())
    faces = plan.face_loads()
    if faces.size:
        ft = to_device(faces)
        _call("gmf_rhs_faces", dev.ws, ft.data_ptr(), y.data_ptr())
    return y.cpu().numpy() if as_numpy else y


def reduced_rhs(plan, as_numpy=True):
    """b_red = P^T (rhs - A_phys shift) on the device (src/assembly.jl:917-933)."""
    torch = _torch()
    dev = _ws(plan)
    if plan.reduced_dof_count == 0:  # every dof constrained
        return np.zeros(0) if as_numpy else torch.zeros(0, dtype=torch.float64, device="cuda")
    rhs = rhs_physical(plan, as_numpy=False)
    b = torch.empty(plan.reduced_dof_count, dtype=torch.float64, device="cuda")
    _call("gmf_reduced_rhs", dev.ws, rhs.data_ptr(), b.data_ptr())
    return b.cpu().numpy() if as_numpy else b


def l2_sums(plan, u_full, exact):
    """(sum w (u_h - u)^2, sum w u^2) over the points of the plan's tensor and
    cut cells on the device (src/verification.jl:201-230)."""
    dev = _ws(plan)
    ut = to_device(u_full, plan.full_dof_count)
    w = to_device(plan.point_weights())
    e = to_device(plan.point_values(exact))
    out = (ctypes.c_double * 2)()
    _call("gmf_l2_error", dev.ws, ut.data_ptr(), w.data_ptr(), e.data_ptr(), out)
    return float(out[0]), float(out[1])


def reduced_apply_into(plan, y, x):
    dev = _ws(plan)
    _call("gmf_reduced_apply", dev.ws, x.data_ptr(), y.data_ptr())
    return y


def expand_reduced(plan, x, include_shift=True):
    dev = _ws(plan)
    xt = to_device(x, plan.reduced_dof_count)
    if plan.reduced_dof_count == 0:  # no free dof: u = shift
        y = to_device(plan.shift if include_shift else np.zeros(plan.full_dof_count))
        return _out_like(x, y)
    y = _torch().empty(plan.full_dof_count, dtype=xt.dtype, device=xt.device)
    _call("gmf_expand_reduced", dev.ws, xt.data_ptr(), y.data_ptr(), int(bool(include_shift)))
    return _out_like(x, y)


def project_reduced(plan, y_full):
    dev = _ws(plan)
    yt = to_device(y_full, plan.full_dof_count)
    r = _torch().empty(plan.reduced_dof_count, dtype=yt.dtype, device=yt.device)
    _call("gmf_project_reduced", dev.ws, yt.data_ptr(), r.data_ptr())
    return _out_like(y_full, r)


def diagonal_physical(plan, as_numpy=True):
    dev = _ws(plan)
    torch = _torch()
    d = torch.empty(plan.full_dof_count, dtype=torch.float64, device="cuda")
    _call("gmf_diagonal_physical", dev.ws, d.data_ptr())
    return d.cpu().numpy() if as_numpy else d


def reduced_diagonal(plan, as_numpy=True):
    dev = _ws(plan)
    torch = _torch()
    d = torch.empty(plan.reduced_dof_count, dtype=torch.float64, device="cuda")
    _call("gmf_reduced_diagonal", dev.ws, d.data_ptr())
    return d.cpu().numpy() if as_numpy else d
