"""Learned per-cell block correction for FGMRES (no reference counterpart).

A small MLP maps the normalized residual on one cell's local modes to a local
error correction; the corrections of all cells are averaged on shared dofs and
added after an inner preconditioner (the hp-GMG V-cycle):

    z0 = V(r),  r1 = r - A z0,  z = z0 + sum_c P_c^T (w * net(r1_c / |r1_c|) |r1_c| / kappa_c)

kappa_c is the cell-mean diffusion coefficient, so one network serves cells
of different conductivity. The preconditioner varies with the data it sees, so
it is used inside FGMRES (src/solve.jl:80-88 allows that); the seam is the
preconditioner policy interface (src/solve.jl:32, 641-670, 753-774).

Training data come from the GMG-CG solves themselves, collected on the device
(gmf_learned_collect): for seeded reference solutions x* and the CG iterates
x_k, the pair (r', e') = (b - A y, x* - y) with y = x_k + V(b - A x_k) is the
residual / error the V-cycle leaves; the residuals are kept as FP32 rows with
the energies e'^T r'. The network (BF16 weights, FP32 accumulation on MFMA) is
trained on the device with Adam to minimise the energy norm of the error left
after its correction, ||e' - c(r')||_A^2 = e'^T A e' - 2 c^T r' + c^T A c,
which is what the outer Krylov method sees (gmf_learned_train). A supervised
fit of per-cell patches (gmf_mlp_train on gmf_cell_patches output) is
available for networks used outside the preconditioner.
"""

import ctypes
import math
import time

import numpy as np

from . import _devabi, _native, ops
from .multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid


def _declare(lib):
    if getattr(lib, "_mlp_declared", False):
        return
    c_void_p, c_int64, c_int32, c_double, c_float = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                                     ctypes.c_double, ctypes.c_float)
    lib.gmf_mlp_create.argtypes = [ctypes.POINTER(MlpDesc), ctypes.POINTER(c_void_p)]
    lib.gmf_mlp_destroy.argtypes = [c_void_p]
    lib.gmf_mlp_set_stream.argtypes = [c_void_p, c_void_p]
    lib.gmf_mlp_parameter_count.argtypes = [c_void_p, ctypes.POINTER(c_int64)]
    lib.gmf_mlp_get_parameters.argtypes = [c_void_p, c_void_p]
    lib.gmf_mlp_forward.argtypes = [c_void_p, c_void_p, c_void_p, c_int64]
    lib.gmf_mlp_gradient.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, ctypes.POINTER(c_double), c_void_p]
    lib.gmf_mlp_train.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_float, c_float, c_float,
                                  c_float, c_void_p]
    lib.gmf_cell_patches.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                     c_void_p]
    lib.gmf_precond_create_learned.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_void_p)]
    lib.gmf_vector_axpby.argtypes = [c_void_p, c_void_p, c_double, c_void_p, c_double, c_int64]
    lib.gmf_vector_dot.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, ctypes.POINTER(c_double)]
    lib.gmf_learned_train.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_float, c_float, c_float,
                                      c_float, c_void_p]
    lib.gmf_learned_gradient.argtypes = [c_void_p, c_void_p, c_double, ctypes.POINTER(c_double), c_void_p]
    lib.gmf_learned_collect.argtypes = [c_void_p, c_int64, c_int32, ctypes.c_uint64, c_void_p, c_void_p]
    for name in ("gmf_learned_train", "gmf_learned_gradient", "gmf_learned_collect", "gmf_mlp_create", "gmf_mlp_destroy", "gmf_mlp_set_stream", "gmf_mlp_parameter_count",
                 "gmf_mlp_get_parameters", "gmf_mlp_forward", "gmf_mlp_gradient", "gmf_mlp_train", "gmf_cell_patches",
                 "gmf_precond_create_learned", "gmf_vector_axpby", "gmf_vector_dot"):
        getattr(lib, name).restype = ctypes.c_int
    lib._mlp_declared = True


class MlpDesc(ctypes.Structure):
    _fields_ = [("input", ctypes.c_int32), ("hidden1", ctypes.c_int32), ("hidden2", ctypes.c_int32),
                ("output", ctypes.c_int32), ("parameters", ctypes.POINTER(ctypes.c_float))]


def init_parameters(widths, seed=0):
    """He-uniform weights, zero biases, packed [W1, b1, W2, b2, W3, b3]."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, n in zip(widths[:-1], widths[1:]):
        bound = math.sqrt(6.0 / k)
        parts.append(rng.uniform(-bound, bound, size=(k, n)).astype(np.float32).reshape(-1))
        parts.append(np.zeros(n, dtype=np.float32))
    out = np.concatenate(parts)
    # the last layer starts small: the untrained correction is close to zero
    n_last = widths[-2] * widths[-1]
    start = out.size - widths[-1] - n_last
    out[start:start + n_last] *= 0.1
    return out


class CellMLP:
    """Device MLP widths = (in, h1, h2, out) (gmf_mlp_*)."""

    def __init__(self, widths=(64, 256, 256, 64), seed=0, parameters=None):
        if len(widths) != 4:
            raise ValueError("widths must be (input, hidden1, hidden2, output)")
        lib = _native.device()
        _declare(lib)
        self.widths = tuple(int(w) for w in widths)
        params = init_parameters(self.widths, seed) if parameters is None else \
            np.ascontiguousarray(parameters, dtype=np.float32)
        d = MlpDesc(*self.widths, params.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        self.handle = ctypes.c_void_p()
        _devabi.check(lib.gmf_mlp_create(ctypes.byref(d), ctypes.byref(self.handle)))
        self._sync_stream()

    def __del__(self):
        if _native is None:  # interpreter shutdown
            return
        lib = _native._device
        if lib is not None and getattr(self, "handle", None) is not None and self.handle.value:
            lib.gmf_mlp_destroy(self.handle)
            self.handle = None

    def _sync_stream(self):
        torch = ops._torch()
        _devabi.check(_native.device().gmf_mlp_set_stream(self.handle,
                                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    @property
    def parameter_count(self):
        n = ctypes.c_int64()
        _devabi.check(_native.device().gmf_mlp_parameter_count(self.handle, ctypes.byref(n)))
        return n.value

    def parameters(self):
        out = np.zeros(self.parameter_count, dtype=np.float32)
        _devabi.check(_native.device().gmf_mlp_get_parameters(self.handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def forward(self, x):
        torch = ops._torch()
        self._sync_stream()
        x = x.to(device="cuda", dtype=torch.float32).contiguous()
        rows = x.shape[0]
        y = torch.empty((rows, self.widths[-1]), dtype=torch.float32, device="cuda")
        _devabi.check(_native.device().gmf_mlp_forward(self.handle, x.data_ptr(), y.data_ptr(), rows))
        return y

    def gradient(self, x, t):
        torch = ops._torch()
        self._sync_stream()
        x = x.to(device="cuda", dtype=torch.float32).contiguous()
        t = t.to(device="cuda", dtype=torch.float32).contiguous()
        loss = ctypes.c_double()
        g = np.zeros(self.parameter_count, dtype=np.float32)
        _devabi.check(_native.device().gmf_mlp_gradient(self.handle, x.data_ptr(), t.data_ptr(), x.shape[0],
                                                        ctypes.byref(loss), g.ctypes.data_as(ctypes.c_void_p)))
        return loss.value, g

    def train(self, x, t, batch=4096, steps=2000, learning_rate=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
              record=True):
        torch = ops._torch()
        self._sync_stream()
        x = x.to(device="cuda", dtype=torch.float32).contiguous()
        t = t.to(device="cuda", dtype=torch.float32).contiguous()
        hist = np.zeros(steps, dtype=np.float64) if record else None
        _devabi.check(_native.device().gmf_mlp_train(
            self.handle, x.data_ptr(), t.data_ptr(), x.shape[0], int(batch), int(steps), float(learning_rate),
            float(beta1), float(beta2), float(eps), hist.ctypes.data_as(ctypes.c_void_p) if record else None))
        return hist


def cell_weights(plan):
    """kappa_c = sum_q W_q c_q / sum_q W_q per cell of the single diffusion batch."""
    if len(plan.buckets) != 1 or plan.buckets[0].form != 0:
        raise ValueError("the learned correction needs one diffusion-form cell batch")
    b = plan.buckets[0]
    w = np.asarray(b.weights).reshape(b.cell_count, -1)
    return (np.asarray(b.point_data).reshape(b.cell_count, -1).sum(axis=1) / w.sum(axis=1)).astype(np.float32)


def _ws(plan):
    return ops._ws(plan).ws


def collect_residuals(learned, pairs=2000, iterations=40, seed=0):
    """Training residuals from GMG-CG solves with the learned preconditioner's
    inner V-cycle, collected on the device (gmf_learned_collect, module
    docstring): a (pairs x n_red) float32 device tensor of residuals r' and a
    host array of energies e'^T r'."""
    torch = ops._torch()
    lib = _native.device()
    _declare(lib)
    n = learned.plan.reduced_dof_count
    R = torch.empty((pairs, n), dtype=torch.float32, device="cuda")
    energy = np.zeros(pairs)
    _ws(learned.plan)  # the plan's stream is torch's current stream
    _devabi.check(lib.gmf_learned_collect(learned.handle, int(pairs), int(iterations), int(seed), R.data_ptr(),
                                          energy.ctypes.data_as(ctypes.c_void_p)))
    return R, energy


class LearnedBlockCorrection:
    """Preconditioner policy: the hp-GMG V-cycle followed by the learned cell
    correction. `network` must have the cell's local mode count as width;
    `inner` is a GeometricMultigridPreconditioner policy or an already
    compiled DevicePreconditioner of the same plan."""

    def __init__(self, network, inner=None):
        self.network = network
        self.inner = inner if inner is not None else GeometricMultigridPreconditioner(smoother="chebyshev")

    def compile(self, plan):
        from .solve import DevicePreconditioner
        lib = _native.device()
        _declare(lib)
        inner_handle, keep = None, ()
        if isinstance(self.inner, DevicePreconditioner):
            inner_handle, keep = self.inner.handle, (self.inner,)
        elif self.inner is not None:
            hierarchy = compile_geometric_multigrid(plan.problem, self.inner, fine_plan=plan)
            M = hierarchy.device_preconditioner()
            inner_handle, keep = M.handle, (M, hierarchy)
        cw = cell_weights(plan)
        h = ctypes.c_void_p()
        _devabi.check(lib.gmf_precond_create_learned(_ws(plan), inner_handle, self.network.handle,
                                                     cw.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        return DevicePreconditioner(plan, h, keep=(keep, self.network, cw))


def train_block_correction(problem, plan=None, widths=None, pairs=2000, steps=3000, learning_rate=1e-3, seed=0,
                           inner=None, iterations=40, log=None, hierarchy=None):
    """Build the GMG hierarchy (policy `inner`, Chebyshev smoothing by
    default; or take a compiled `hierarchy` on `plan`), collect V-cycle
    residuals from GMG-CG solves on the device and train a CellMLP in the
    energy norm. Returns (network, report, compiled learned preconditioner,
    compiled V-cycle)."""
    from .plan import compile as compile_plan
    torch = ops._torch()
    plan = plan or (hierarchy.plans[-1] if hierarchy is not None else compile_plan(problem))
    inner = inner or GeometricMultigridPreconditioner(smoother="chebyshev")
    t0 = time.perf_counter()
    if hierarchy is None:
        hierarchy = compile_geometric_multigrid(problem, inner, fine_plan=plan)
    elif hierarchy.plans[-1] is not plan:
        raise ValueError("the hierarchy's fine level must be the given plan")
    V = hierarchy.device_preconditioner()
    width = plan.buckets[0].n_modes
    widths = widths or (width, 256, 256, width)
    net = CellMLP(widths, seed=seed)
    M = LearnedBlockCorrection(net, inner=V).compile(plan)
    t1 = time.perf_counter()
    R, energy = collect_residuals(M, pairs=pairs, iterations=iterations, seed=seed)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    hist = np.zeros(steps)
    _devabi.check(_native.device().gmf_learned_train(M.handle, R.data_ptr(), energy.ctypes.data_as(ctypes.c_void_p),
                                                     pairs, steps, float(learning_rate), 0.9, 0.999, 1e-8,
                                                     hist.ctypes.data_as(ctypes.c_void_p)))
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    del R
    k = max(1, min(50, steps // 10))
    report = {"pairs": pairs, "hierarchy_seconds": t1 - t0, "collect_seconds": t2 - t1, "train_seconds": t3 - t2,
              "steps": steps, "energy_ratio_first": float(hist[:k].mean()), "energy_ratio_last": float(hist[-k:].mean())}
    if log:
        log(f"learned correction: {report}")
    M._keep = (M._keep, hierarchy)
    return net, report, M, V
