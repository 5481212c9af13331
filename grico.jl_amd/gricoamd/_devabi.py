"""ctypes mirror of include/grico_mf.h and the device-plan wrapper."""

import ctypes

import numpy as np

from . import _native

c_int32, c_int64, c_double, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_f64 = ctypes.POINTER(ctypes.c_double)

STATUS = {
    0: "GMF_OK", -1: "GMF_EINVAL", -2: "GMF_EALIAS", -3: "GMF_ENOCONV", -4: "GMF_EBREAKDOWN",
    -5: "GMF_EDIAGONAL", -6: "GMF_ENOTSPD", -7: "GMF_EDEVICE", -8: "GMF_EUNSUPPORTED",
}


class BucketDesc(ctypes.Structure):
    _fields_ = [
        ("dim", c_int32), ("modes", c_int32 * 3), ("points", c_int32 * 3), ("form", c_int32),
        ("cell_count", c_int64),
        ("values", P_f64), ("derivatives", P_f64), ("inverse_jacobian", P_f64), ("point_data", P_f64),
        ("local_index", P_i32), ("local_coefficient", P_f64),
        ("multi_count", c_int64), ("multi_offsets", P_i64), ("multi_indices", P_i32),
        ("multi_coefficients", P_f64), ("point_offsets", P_i64), ("ref_points", P_f64),
        ("sides", c_int32), ("plus_modes", c_int32 * 3),
        ("brick_cells", P_i32), ("brick_count", c_int64),
    ]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("dim", c_int32), ("full_dof_count", c_int64), ("bucket_count", c_int32),
        ("buckets", ctypes.POINTER(BucketDesc)),
        ("fixed_count", c_int64), ("fixed_dofs", P_i64),
        ("affine_count", c_int64), ("affine_pivots", P_i64), ("affine_offsets", P_i64),
        ("affine_indices", P_i64), ("affine_coefficients", P_f64),
        ("mean_count", c_int64), ("mean_pivots", P_i64), ("mean_offsets", P_i64),
        ("mean_indices", P_i64), ("mean_coefficients", P_f64),
        ("reduced_dof_count", c_int64), ("map_offsets", P_i64), ("map_indices", P_i64),
        ("map_coefficients", P_f64), ("shift", P_f64),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int64), ("operator_applications", c_int64),
        ("preconditioner_applications", c_int64), ("initial_residual_norm", c_double),
        ("final_residual_norm", c_double), ("rhs_norm", c_double), ("converged", c_int32),
    ]


class BucketTuning(ctypes.Structure):
    _fields_ = [("kernel", c_int32), ("param0", c_int32), ("param1", c_int32), ("candidates", c_int32),
                ("best_us", c_double), ("default_us", c_double)]


# gmf_bucket_tuning.kernel
TUNED_KERNELS = {0: "empty", 1: "brick", 2: "cell", 3: "generic", 4: "point-cloud", 5: "face-tensor",
                 6: "point-cloud-matrix"}


class TransferDesc(ctypes.Structure):
    _fields_ = [("nnz", c_int64), ("coarse_indices", P_i64), ("fine_indices", P_i64), ("coefficients", P_f64)]


class BlockDesc(ctypes.Structure):
    _fields_ = [("count", c_int64), ("offsets", P_i64), ("dofs", P_i32), ("inverse", P_f64)]


class MgDesc(ctypes.Structure):
    _fields_ = [
        ("level_count", c_int32), ("levels", ctypes.POINTER(c_void_p)), ("transfers", ctypes.POINTER(TransferDesc)),
        ("smoother", c_int32), ("pre_steps", c_int32), ("post_steps", c_int32), ("damping", c_double),
        ("chebyshev_degree", c_int32), ("chebyshev_range", c_double), ("eigen_iterations", c_int32),
        ("coarse_direct_limit", c_int64), ("coarse_rtol", c_double), ("coarse_atol", c_double),
        ("coarse_maxiter", c_int64), ("graph", c_int32), ("precision", c_int32),
        ("smoother_restart", c_int32), ("spd", c_int32), ("level_diagonals", ctypes.POINTER(P_f64)),
        ("level_blocks", ctypes.POINTER(BlockDesc)),
    ]


def declare_mg(lib):
    if getattr(lib, "_mg_declared", False):
        return
    lib.gmf_precond_create_mg.argtypes = [ctypes.POINTER(MgDesc), ctypes.POINTER(c_void_p)]
    lib.gmf_precond_create_mg.restype = ctypes.c_int
    lib.gmf_mg_level_info.argtypes = [c_void_p, c_int32, P_i64, P_f64]
    lib.gmf_mg_level_info.restype = ctypes.c_int
    lib.gmf_reduced_blocks.argtypes = [c_void_p, ctypes.POINTER(BlockDesc), P_i32, c_int32, P_f64]
    lib.gmf_reduced_blocks.restype = ctypes.c_int
    lib._mg_declared = True


def declare(lib):
    lib.gmf_last_error.restype = ctypes.c_char_p
    lib.gmf_device_info.argtypes = [ctypes.c_char_p, P_i32, P_i64]
    lib.gmf_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(c_void_p)]
    lib.gmf_plan_destroy.argtypes = [c_void_p]
    lib.gmf_plan_sizes.argtypes = [c_void_p, P_i64, P_i64, P_i64]
    lib.gmf_workspace_create.argtypes = [c_void_p, ctypes.POINTER(c_void_p)]
    lib.gmf_workspace_destroy.argtypes = [c_void_p]
    lib.gmf_workspace_set_stream.argtypes = [c_void_p, c_void_p]
    for name in ("gmf_apply", "gmf_apply_async", "gmf_apply_physical", "gmf_reduced_apply",
                 "gmf_reduced_apply_f32", "gmf_project_reduced"):
        getattr(lib, name).argtypes = [c_void_p, c_void_p, c_void_p]
    lib.gmf_vector_copy.argtypes = [c_void_p, c_void_p, c_void_p, c_int64]
    lib.gmf_plan_describe.argtypes = [c_void_p, ctypes.c_char_p, c_int64]
    lib.gmf_apply_timed.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, P_f64, P_f64]
    lib.gmf_precond_create_jacobi.argtypes = [c_void_p, c_void_p, ctypes.POINTER(c_void_p)]
    lib.gmf_precond_destroy.argtypes = [c_void_p]
    lib.gmf_precond_apply.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.gmf_precond_apply_block.argtypes = [c_void_p, c_void_p, c_void_p, c_int32]
    lib.gmf_expand_reduced.argtypes = [c_void_p, c_void_p, c_void_p, c_int32]
    lib.gmf_rhs_physical.argtypes = [c_void_p, c_void_p, c_int32, c_void_p]
    lib.gmf_reduced_rhs.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.gmf_l2_error.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, P_f64]
    lib.gmf_interpolate_channels.argtypes = [c_void_p, c_void_p, c_int32, c_void_p]
    lib.gmf_rhs_faces.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.gmf_plan_set_point_data.argtypes = [c_void_p, c_int32, P_f64]
    lib.gmf_diagonal_physical.argtypes = [c_void_p, c_void_p]
    lib.gmf_reduced_diagonal.argtypes = [c_void_p, c_void_p]
    lib.gmf_invert_diagonal.argtypes = [c_void_p, c_void_p, c_void_p, c_int64]
    lib.gmf_cg_solve.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double,
                                 c_double, c_int64, ctypes.POINTER(SolveInfo)]
    lib.gmf_fgmres_solve.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32,
                                     c_double, c_double, c_int64, ctypes.POINTER(SolveInfo)]
    lib.gmf_reduced_apply_block.argtypes = [c_void_p, c_void_p, c_void_p, c_int32]
    lib.gmf_cg_solve_block.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_double,
                                       c_double, c_int64, ctypes.POINTER(SolveInfo)]
    lib.gmf_precond_create_jacobi_f32.argtypes = [c_void_p, c_void_p, ctypes.POINTER(c_void_p)]
    lib.gmf_plan_tuning.argtypes = [c_void_p, c_int32, ctypes.POINTER(BucketTuning)]
    for name in ("gmf_plan_tuning", "gmf_plan_create", "gmf_plan_destroy", "gmf_plan_sizes", "gmf_workspace_create",
                 "gmf_workspace_destroy", "gmf_workspace_set_stream", "gmf_apply", "gmf_apply_async",
                 "gmf_apply_physical", "gmf_reduced_apply", "gmf_reduced_apply_f32", "gmf_project_reduced",
                 "gmf_expand_reduced",
                 "gmf_diagonal_physical", "gmf_reduced_diagonal", "gmf_invert_diagonal", "gmf_cg_solve",
                 "gmf_fgmres_solve", "gmf_rhs_physical", "gmf_reduced_rhs", "gmf_l2_error",
                 "gmf_interpolate_channels", "gmf_plan_set_point_data", "gmf_rhs_faces", "gmf_device_info", "gmf_apply_timed", "gmf_precond_create_jacobi",
                 "gmf_precond_destroy", "gmf_precond_apply", "gmf_precond_apply_block", "gmf_plan_describe", "gmf_vector_copy",
                 "gmf_reduced_apply_block", "gmf_cg_solve_block", "gmf_precond_create_jacobi_f32"):
        getattr(lib, name).restype = ctypes.c_int


class DeviceError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def check(status):
    if status != 0:
        raise DeviceError(status, _native.device().gmf_last_error().decode())


def _p(a, t):
    return _native.ptr(a, t)


class DevicePlan:
    """Uploads an AssemblyPlan's tables once and owns a workspace."""

    def __init__(self, plan):
        lib = _native.device()
        keep = []
        buckets = getattr(plan, "all_buckets", plan.buckets)
        descs = (BucketDesc * len(buckets))()
        for i, b in enumerate(buckets):
            d = descs[i]
            d.dim = b.dim
            cloud = getattr(b, "cloud", False)
            for a in range(3):
                d.modes[a] = b.modes[a] if a < b.dim else 1
                d.points[a] = (1 if cloud else b.points[a]) if a < b.dim else 1
            d.form = b.form
            d.cell_count = b.cell_count
            d.sides = getattr(b, "sides", 1)
            if d.sides == 2:
                for a in range(3):
                    d.plus_modes[a] = b.plus_modes[a] if a < b.dim else 1
            pdata = np.ascontiguousarray(b.point_data, dtype=np.float64)
            coef = None if b.unit_coefficients else b.local_coefficient
            keep += [pdata, b.inverse_jacobian, b.local_index, coef, b.multi_offsets,
                     b.multi_indices, b.multi_coefficients]
            if cloud:
                po = np.ascontiguousarray(b.point_offsets, dtype=np.int64)
                rp = np.ascontiguousarray(b.ref, dtype=np.float64)
                keep += [po, rp]
                d.point_offsets = _p(po, c_int64)
                d.ref_points = _p(rp, c_double)
                d.values = None
                d.derivatives = None
            else:
                vals = np.ascontiguousarray(np.concatenate([t.reshape(-1) for t in b.B]))
                ders = np.ascontiguousarray(np.concatenate([t.reshape(-1) for t in b.G]))
                keep += [vals, ders]
                d.values = _p(vals, c_double)
                d.derivatives = _p(ders, c_double)
                d.point_offsets = None
                d.ref_points = None
            d.inverse_jacobian = _p(b.inverse_jacobian, c_double)
            d.point_data = _p(pdata, c_double)
            d.local_index = _p(b.local_index, c_int32)
            d.local_coefficient = _p(coef, c_double)
            d.multi_count = len(b.multi_offsets) - 1
            d.multi_offsets = _p(b.multi_offsets, c_int64)
            d.multi_indices = _p(b.multi_indices, c_int32)
            d.multi_coefficients = _p(b.multi_coefficients, c_double)
            bricks = None if cloud else getattr(b, "brick_cells", None)
            if bricks is not None:
                bricks = np.ascontiguousarray(bricks, dtype=np.int32)
                keep.append(bricks)
                d.brick_cells = _p(bricks, c_int32)
                d.brick_count = len(bricks)
        pd = PlanDesc()
        pd.dim = plan.dim
        pd.full_dof_count = plan.full_dof_count
        pd.bucket_count = len(buckets)
        pd.buckets = descs
        fixed = _native.as_i64(plan.fixed_dofs)
        keep.append(fixed)
        pd.fixed_count = len(fixed)
        pd.fixed_dofs = _p(fixed, c_int64)
        for name, rows in (("affine", getattr(plan, "affine_rows", [])), ("mean", getattr(plan, "mean_rows", []))):
            setattr(pd, name + "_count", len(rows))
            if not rows:
                continue
            piv = np.asarray([r[0] for r in rows], dtype=np.int64)
            off = np.zeros(len(rows) + 1, dtype=np.int64)
            off[1:] = np.cumsum([len(r[1]) for r in rows])
            idx = np.ascontiguousarray(np.concatenate([r[1] for r in rows]).astype(np.int64))
            cf = np.ascontiguousarray(np.concatenate([r[2] for r in rows]).astype(np.float64))
            keep += [piv, off, idx, cf]
            setattr(pd, name + "_pivots", _p(piv, c_int64))
            setattr(pd, name + "_offsets", _p(off, c_int64))
            setattr(pd, name + "_indices", _p(idx, c_int64))
            setattr(pd, name + "_coefficients", _p(cf, c_double))
        pd.reduced_dof_count = plan.reduced_dof_count
        mo, mi, mc, sh = (_native.as_i64(plan.map_offsets), _native.as_i64(plan.map_indices),
                          _native.as_f64(plan.map_coefficients), _native.as_f64(plan.shift))
        keep += [mo, mi, mc, sh]
        pd.map_offsets = _p(mo, c_int64)
        pd.map_indices = _p(mi, c_int64)
        pd.map_coefficients = _p(mc, c_double)
        pd.shift = _p(sh, c_double)
        handle = c_void_p()
        check(lib.gmf_plan_create(ctypes.byref(pd), ctypes.byref(handle)))
        self.handle = handle
        self.bucket_count = len(buckets)
        ws = c_void_p()
        check(lib.gmf_workspace_create(handle, ctypes.byref(ws)))
        self.ws = ws
        self.full_dof_count = plan.full_dof_count
        self.reduced_dof_count = plan.reduced_dof_count

    def __del__(self):
        if _native is None:  # interpreter shutdown
            return
        lib = _native._device
        if lib is None:
            return
        if getattr(self, "ws", None) is not None and self.ws.value:
            lib.gmf_workspace_destroy(self.ws)
            self.ws = None
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib.gmf_plan_destroy(self.handle)
            self.handle = None

    def set_stream(self, stream_ptr):
        check(_native.device().gmf_workspace_set_stream(self.ws, c_void_p(stream_ptr)))

    def describe(self):
        buf = ctypes.create_string_buffer(1 << 16)
        check(_native.device().gmf_plan_describe(self.handle, buf, len(buf)))
        return buf.value.decode()

    def tuning(self):
        """Kernel family and launch shape of every device bucket as picked by
        the plan-time tuner (gmf_plan_tuning), in bucket order."""
        out = []
        lib = _native.device()
        for i in range(self.bucket_count):
            t = BucketTuning()
            check(lib.gmf_plan_tuning(self.handle, i, ctypes.byref(t)))
            out.append({"kernel": TUNED_KERNELS.get(t.kernel, t.kernel), "param0": t.param0, "param1": t.param1,
                        "candidates": t.candidates, "best_us": t.best_us, "default_us": t.default_us})
        return out

    def device_bytes(self):
        a, b, c = c_int64(), c_int64(), c_int64()
        check(_native.device().gmf_plan_sizes(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return c.value
