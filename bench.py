"""Benchmark: matrix-free FP64 apply of 3D Poisson, 64^3 cells, p = 4.

Workload (the headline metric "MF apply GDOF/s"): [0,1]^3 built as six
uniform all-axis refinements of one root cell (64^3 = 262,144 cells),
FullTensorBasis degree 4, Gauss 5^3 points per cell, kappa = 1 + 0.5
sin(2 pi x) sin(2 pi y) sin(2 pi z), zero Dirichlet on all faces:
16,974,593 full-layout dofs. One step = one full-layout apply (cell pass +
constraint rows) of a seeded N(0,1) vector resident in HBM.

Prints one JSON line. `roofline` is the dominant kernel (the cell pass)
against MI355X's 8 TB/s HBM peak using the algorithmic bytes
8 N_full (read x) + 8 N_full (write y) + 8 n_cells n_q (per-point
coefficients); its duration is measured with HIP events on the launch
stream. `cpu_baseline` times the oracle (the reference algorithm restated
in C, OpenMP over DOF-disjoint colors) on a bounded 16^3-cell sample.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "grico.jl_amd"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_problem(depth, degree=4):
    import numpy as np

    import gricoamd as g

    class Diffusion:
        def cell_accumulate(self, q, trial, test_component):
            x, y, z = q.point
            k = 1.0 + 0.5 * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
            return g.TestChannels(0.0 * trial.value, tuple(k * c for c in trial.gradient))

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            x, y, z = q.point
            return 3.0 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)

    domain = g.Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    g.uniform_refine(domain, depth)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, Diffusion())
    problem.add_cell_accumulator(u, Load())
    for axis in range(3):
        for side in (g.LOWER, g.UPPER):
            problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    return problem


def cpu_baseline(budget_s=12.0):
    """Oracle apply on a 16^3-cell p=4 sample, all host threads, GDOF/s."""
    import numpy as np

    import gricoamd as g
    import pyoracle

    threads = host_threads()
    plan = g.compile(build_problem(4))
    mats = [pyoracle.action_matrices(plan, bk) for bk in plan.buckets]
    cols = [pyoracle.color_batches(plan, pyoracle._Bucket(bk)) for bk in plan.buckets]
    x = np.random.default_rng(1).standard_normal(plan.full_dof_count)
    pyoracle.apply_lowered(plan, x, threads=threads, matrices=mats, colors=cols)
    reps = 0
    t0 = time.perf_counter()
    while True:
        pyoracle.apply_lowered(plan, x, threads=threads, matrices=mats, colors=cols)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    per = el / reps
    # The device apply on the same sample and input, against the oracle.
    y_ref = pyoracle.apply_lowered(plan, x, threads=threads, matrices=mats, colors=cols)
    y_dev = g.apply(plan, x)
    err = float(np.abs(y_dev - y_ref).max() / np.abs(y_ref).max())
    return {
        "value": plan.full_dof_count / per / 1e9,
        "unit": "GDOF/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle full-layout apply, 16^3 cells p=4 ({plan.full_dof_count} dofs), "
                  f"{reps} reps in {el:.1f}s, {threads} OpenMP threads over DOF-disjoint colors",
        "device_max_relative_error_vs_oracle": err,
    }


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(int(n), 64))


def oracle_relative_residuals(plan, solutions, b_red):
    """{name: ||b - A x|| / ||b||} of device solutions with A the C oracle's
    FP64 operator (lowered apply, all host threads): each solution checked
    against the reference algorithm's operator rather than the device's own."""
    import numpy as np

    import pyoracle

    mats = [pyoracle.action_matrices(plan, bk) for bk in plan.buckets]
    cols = [pyoracle.color_batches(plan, pyoracle._Bucket(bk)) for bk in plan.buckets]
    out = {}
    for name, x_red in solutions.items():
        full = np.zeros(plan.full_dof_count)
        full[plan.solve_dofs] = x_red.cpu().numpy()
        y = pyoracle.apply_lowered(plan, full, threads=host_threads(), matrices=mats, colors=cols,
                                   physical=True)[plan.solve_dofs]
        out[name] = float(np.linalg.norm(b_red - y) / np.linalg.norm(b_red))
    return out


def _device_solver_kit(plan, repeats):
    """(b, x, timed(fn) -> dict, cg(handle) -> info) for device CG solves of
    the plan's reduced system from x0 = 0 to ||r|| <= 1e-8 ||b||."""
    import ctypes

    import torch

    from gricoamd import _devabi, _native, ops

    lib = _native.device()
    b = ops.to_device(plan.reduced_rhs(), plan.reduced_dof_count)
    x = torch.zeros_like(b)
    dev = ops._ws(plan)

    def timed(fn):
        fn()
        ts = []
        info = None
        for _ in range(repeats):
            torch.cuda.synchronize()
            t = time.perf_counter()
            info = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return {"seconds": sorted(ts)[len(ts) // 2], "iterations": info.iterations,
                "final_relative_residual": info.final_residual_norm / info.rhs_norm}

    def cg(handle, maxiter=20000):
        info = _devabi.SolveInfo()
        _devabi.check(lib.gmf_cg_solve(dev.ws, handle, b.data_ptr(), x.data_ptr(), 0, 1e-8, 0.0, maxiter,
                                       ctypes.byref(info)))
        return info

    return b, x, timed, cg


def _vcycle_seconds(M, n, reps=10):
    import torch

    from gricoamd import _devabi, _native

    lib = _native.device()
    r = torch.randn(n, dtype=torch.float64, device="cuda")
    z = torch.empty_like(r)
    for _ in range(2):
        _devabi.check(lib.gmf_precond_apply(M.handle, r.data_ptr(), z.data_ptr()))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        _devabi.check(lib.gmf_precond_apply(M.handle, r.data_ptr(), z.data_ptr()))
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def gmg_policy_record(pol, levels):
    return {"smoother": f"Chebyshev-Jacobi degree {pol.chebyshev_degree} on [lambda_max/{pol.chebyshev_range:g}, "
                        f"lambda_max], V({pol.pre_smoothing_steps},{pol.post_smoothing_steps})",
            "precision": pol.precision, "p_sequence": pol.p_sequence, "h_steps_per_level": pol.h_steps_per_level,
            "coarse": f"device dense inverse at <= {pol.coarse_direct_dof_limit} dofs", "levels": levels}


def solve_leg(problem, plan, repeats=3):
    """GMG-CG (FP32 and FP64 Chebyshev V-cycles) and Jacobi-CG from x0 = 0 to
    ||r|| <= 1e-8 ||b|| on the device, hierarchy already compiled; every
    solution's residual is re-evaluated with the FP64 oracle operator. The
    V-cycle (degree 3, hierarchy cut at the first level of at most 4096 dofs,
    solved there by a device dense inverse) is the fastest of a sweep over
    degree 2-6, range 20/50, p-sequence and h-steps per level on this
    problem."""
    import torch

    import gricoamd as g
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid
    from gricoamd.solve import compile_preconditioner

    t0 = time.perf_counter()
    hierarchy = compile_geometric_multigrid(problem, GeometricMultigridPreconditioner(smoother="chebyshev"),
                                            fine_plan=plan)
    t_host = time.perf_counter() - t0
    counts = hierarchy.reduced_dof_counts()
    keep = len(counts) - max(i for i, c in enumerate(counts) if c <= 4096)
    b, x, timed, cg = _device_solver_kit(plan, repeats)
    b_host = b.cpu().numpy()
    out = {"metric": "time-to-1e-8 relative residual, reduced system (s)", "hierarchy_host_seconds": t_host,
           "load": "f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z)"}
    sols = {}
    for tag, prec in (("gmg_cg", "fp32"), ("gmg_cg_fp64_vcycle", "fp64")):
        pol = GeometricMultigridPreconditioner(smoother="chebyshev", precision=prec, chebyshev_degree=3,
                                               max_levels=keep, coarse_direct_dof_limit=4096)
        h = hierarchy.with_policy(pol)
        M = h.device_preconditioner()
        res = timed(lambda: cg(M.handle))
        res["vcycle_seconds"] = _vcycle_seconds(M, plan.reduced_dof_count)
        res["policy"] = gmg_policy_record(pol, h.reduced_dof_counts())
        sols[tag] = x.clone()
        out[tag] = res
        log(f"C2 {tag}: {res}")
        del M
    J = compile_preconditioner(g.JacobiPreconditioner(), plan)
    out["jacobi_cg"] = timed(lambda: cg(J.handle))
    sols["jacobi_cg"] = x.clone()
    out["gmg_cg_speedup_vs_jacobi_cg"] = out["jacobi_cg"]["seconds"] / out["gmg_cg"]["seconds"]
    nrm = float(torch.linalg.norm(sols["jacobi_cg"]))
    out["solution_relative_difference_vs_jacobi_cg"] = {
        k: float(torch.linalg.norm(v - sols["jacobi_cg"])) / nrm for k, v in sols.items() if k != "jacobi_cg"}
    t0 = time.perf_counter()
    out["oracle_relative_residual"] = oracle_relative_residuals(plan, sols, b_host)
    out["oracle_check_seconds"] = time.perf_counter() - t0
    log(f"C2 solve leg: {out}")
    return out


def build_adaptive_problem(base=16, levels=5, pmax=8):
    """3D adaptive hp Poisson: base^3 cells from uniform refinements of one root,
    `levels` rounds of dyadic refinement of the leaf at (0,0,0) (all axes on
    even rounds, x and y on odd rounds: anisotropic cells, hanging nodes),
    per-leaf degree 1..pmax rising with the distance of the leaf center from
    the corner, u = r^(2/3) manufactured (-div grad u = -(10/9) r^(-4/3)),
    Dirichlet data u on every face."""
    import numpy as np

    import gricoamd as g

    def exact(p):
        return np.sqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2) ** (2.0 / 3.0)

    class Diffusion:
        def cell_accumulate(self, q, trial, test_component):
            return g.TestChannels(0.0 * trial.value, tuple(trial.gradient))

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            x, y, z = q.point
            return -(10.0 / 9.0) * (x * x + y * y + z * z) ** (-2.0 / 3.0)

    domain = g.Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    g.uniform_refine(domain, int(round(np.log2(base))))
    for lev in range(levels):
        axes = (0, 1, 2) if lev % 2 == 0 else (0, 1)
        leaves = domain.active_leaves()
        for leaf in [int(l) for l, lo in zip(leaves, domain.cell_lowers(leaves)) if np.all(lo == 0.0)]:
            cur = [leaf]
            for a in axes:
                nxt = []
                for c in cur:
                    f = domain.refine(c, a)
                    nxt += [f, f + 1]
                cur = nxt

    def degree(dom, leaf):
        lo = dom.cell_lowers([leaf])[0]
        h = dom.cell_sizes([leaf])[0]
        return int(min(pmax, 1 + np.floor(pmax * float(np.linalg.norm(lo + 0.5 * h)) / np.sqrt(3.0))))

    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.ByLeafDegrees(degree)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, Diffusion())
    problem.add_cell_accumulator(u, Load())
    for axis in range(3):
        for side in (g.LOWER, g.UPPER):
            problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), exact))
    return problem, exact


def adaptive_leg(repeats=3, steps=20):
    """Adaptive hp workload: full-layout apply throughput, GMG-CG (FP32 and FP64
    Chebyshev V-cycles) and Jacobi-CG to 1e-8 from x0 = 0, relative L2 error
    of the GMG solution against u = r^(2/3)."""
    import ctypes

    import numpy as np
    import torch

    import gricoamd as g
    from gricoamd import _devabi, _native, ops
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid
    from gricoamd.solve import compile_preconditioner

    lib = _native.device()
    t0 = time.perf_counter()
    problem, exact = build_adaptive_problem()
    plan = g.compile(problem)
    t_plan = time.perf_counter() - t0
    dev = ops._ws(plan)
    log(dev.describe().strip())
    degrees = sorted({b.degrees[0] for b in plan.buckets})
    log(f"adaptive: host compile {t_plan:.1f}s, full dofs {plan.full_dof_count}, reduced {plan.reduced_dof_count}, "
        f"cells {sum(b.cell_count for b in plan.buckets)}, degrees {degrees}")
    n = plan.full_dof_count
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        _devabi.check(lib.gmf_apply_async(dev.ws, x.data_ptr(), y.data_ptr()))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        _devabi.check(lib.gmf_apply_async(dev.ws, x.data_ptr(), y.data_ptr()))
    torch.cuda.synchronize()
    apply_s = (time.perf_counter() - t) / steps
    b = ops.to_device(plan.reduced_rhs(), plan.reduced_dof_count)
    xs = torch.zeros_like(b)

    def run(handle, maxiter):
        info = _devabi.SolveInfo()
        torch.cuda.synchronize()
        t = time.perf_counter()
        _devabi.check(lib.gmf_cg_solve(dev.ws, handle, b.data_ptr(), xs.data_ptr(), 0, 1e-8, 0.0, maxiter,
                                       ctypes.byref(info)))
        torch.cuda.synchronize()
        return time.perf_counter() - t, info

    t0 = time.perf_counter()
    h64 = compile_geometric_multigrid(problem, GeometricMultigridPreconditioner(smoother="chebyshev"), fine_plan=plan)
    t_h = time.perf_counter() - t0
    out = {"workload": "adaptive hp Poisson, 16^3 base cells (4 uniform refinements of one root), 5 rounds of "
                       "anisotropic dyadic refinement at (0,0,0), degree 1..8 rising with distance, u = r^(2/3)",
           "full_dofs": n, "reduced_dofs": plan.reduced_dof_count,
           "buckets": [list(bk.degrees) + [bk.cell_count] for bk in plan.buckets],
           "kernel_tuning": dev.tuning(),
           "apply_gdofs": n / apply_s / 1e9, "apply_seconds": apply_s,
           "levels": h64.reduced_dof_counts(), "host_compile_seconds": t_plan, "hierarchy_seconds": t_h}
    counts = h64.reduced_dof_counts()
    keep = len(counts) - max(i for i, c in enumerate(counts) if c <= 4096)
    dense = h64.with_policy(GeometricMultigridPreconditioner(smoother="chebyshev", precision="fp32", max_levels=keep,
                                                             coarse_direct_dof_limit=4096))
    out["dense_coarse_levels"] = dense.reduced_dof_counts()
    for tag, h in (("gmg_cg_fp32_vcycle", h64.with_policy(GeometricMultigridPreconditioner(smoother="chebyshev",
                                                                                            precision="fp32"))),
                   ("gmg_cg", h64), ("gmg_cg_fp32_vcycle_dense_coarse", dense)):
        M = h.device_preconditioner()
        run(M.handle, 2000)
        times = []
        for _ in range(repeats):
            dt, info = run(M.handle, 2000)
            times.append(dt)
        full = ops.expand_reduced(plan, xs, include_shift=True).cpu().numpy()
        err = g.relative_l2_error(g.State(plan, full), exact)
        out[tag] = {"seconds": sorted(times)[len(times) // 2], "iterations": info.iterations,
                    "final_relative_residual": info.final_residual_norm / info.rhs_norm,
                    "relative_l2_error": err}
        log(f"adaptive {tag}: {out[tag]}")
    J = compile_preconditioner(g.JacobiPreconditioner(), plan)
    dt, info = run(J.handle, 50000)
    full = ops.expand_reduced(plan, xs, include_shift=True).cpu().numpy()
    out["jacobi_cg"] = {"seconds": dt, "iterations": info.iterations,
                        "final_relative_residual": info.final_residual_norm / info.rhs_norm,
                        "relative_l2_error": g.relative_l2_error(g.State(plan, full), exact)}
    log(f"adaptive jacobi_cg: {out['jacobi_cg']}")
    return out


def rhs_l2_leg(plan, reps=5):
    """Device right-hand side, reduced right-hand side and verification sums on
    the apply workload with inputs resident in HBM (loads lowered on the host
    once); roofline against the algorithmic bytes; the host sum-factorized
    right-hand side and the oracle (dense basis, CPU) beside it."""
    import ctypes

    import numpy as np
    import torch

    import pyoracle
    from gricoamd import _devabi, _native, ops

    lib = _native.device()
    dev = ops._ws(plan)
    t0 = time.perf_counter()
    loads, channels = plan.lowered_loads()
    t_lower = time.perf_counter() - t0
    lt = ops.to_device(loads)
    y = torch.empty(plan.full_dof_count, dtype=torch.float64, device="cuda")
    b = torch.empty(plan.reduced_dof_count, dtype=torch.float64, device="cuda")
    w = ops.to_device(plan.point_weights())
    e = ops.to_device(plan.point_values(lambda p: np.sin(np.pi * p[0]) * np.sin(np.pi * p[1]) * np.sin(np.pi * p[2])))
    u = torch.randn(plan.full_dof_count, dtype=torch.float64, device="cuda")
    sums = (ctypes.c_double * 2)()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    t_rhs = timed(lambda: _devabi.check(lib.gmf_rhs_physical(dev.ws, lt.data_ptr(), channels, y.data_ptr())))
    t_red = timed(lambda: _devabi.check(lib.gmf_reduced_rhs(dev.ws, y.data_ptr(), b.data_ptr())))
    t_l2 = timed(lambda: _devabi.check(lib.gmf_l2_error(dev.ws, u.data_ptr(), w.data_ptr(), e.data_ptr(), sums)))
    npts = int(lt.numel() // channels)
    rhs_bytes = 8 * npts * channels + 8 * plan.full_dof_count
    l2_bytes = 8 * plan.full_dof_count + 16 * npts
    plan._rhs = None  # time the host path from scratch
    t0 = time.perf_counter()
    host_rhs = plan.rhs_physical()
    t_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle_rhs = pyoracle.rhs_from_loads(plan, loads, channels)
    t_oracle = time.perf_counter() - t0
    y_np = y.cpu().numpy()
    err = float(np.abs(y_np - oracle_rhs).max() / np.abs(oracle_rhs).max())
    out = {
        "rhs_seconds": t_rhs, "rhs_roofline_frac": rhs_bytes / t_rhs / (HBM_PEAK_GBS * 1e9),
        "reduced_rhs_seconds": t_red, "l2_seconds": t_l2,
        "l2_roofline_frac": l2_bytes / t_l2 / (HBM_PEAK_GBS * 1e9),
        "host_lowering_seconds": t_lower, "host_sum_factorized_rhs_seconds": t_host,
        "oracle_dense_rhs_seconds": t_oracle, "rhs_max_relative_error_vs_oracle": err,
        "host_rhs_max_relative_difference": float(np.abs(y_np - host_rhs).max() / np.abs(host_rhs).max()),
    }
    log(f"rhs/l2: {out}")
    return out


class _SIPGCells:
    def cell_accumulate(self, q, trial, test_component):
        import gricoamd as g
        return g.TestChannels(0.0 * trial.value, tuple(trial.gradient))

    def cell_rhs_accumulate(self, q, test_component):
        return -8.0 + 0.0 * q.point[0]


class _SIPGFaces:
    """Symmetric interior penalty (benchmark/sine_poisson_sipg.jl:54-150):
    interfaces -{d_n u}[v] - {d_n v}[u] + sigma [u][v], boundary faces the
    one-sided analogue with Dirichlet data g = x^2 + 2y^2 - xy + z^2 - yz,
    sigma = 10 (p+1)^2 / h."""

    def __init__(self, degree):
        self.degree = degree

    def _sigma(self, q):
        return 10.0 * (self.degree + 1) ** 2 / q.cell_size[q.face_axis]

    @staticmethod
    def exact(p):
        return p[0] ** 2 + 2.0 * p[1] ** 2 - p[0] * p[1] + p[2] ** 2 - p[1] * p[2]

    def interface_accumulate(self, q, trial, test_component):
        import gricoamd as g
        jmp = g.jump(trial.value)
        avg = g.average(g.normal_component(trial.gradient, q.normal))
        sig = self._sigma(q)
        shared = tuple(0.5 * jmp * c for c in q.normal)
        return g.TraceTestChannels(g.TestChannels(-avg - sig * jmp, shared), g.TestChannels(avg + sig * jmp, shared))

    def boundary_accumulate(self, q, trial, test_component):
        import gricoamd as g
        dn = g.normal_component(trial.gradient, q.normal)
        return g.TestChannels(-dn + self._sigma(q) * trial.value, tuple(-trial.value * c for c in q.normal))

    def boundary_rhs_accumulate(self, q, test_component):
        import gricoamd as g
        gv = self.exact(q.point)
        return g.TestChannels(self._sigma(q) * gv, tuple(-gv * c for c in q.normal))


def build_sipg_problem(cells, degree, interfaces_only=False):
    """cells = 2^k per axis, built by k uniform refinements of one root (so the
    multigrid hierarchy has h-levels)."""
    import gricoamd as g
    depth = int(round(math.log2(cells)))
    if 2 ** depth != cells:
        raise ValueError("the SIPG workload needs a power-of-two cell count per axis")
    domain = g.Domain((0.0,) * 3, (1.0,) * 3, (1, 1, 1))
    g.uniform_refine(domain, depth)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree),
                                             continuity="dg"))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    faces = _SIPGFaces(degree)
    problem.add_interface_accumulator(u, u, faces)
    if interfaces_only:
        return problem
    problem.add_cell_accumulator(u, u, _SIPGCells())
    problem.add_cell_accumulator(u, _SIPGCells())
    for axis in range(3):
        for side in (g.LOWER, g.UPPER):
            problem.add_boundary_accumulator(g.BoundaryFace(axis, side), u, u, faces)
            problem.add_boundary_accumulator(g.BoundaryFace(axis, side), u, faces)
    return problem


def _rebuild_sipg_on(domain, degree):
    """The interface-only SIPG problem on an existing (refined) domain."""
    import gricoamd as g
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree),
                                             continuity="dg"))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_interface_accumulator(u, u, _SIPGFaces(degree))
    return problem


def sipg_leg(cells=32, degree=3, reps=20, cpu_cells=8):
    """Symmetric interior penalty DG Poisson (3D, DG p=3): device apply of the
    whole operator and of the interface items alone (two-sided point-cloud
    kernel), roofline of the interface kernel against its algorithmic bytes,
    and a device CG + Jacobi solve to 1e-8 that must reproduce the quadratic
    exact solution; the oracle apply on a small mesh timed beside it."""
    import numpy as np
    import torch

    import gricoamd as g
    import pyoracle
    from gricoamd import _devabi, _native, ops

    lib = _native.device()
    out = {"workload": f"3D SIPG DG Poisson, {cells}^3 cells, p={degree}, FullTensorBasis"}
    t0 = time.perf_counter()
    plan = g.compile(build_sipg_problem(cells, degree))
    iplan = g.compile(build_sipg_problem(cells, degree, interfaces_only=True))
    out["host_compile_seconds"] = time.perf_counter() - t0
    n = plan.full_dof_count
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)

    def timed(pl):
        dev = ops._ws(pl)
        _devabi.check(lib.gmf_apply_physical(dev.ws, x.data_ptr(), y.data_ptr()))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            _devabi.check(lib.gmf_apply_physical(dev.ws, x.data_ptr(), y.data_ptr()))
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    t_all = timed(plan)
    t_if = timed(iplan)
    # Face kernels alone (HIP events around the interface-only plan's cell pass).
    import ctypes
    cell_ms, total_ms = ctypes.c_double(), ctypes.c_double()
    _devabi.check(lib.gmf_apply_timed(ops._ws(iplan).ws, x.data_ptr(), y.data_ptr(), reps,
                                      ctypes.byref(cell_ms), ctypes.byref(total_ms)))
    t_if_kernels = cell_ms.value * 1e-3
    # The same interface items on the point-cloud kernel (no tensor detection).
    os.environ["GMF_FACE_TENSOR"] = "0"
    try:
        cplan = g.compile(build_sipg_problem(cells, degree, interfaces_only=True))
        ops._ws(cplan)
    finally:
        del os.environ["GMF_FACE_TENSOR"]
    reps_saved = reps
    reps = 3
    t_cloud = timed(cplan)
    reps = reps_saved
    del cplan
    items = sum(b.cell_count for b in iplan.interface_buckets)
    pts = sum(b.n_points_total for b in iplan.interface_buckets)
    ntot = iplan.interface_buckets[0].n_modes
    # Algorithmic bytes of the interface items: gather x and read-modify-write y
    # on both sides' modes, both sides' J^-1, and the lowered point action
    # restricted to its nonzero channel pairs (value and normal derivative of
    # both sides for this form: 16 of 64).
    nnz = max(int(np.count_nonzero(np.any(b.point_data != 0, axis=0))) for b in iplan.interface_buckets)
    if_bytes = items * (24 * ntot + 8 * 6) + pts * 8 * nnz
    out.update({"full_dofs": n, "interface_items": items, "interface_points": pts, "action_nonzeros": nnz,
                "apply_seconds": t_all, "apply_gdofs": n / t_all / 1e9,
                "interface_apply_seconds": t_if, "interface_kernel_seconds": t_if_kernels,
                "interface_algorithmic_bytes": if_bytes,
                "interface_apply_seconds_point_cloud_kernel": t_cloud,
                "interface_path": "face-tensor" if "face-tensor" in ops._ws(iplan).describe() else "point-cloud",
                "interface_roofline_frac": if_bytes / t_if_kernels / (HBM_PEAK_GBS * 1e9)})
    # Right-hand side on the host (face loads are not on the device path), then
    # Jacobi-CG and GMG-CG timed on the Krylov loop alone.
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid
    from gricoamd.solve import solve_reduced
    t0 = time.perf_counter()
    b = ops.to_device(plan.reduced_rhs(), plan.reduced_dof_count)
    out["host_rhs_seconds"] = time.perf_counter() - t0
    plan._loads = None
    t0 = time.perf_counter()
    bd = ops.reduced_rhs(plan, as_numpy=False)  # loads lowered on the host, projected on the device
    torch.cuda.synchronize()
    out["device_rhs_seconds_including_host_lowering"] = time.perf_counter() - t0
    out["device_vs_host_rhs_max_relative_difference"] = float(torch.max(torch.abs(bd - b)) / torch.max(torch.abs(b)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xj, info = solve_reduced(plan, g.CGSolver(preconditioner=g.JacobiPreconditioner()), b, 1e-8, 0.0, 50000)
    torch.cuda.synchronize()
    full = ops.expand_reduced(plan, xj, include_shift=True).cpu().numpy()
    out.update({"solve_seconds": time.perf_counter() - t0, "solve_iterations": int(info.iterations),
                "solve_relative_l2_error": g.relative_l2_error(g.State(plan, full), _SIPGFaces.exact)})
    # hp-multigrid on the DG space, faces rediscretized on every level.
    t0 = time.perf_counter()
    policy = GeometricMultigridPreconditioner(smoother="chebyshev")  # damped Jacobi diverges under the penalty
    h = compile_geometric_multigrid(build_sipg_problem(cells, degree), policy, fine_plan=plan)
    M = h.device_preconditioner()
    t_setup = time.perf_counter() - t0
    solve_reduced(plan, g.CGSolver(preconditioner=M), b, 1e-8, 0.0, 2000)  # warm-up (graph capture)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xg, info = solve_reduced(plan, g.CGSolver(preconditioner=M), b, 1e-8, 0.0, 2000)
    torch.cuda.synchronize()
    t_gmg = time.perf_counter() - t0
    full = ops.expand_reduced(plan, xg, include_shift=True).cpu().numpy()
    out.update({"gmg_cg_seconds": t_gmg, "gmg_cg_iterations": int(info.iterations),
                "gmg_setup_seconds": t_setup, "gmg_levels": h.reduced_dof_counts(),
                "gmg_cg_relative_l2_error": g.relative_l2_error(g.State(plan, full), _SIPGFaces.exact)})
    counts = h.reduced_dof_counts()
    keep = len(counts) - max(i for i, c in enumerate(counts) if c <= 4096)
    ht = h.with_policy(GeometricMultigridPreconditioner(smoother="chebyshev", max_levels=keep,
                                                        coarse_direct_dof_limit=4096))
    Mt = ht.device_preconditioner()
    solve_reduced(plan, g.CGSolver(preconditioner=Mt), b, 1e-8, 0.0, 2000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = solve_reduced(plan, g.CGSolver(preconditioner=Mt), b, 1e-8, 0.0, 2000)
    torch.cuda.synchronize()
    out.update({"gmg_cg_dense_coarse_seconds": time.perf_counter() - t0,
                "gmg_cg_dense_coarse_iterations": int(info.iterations),
                "gmg_dense_coarse_levels": ht.reduced_dof_counts()})
    # One level per 2x2x2 coarsening (h_steps_per_level = 3) instead of one per
    # binary merge, dense coarse level: far fewer latency-bound DG levels.
    t0 = time.perf_counter()
    pol3 = GeometricMultigridPreconditioner(smoother="chebyshev", h_steps_per_level=3, chebyshev_degree=3,
                                            coarse_direct_dof_limit=4096)
    h3 = compile_geometric_multigrid(build_sipg_problem(cells, degree), pol3, fine_plan=plan)
    counts3 = h3.reduced_dof_counts()
    keep3 = len(counts3) - max(i for i, c in enumerate(counts3) if c <= 4096)
    h3 = h3.with_policy(GeometricMultigridPreconditioner(smoother="chebyshev", h_steps_per_level=3,
                                                         chebyshev_degree=3, max_levels=keep3,
                                                         coarse_direct_dof_limit=4096))
    M3 = h3.device_preconditioner()
    t_setup3 = time.perf_counter() - t0
    solve_reduced(plan, g.CGSolver(preconditioner=M3), b, 1e-8, 0.0, 2000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x3, info = solve_reduced(plan, g.CGSolver(preconditioner=M3), b, 1e-8, 0.0, 2000)
    torch.cuda.synchronize()
    t3 = time.perf_counter() - t0
    full = ops.expand_reduced(plan, x3, include_shift=True).cpu().numpy()
    out.update({"gmg_cg_merged_levels_seconds": t3, "gmg_cg_merged_levels_iterations": int(info.iterations),
                "gmg_merged_levels": h3.reduced_dof_counts(), "gmg_merged_setup_seconds": t_setup3,
                "gmg_cg_merged_levels_relative_l2_error": g.relative_l2_error(g.State(plan, full), _SIPGFaces.exact),
                "gmg_cg_merged_levels_speedup_vs_jacobi_cg": out["solve_seconds"] / t3})
    log(f"sipg GMG-CG (one level per 2x2x2 coarsening): {int(info.iterations)} iterations, {t3:.3f} s "
        f"(Jacobi-CG {out['solve_seconds']:.3f} s)")
    # Hanging interfaces: the same operator on 16^3 cells with the cells of the
    # octant [0, 1/2]^3 refined once in every axis.
    hang = build_sipg_problem(16, degree, interfaces_only=True)
    dom = hang.field.space.domain
    cells = [int(c) for c in dom.active_leaves()]
    centers, _ = dom.cell_centers_and_half_sizes(cells)
    for c, ctr in zip(cells, centers):
        if np.all(ctr < 0.5):
            kids = [c]
            for axis in range(3):
                kids = [k2 for k in kids for k2 in (lambda f: (f, f + 1))(dom.refine(k, axis))]
    hang = _rebuild_sipg_on(hang.field.space.domain, degree)
    t_h = {}
    xh = None
    for tag, env in (("face_tensor", None), ("point_cloud_kernel", "0")):
        if env is not None:
            os.environ["GMF_FACE_TENSOR"] = env
        try:
            hp = g.compile(hang)
            ops._ws(hp)
        finally:
            os.environ.pop("GMF_FACE_TENSOR", None)
        cm, tm = ctypes.c_double(), ctypes.c_double()
        if xh is None:  # one input for both paths
            gen = torch.Generator(device="cuda")
            gen.manual_seed(7)
            xh = torch.randn(hp.full_dof_count, dtype=torch.float64, device="cuda", generator=gen)
        yh = torch.empty_like(xh)
        _devabi.check(lib.gmf_apply_timed(ops._ws(hp).ws, xh.data_ptr(), yh.data_ptr(), 5, ctypes.byref(cm),
                                          ctypes.byref(tm)))
        t_h[tag] = cm.value * 1e-3
        if tag == "face_tensor":
            hang_items = sum(b.cell_count for b in hp.interface_buckets)
            hang_batches = len(hp.interface_buckets)
            yt = yh.cpu().numpy()
        else:
            hang_diff = float(np.abs(yh.cpu().numpy() - yt).max() / np.abs(yt).max())
        del hp
    out.update({"hanging_interface_items": hang_items, "hanging_interface_batches": hang_batches,
                "hanging_interface_kernel_seconds": t_h["face_tensor"],
                "hanging_interface_kernel_seconds_point_cloud": t_h["point_cloud_kernel"],
                "hanging_tensor_vs_point_cloud_max_relative_difference": hang_diff})
    small = g.compile(build_sipg_problem(cpu_cells, degree))
    xv = np.random.default_rng(5).standard_normal(small.full_dof_count)
    t0 = time.perf_counter()
    yo = pyoracle.apply(small, xv, physical=True)
    t_cpu = time.perf_counter() - t0
    yd = g.apply_physical(small, xv)
    out.update({"oracle_apply_seconds_small": t_cpu, "oracle_small_cells": cpu_cells ** 3,
                "device_vs_oracle_max_relative_error_small": float(np.abs(yd - yo).max() / np.abs(yo).max())})
    log(f"sipg: {out}")
    return out


class _Bratu:
    """examples/nonlinear_bratu/driver.jl: residual (lam e^u - f, grad u),
    tangent (lam e^u du, grad du), u_e = A sin(pi x) sin(pi y)."""

    def __init__(self, lam=1.0, amplitude=0.2):
        self.lam, self.amplitude = lam, amplitude

    def exact(self, p):
        import numpy as np
        return self.amplitude * np.sin(np.pi * p[0]) * np.sin(np.pi * p[1])

    def cell_residual_accumulate(self, q, state, test_component):
        import numpy as np

        import gricoamd as g
        e = self.exact(q.point)
        f = 2.0 * np.pi ** 2 * e + self.lam * np.exp(e)
        return g.TestChannels(self.lam * np.exp(state.value) - f, tuple(state.gradient))

    def cell_tangent_accumulate(self, q, state, increment, test_component):
        import numpy as np

        import gricoamd as g
        return g.TestChannels(self.lam * np.exp(state.value) * increment.value, tuple(increment.gradient))


def newton_leg(cells=256, degree=3):
    """Nonlinear Bratu example (2D, CG p=3, zero Dirichlet) at 256^2 cells:
    device Newton (residual from host-evaluated point loads projected on the
    device, tangent re-lowered per step, unpreconditioned device CG as the
    reference's default tangent solve); the oracle residual on the same mesh
    timed beside it."""
    import numpy as np

    import gricoamd as g
    import pyoracle
    from gricoamd import nonlinear as NL

    op = _Bratu()
    domain = g.Domain((0.0, 0.0), (1.0, 1.0), (cells, cells))
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    problem = NL.ResidualProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, op)
    for axis in range(2):
        for side in (g.LOWER, g.UPPER):
            problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    t0 = time.perf_counter()
    rplan = NL.ResidualPlan(problem)
    t_compile = time.perf_counter() - t0
    NL.newton_solve(NL.ResidualPlan(problem), maxiter=1, relative_tolerance=1.0, absolute_tolerance=1e300)  # warm-up
    t0 = time.perf_counter()
    state = NL.newton_solve(rplan, maxiter=10, relative_tolerance=1e-10, absolute_tolerance=1e-10,
                            linear_relative_tolerance=1e-10)
    t_newton = time.perf_counter() - t0
    x = 0.5 * state.coefficients  # away from the solution, where the residual is not round-off
    t0 = time.perf_counter()
    r_dev = NL.residual(rplan, x)
    t_res = time.perf_counter() - t0
    t0 = time.perf_counter()
    r_or = pyoracle.residual(rplan, x)
    t_or = time.perf_counter() - t0
    scale = max(np.abs(r_or).max(), 1e-300)
    out = {"workload": f"2D Bratu (lambda=1, A=0.2), {cells}^2 cells, p={degree}, zero Dirichlet",
           "full_dofs": rplan.full_dof_count, "host_compile_seconds": t_compile,
           "newton_seconds": t_newton, "newton_iterations": state.info["iterations"],
           "residual_norms": state.info["residual_norms"],
           "l2_error": g.l2_error(state, op.exact),
           "device_residual_seconds": t_res, "oracle_residual_seconds": t_or,
           "residual_max_difference_vs_oracle_scaled": float(np.abs(r_dev - r_or).max() / scale)}
    log(f"newton: {out}")
    return out


def build_lognormal_problem(depth=5, degree=3, seed=2024):
    """32^3 cells (five uniform refinements of one root), p = 3, kappa = exp(g)
    with g a Gaussian random field (squared-exponential covariance, correlation
    length 0.1, sigma = 1) from 256 seeded random Fourier features, unit load,
    zero Dirichlet."""
    import numpy as np

    import gricoamd as g

    rng = np.random.default_rng(seed)
    w = rng.standard_normal((256, 3)) / 0.1
    phi = rng.uniform(0.0, 2.0 * np.pi, size=256)
    amp = np.sqrt(2.0 / 256)

    def kappa(x, y, z):
        pts = np.stack([np.asarray(x).reshape(-1), np.asarray(y).reshape(-1), np.asarray(z).reshape(-1)], axis=1)
        return np.exp(amp * np.cos(pts @ w.T + phi).sum(axis=1)).reshape(np.shape(x))

    class Diffusion:
        def cell_accumulate(self, q, trial, test_component):
            k = kappa(*q.point)
            return g.TestChannels(0.0 * trial.value, tuple(k * c for c in trial.gradient))

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            return 1.0 + 0.0 * q.point[0]

    domain = g.Domain((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    g.uniform_refine(domain, depth)
    space = g.HpSpace(domain, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    problem = g.AffineProblem(u, g.SPD())
    problem.add_cell_accumulator(u, u, Diffusion())
    problem.add_cell_accumulator(u, Load())
    for axis in range(3):
        for side in (g.LOWER, g.UPPER):
            problem.add_constraint(g.Dirichlet(u, g.BoundaryFace(axis, side), 0.0))
    return problem


def lognormal_leg(repeats=3, depth=5, learned=True, pairs=2000, steps=3000):
    """The log-normal p = 3 problem (32^3 cells at depth 5) from x0 = 0 to
    ||r|| <= 1e-8 ||b||, hierarchy and network already built: GMG-CG with FP32
    and FP64 Chebyshev V-cycles, Jacobi-CG, and FGMRES with the learned block
    correction around the same FP32 V-cycle (trained on `pairs` GMG-CG
    residuals, `steps` Adam steps; training reported separately), against
    GMG-CG. Every solution's residual is re-evaluated with the FP64 oracle
    operator. The V-cycle (degree 3, range 50, one h-level per full 2x2x2
    coarsening, hierarchy cut at the first level of at most 4096 dofs and
    solved there by a device dense inverse) is the fastest of a sweep over
    degree, range, p-sequence and h-steps per level on this problem."""
    import ctypes

    import torch

    import gricoamd as g
    from gricoamd import _devabi, _native, ops
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid
    from gricoamd.solve import compile_preconditioner

    lib = _native.device()
    t0 = time.perf_counter()
    problem = build_lognormal_problem(depth=depth)
    plan = g.compile(problem)
    t_compile = time.perf_counter() - t0
    log(f"log-normal problem: compile {t_compile:.1f}s, reduced dofs {plan.reduced_dof_count}")
    t0 = time.perf_counter()
    hierarchy = compile_geometric_multigrid(problem, GeometricMultigridPreconditioner(smoother="chebyshev",
                                                                                      h_steps_per_level=3),
                                            fine_plan=plan)
    t_host = time.perf_counter() - t0
    counts = hierarchy.reduced_dof_counts()
    keep = len(counts) - max([i for i, c in enumerate(counts) if c <= 4096] or [0])
    b, x, timed, cg = _device_solver_kit(plan, repeats)
    b_host = b.cpu().numpy()
    out = {"workload": f"3D Poisson {2 ** depth}^3 cells (uniform refinement of one root), p=3, log-normal kappa "
                       "(corr. length 0.1, sigma 1, seed 2024), unit load, zero Dirichlet",
           "reduced_dofs": plan.reduced_dof_count, "host_compile_seconds": t_compile,
           "hierarchy_host_seconds": t_host}
    sols = {}
    policies = {}
    for tag, prec in (("gmg_cg", "fp32"), ("gmg_cg_fp64_vcycle", "fp64")):
        pol = GeometricMultigridPreconditioner(smoother="chebyshev", precision=prec, chebyshev_degree=3,
                                               chebyshev_range=50.0, h_steps_per_level=3, max_levels=keep,
                                               coarse_direct_dof_limit=4096)
        policies[tag] = pol
        h = hierarchy.with_policy(pol)
        M = h.device_preconditioner()
        res = timed(lambda: cg(M.handle))
        res["vcycle_seconds"] = _vcycle_seconds(M, plan.reduced_dof_count)
        res["policy"] = gmg_policy_record(pol, h.reduced_dof_counts())
        sols[tag] = x.clone()
        out[tag] = res
        log(f"C4 {tag}: {res}")
        del M
    J = compile_preconditioner(g.JacobiPreconditioner(), plan)
    out["jacobi_cg"] = timed(lambda: cg(J.handle))
    sols["jacobi_cg"] = x.clone()
    out["gmg_cg_speedup_vs_jacobi_cg"] = out["jacobi_cg"]["seconds"] / out["gmg_cg"]["seconds"]
    if learned:
        from gricoamd.learned import train_block_correction

        h = hierarchy.with_policy(policies["gmg_cg"])
        net, report, M, V = train_block_correction(problem, plan=plan, pairs=pairs, steps=steps, hierarchy=h, log=log)

        def fgmres(handle):
            info = _devabi.SolveInfo()
            status = lib.gmf_fgmres_solve(ops._ws(plan).ws, handle, b.data_ptr(), x.data_ptr(), 0, 30, 1e-8, 0.0,
                                          5000, ctypes.byref(info))
            if status == -3:  # GMF_ENOCONV
                return None
            _devabi.check(status)
            return info

        first = fgmres(M.handle)
        if first is None:
            out["learned_fgmres"] = {"converged": False}
        else:
            out["learned_fgmres"] = timed(lambda: fgmres(M.handle))
            sols["learned_fgmres"] = x.clone()
            out["learned_speedup_vs_gmg_cg"] = out["gmg_cg"]["seconds"] / out["learned_fgmres"]["seconds"]
        out["fgmres_vcycle_only"] = timed(lambda: fgmres(V.handle))
        sols["fgmres_vcycle_only"] = x.clone()
        out["learned_target_speedup"] = 2.0
        out["learned_training"] = report
        out["learned_network"] = "per-cell MLP 64-256-256-64, BF16 weights on MFMA, FP32 accumulation"
        del M, V, net
    nrm = float(torch.linalg.norm(sols["jacobi_cg"]))
    out["solution_relative_difference_vs_jacobi_cg"] = {
        k: float(torch.linalg.norm(v - sols["jacobi_cg"])) / nrm for k, v in sols.items() if k != "jacobi_cg"}
    t0 = time.perf_counter()
    out["oracle_relative_residual"] = oracle_relative_residuals(plan, sols, b_host)
    out["oracle_check_seconds"] = time.perf_counter() - t0
    log(f"log-normal leg: {out}")
    return out


# Sources of the cell-pass kernels (brick kernel, halo pass, their stores and
# one-dimensional sweeps) whose HBM traffic profiles/apply_traffic.json records.
TRAFFIC_SOURCES = ("grico.jl_amd/csrc/device/cell_fast.hip", "grico.jl_amd/csrc/device/eo.hpp")


def kernel_source_hash():
    """First 16 hex digits of the SHA-256 over the cell-pass kernel sources."""
    import hashlib
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _traffic_from_profiles(n, cells):
    """HBM bytes per cell pass measured with rocprofv3 PMC counters (FETCH_SIZE +
    WRITE_SIZE, scaled by a 1 GiB copy), recorded in profiles/apply_traffic.json
    for this workload and these kernel sources; None when absent, for another
    workload, or when the kernel sources changed since the measurement."""
    path = os.path.join(ROOT, "profiles", "apply_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("full_dofs") != n or rec.get("cells") != cells:
        return None
    if rec.get("kernel_source_sha256_16") != kernel_source_hash():
        return None
    return rec.get("traffic_bytes_per_cell_pass")


def build_fcm_problem(cells, degree, depth, alpha=1e-3):
    """Finite-cell Poisson on the unit ball embedded in [-1.2, 1.2]^3: cells^3
    background leaves of a PhysicalDomain (cut-cell rules moment fitted by
    compile() on every leaf set it sees, subdivision depth `depth`, finite-cell
    extension alpha), Nitsche (symmetric, gamma = 10 (p + 1)^2 / h) on a sampled
    sphere, load 6 (exact solution 1 - r^2). alpha = 1e-3 keeps the Nitsche form
    positive definite (with 1e-5 the 8^3 ball already has negative eigenvalues
    on small cuts, so CG has no guarantee)."""
    import gricoamd as g
    from gricoamd.embedded import FiniteCellExtension, ImplicitRegion, PhysicalDomain, sphere_surface

    class Diffusion:
        def cell_accumulate(self, q, trial, test_component):
            return g.TestChannels(0.0 * trial.value, tuple(trial.gradient))

    class Load:
        def cell_rhs_accumulate(self, q, test_component):
            return 6.0 + 0.0 * q.point[0]

    class Nitsche:
        def __init__(self, gamma):
            self.gamma = gamma

        def surface_accumulate(self, q, trial, test_component):
            n = q.normal
            dn = sum(nc * gc for nc, gc in zip(n, trial.gradient))
            return g.TestChannels(-dn + self.gamma / q.cell_size[0] * trial.value,
                                  tuple(-trial.value * nc for nc in n))

    region = ImplicitRegion(lambda p: p[0] ** 2 + p[1] ** 2 + p[2] ** 2 - 1.0, subdivision_depth=depth)
    # cells = roots * 2^k with roots odd: roots^3 root cells refined k times, so
    # that multigrid finds h-levels (the same leaves as cells^3 root cells).
    roots, k = cells, 0
    while roots % 2 == 0:
        roots, k = roots // 2, k + 1
    dom = g.Domain((-1.2, -1.2, -1.2), (2.4, 2.4, 2.4), (roots, roots, roots))
    g.uniform_refine(dom, k)
    pdom = PhysicalDomain(dom, region, FiniteCellExtension(alpha))
    last = [0.0]

    def progress(done, total):
        if time.perf_counter() - last[0] > 20.0 or done == total:
            last[0] = time.perf_counter()
            log(f"cut-cell moment fitting: {done}/{total} cells")

    pdom.fit_progress = progress
    space = g.HpSpace(pdom, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    pr = g.AffineProblem(u, g.SPD())
    pr.add_cell_accumulator(u, u, Diffusion())
    pr.add_cell_accumulator(u, Load())
    pr.add_embedded_surface(sphere_surface((0.0, 0.0, 0.0), 1.0, n_theta=4 * cells, n_phi=8 * cells))
    pr.add_surface_accumulator(u, u, Nitsche(10.0 * (degree + 1) ** 2))
    return pr


def fcm_leg(cells=48, degree=3, depth=3, nrhs=64, maxiter=2000, sequential=2, oracle_cells=4, precision="fp64",
            max_levels=3):
    """Finite-cell sphere workload with nrhs seeded N(0,1) reduced right-hand
    sides solved to 1e-8 by block CG (FP64 Krylov, lockstep columns) with the
    hp-multigrid preconditioner on physical-domain levels (cut-cell rules
    rediscretized per level, Chebyshev smoothing over cell blocks on the cut
    cells); the block apply against the one-column apply; `sequential` columns
    solved one at a time; the physical load's solution against 1 - r^2; the
    block apply against the oracle on a small ball."""
    import numpy as np
    import torch

    import gricoamd as g
    import pyoracle
    from gricoamd import _devabi, _native, ops
    from gricoamd.multigrid import GeometricMultigridPreconditioner, compile_geometric_multigrid

    lib = _native.device()
    out = {"workload": f"finite-cell Poisson, unit ball in [-1.2,1.2]^3, {cells}^3 background cells, p={degree}, "
                       f"PhysicalDomain cut cells subdivided to depth {depth} and moment fitted, finite-cell "
                       f"extension alpha 1e-3, Nitsche on a sampled sphere; {nrhs} seeded N(0,1) right-hand sides, "
                       f"block CG to 1e-8 with the hp-multigrid preconditioner ({precision} V-cycle)"}
    t0 = time.perf_counter()
    problem = build_fcm_problem(cells, degree, depth)
    plan = g.compile(problem)
    dev = ops._ws(plan)
    torch.cuda.synchronize()
    out["host_compile_seconds"] = time.perf_counter() - t0
    n = plan.reduced_dof_count
    cut = plan.cloud_buckets
    out.update({"reduced_dofs": n, "tensor_cells": sum(b.cell_count for b in plan.buckets),
                "cut_cells": sum(b.cell_count for b in cut),
                "cut_cell_rules": "moment-fitted (src/regions.jl:519-708) plus alpha times the tensor rule",
                "cut_cell_points": int(sum(b.n_points_total for b in cut)),
                "kernel_tuning": [t for t in dev.tuning() if t["kernel"] != "empty"],
                "surface_pieces": sum(b.cell_count for b in plan.surface_buckets),
                "surface_points": int(sum(b.n_points_total for b in plan.surface_buckets))})
    log(f"fcm: {out}")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    B = torch.randn((nrhs, n), dtype=torch.float64, device="cuda", generator=gen)  # column j = B[j]
    Y = torch.empty_like(B)
    y1 = torch.empty(n, dtype=torch.float64, device="cuda")

    def clock(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    t_one = clock(lambda: _devabi.check(lib.gmf_reduced_apply(dev.ws, B[0].data_ptr(), y1.data_ptr())), 5)
    t_block = clock(lambda: _devabi.check(lib.gmf_reduced_apply_block(dev.ws, B.data_ptr(), Y.data_ptr(), nrhs)), 2)
    _devabi.check(lib.gmf_reduced_apply(dev.ws, B[nrhs - 1].data_ptr(), y1.data_ptr()))
    torch.cuda.synchronize()
    out.update({"apply_one_column_seconds": t_one, "apply_block_seconds": t_block,
                "apply_block_seconds_per_column": t_block / nrhs, "apply_block_speedup_per_column": t_one / (t_block / nrhs),
                "block_vs_one_column_max_relative_difference":
                    float(torch.max(torch.abs(Y[nrhs - 1] - y1)) / torch.max(torch.abs(y1)))})
    log(f"fcm apply: one column {t_one * 1e3:.2f} ms, block {t_block / nrhs * 1e3:.2f} ms per column")
    del Y
    # hp-multigrid on physical-domain levels with cell-block smoothing.
    t0 = time.perf_counter()
    pol = GeometricMultigridPreconditioner(smoother="chebyshev", cell_blocks="cut", h_steps_per_level=3,
                                           chebyshev_degree=3, precision=precision, coarse_direct_dof_limit=40000,
                                           max_levels=max_levels)
    h = compile_geometric_multigrid(problem, pol, fine_plan=plan)
    M = h.device_preconditioner()
    torch.cuda.synchronize()
    out.update({"hierarchy_seconds": time.perf_counter() - t0, "levels": h.reduced_dof_counts(),
                "cut_blocks_per_level": [sum(b.cell_count for b in p.cloud_buckets) for p in h.plans],
                "block_colors_per_level": getattr(h, "block_colors", None), "coarse_solver": h.coarse_solver_kind()})
    log(f"fcm hierarchy: {h.reduced_dof_counts()} in {out['hierarchy_seconds']:.1f}s")
    X = torch.zeros_like(B)
    infos = (_devabi.SolveInfo * nrhs)()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    status = lib.gmf_cg_solve_block(dev.ws, M.handle, B.data_ptr(), X.data_ptr(), nrhs, 0, 1e-8, 0.0, maxiter, infos)
    torch.cuda.synchronize()
    t_bsolve = time.perf_counter() - t0
    _devabi.check(status)
    its = [int(i.iterations) for i in infos]
    # true residuals of the block solution
    R = torch.empty_like(B)
    _devabi.check(lib.gmf_reduced_apply_block(dev.ws, X.data_ptr(), R.data_ptr(), nrhs))
    rel_true = (torch.linalg.norm(B - R, dim=1) / torch.linalg.norm(B, dim=1)).cpu().numpy()
    out.update({"gmg_block_cg_status": int(status), "gmg_block_cg_seconds": t_bsolve,
                "gmg_block_cg_seconds_per_rhs": t_bsolve / nrhs,
                "gmg_block_cg_iterations_max": max(its), "gmg_block_cg_iterations_mean": float(np.mean(its)),
                "all_rhs_converged": bool(all(int(i.converged) == 1 for i in infos)),
                "true_relative_residual_max": float(rel_true.max())})
    log(f"fcm GMG block CG: {t_bsolve:.2f} s for {nrhs} rhs, iterations max {max(its)}, "
        f"true residual max {rel_true.max():.2e}")
    del R
    seq = []
    for j in range(sequential):
        info = _devabi.SolveInfo()
        x1 = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _devabi.check(lib.gmf_cg_solve(dev.ws, M.handle, B[j].data_ptr(), x1.data_ptr(), 0, 1e-8, 0.0, maxiter,
                                       ctypes_byref(info)))
        torch.cuda.synchronize()
        seq.append((time.perf_counter() - t0, int(info.iterations),
                    float(torch.max(torch.abs(x1 - X[j])) / torch.max(torch.abs(x1)))))
    out.update({"one_column_gmg_cg_seconds": float(np.mean([s_[0] for s_ in seq])),
                "one_column_gmg_cg_iterations": [s_[1] for s_ in seq],
                "one_column_vs_block_solution_max_relative_difference": max(s_[2] for s_ in seq)})
    del X, B
    # the physical load (exact solution 1 - r^2): right-hand side on the host
    # (lowering + projection) and on the device (host lowering of the point
    # loads, gmf_rhs_physical + gmf_rhs_faces)
    plan._rhs = None
    t0 = time.perf_counter()
    host_rhs = plan.rhs_physical()
    t_host_rhs = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev_rhs = ops.rhs_physical(plan, as_numpy=False)
    torch.cuda.synchronize()
    t_dev_rhs = time.perf_counter() - t0
    tl, ch = plan.lowered_loads()
    fl = torch.as_tensor(plan.face_loads(), device="cuda")
    tl = torch.as_tensor(tl, device="cuda")
    y = torch.zeros(plan.full_dof_count, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if tl.numel():
        _devabi.check(lib.gmf_rhs_physical(dev.ws, tl.data_ptr(), int(ch), y.data_ptr()))
    _devabi.check(lib.gmf_rhs_faces(dev.ws, fl.data_ptr(), y.data_ptr()))
    torch.cuda.synchronize()
    t_dev_kernels = time.perf_counter() - t0
    out.update({"rhs_host_seconds": t_host_rhs, "rhs_device_seconds_with_host_lowering": t_dev_rhs,
                "rhs_device_kernel_seconds": t_dev_kernels,
                "rhs_device_vs_host_max_relative_difference":
                    float(np.abs(dev_rhs.cpu().numpy() - host_rhs).max() / np.abs(host_rhs).max())})
    log(f"fcm right-hand side: host {t_host_rhs:.2f} s, device {t_dev_rhs:.2f} s with lowering, "
        f"kernels {t_dev_kernels * 1e3:.2f} ms")
    b = ops.reduced_rhs(plan, as_numpy=False)
    xs = torch.zeros(n, dtype=torch.float64, device="cuda")
    info = _devabi.SolveInfo()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _devabi.check(lib.gmf_cg_solve(dev.ws, M.handle, b.data_ptr(), xs.data_ptr(), 0, 1e-8, 0.0, maxiter,
                                   ctypes_byref(info)))
    torch.cuda.synchronize()
    t_phys = time.perf_counter() - t0
    state = g.State(plan, ops.expand_reduced(plan, xs, include_shift=True).cpu().numpy())
    from gricoamd.verification import device_relative_l2_error
    exact = lambda p: 1.0 - (p[0] ** 2 + p[1] ** 2 + p[2] ** 2)  # noqa: E731
    t0 = time.perf_counter()
    l2_host = g.relative_l2_error(state, exact)
    t_l2_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    l2_dev = device_relative_l2_error(state, exact)
    t_l2_dev = time.perf_counter() - t0
    out.update({"physical_load_seconds": t_phys, "physical_load_iterations": int(info.iterations),
                "physical_load_relative_l2_error_vs_1_minus_r2": l2_dev,
                "l2_host_seconds": t_l2_host, "l2_device_seconds_with_host_point_values": t_l2_dev,
                "l2_device_vs_host_relative_difference": abs(l2_dev - l2_host) / l2_host})
    log(f"fcm physical load: {int(info.iterations)} iterations, {t_phys:.2f} s, "
        f"L2 {out['physical_load_relative_l2_error_vs_1_minus_r2']:.2e}")
    # The block apply against the oracle on a small ball of the same kind.
    small = g.compile(build_fcm_problem(oracle_cells, degree, depth))
    Xo = np.random.default_rng(9).standard_normal((small.reduced_dof_count, 3))
    Yo = g.reduced_apply_block(small, Xo).cpu().numpy()
    ref = np.stack([pyoracle.reduced_apply(small, Xo[:, j]) for j in range(3)], axis=1)
    out["oracle_max_relative_error"] = float(np.abs(Yo - ref).max() / np.abs(ref).max())
    out["oracle_sample"] = f"{oracle_cells}^3 background cells, same degree and depth, 3 columns"
    return out


def ctypes_byref(obj):
    import ctypes
    return ctypes.byref(obj)


def transfer_leg(roots=16, degree=2):
    """State transfer across an hp transition (3D CG, roots^3 cells at p=2;
    the corner octant, split once along x in the source, is merged back onto
    composite rules, the opposite octant is split along x and the middle
    layer is raised to p=3): host lowering of the L2 projection, device
    Jacobi-CG of the global projection to 1e-12 (transfer_state with a
    linear_solve), the oracle Jacobi-CG of the same system timed beside it,
    and the default local-projection transfer (per-cell solves reconciled on
    shared dofs, host) with its distance from the global projection."""
    import numpy as np

    import gricoamd as g
    import pyoracle
    from gricoamd import adaptivity as ad

    class Mass:
        def cell_accumulate(self, q, trial, test_component):
            return g.TestChannels(trial.value, tuple(0.0 * x for x in trial.gradient))

    dom = g.Domain((0.0,) * 3, (1.0,) * 3, (roots,) * 3)
    half = roots // 2
    crd = dom.grid.coords()
    corner = [int(c) for c in dom.active_leaves() if np.all(crd[c] < half)]
    for c in corner:
        dom.refine(c, 0)
    space = g.HpSpace(dom, g.SpaceOptions(basis=g.FullTensorBasis(), degree=g.UniformDegree(degree)))
    u = g.ScalarField(space, "u")
    src_problem = g.AffineProblem(u, g.SPD())
    src_problem.add_cell_accumulator(u, u, Mass())
    src = g.State(g.compile(src_problem), np.random.default_rng(7).standard_normal(space.scalar_dof_count()))
    plan = ad.AdaptivityPlan(space)
    for c in corner:
        plan.request_h_derefinement(c, 0)
    crd = plan.target_domain.grid.coords()
    for leaf in list(plan.degrees):
        lv = plan.target_domain.grid.levels()[leaf]
        if np.all(lv == 0) and np.all(crd[leaf] >= half):
            plan.request_h_refinement(leaf, 0)
        elif np.all(lv == 0) and crd[leaf][2] == half - 1:
            plan.request_p_refinement(leaf)
    t0 = time.perf_counter()
    tr = ad.transition(plan)
    t_space = time.perf_counter() - t0
    t0 = time.perf_counter()
    pplan = g.compile(ad.projection_problem(tr, src))
    b = pplan.rhs_physical()[pplan.solve_dofs]
    t_lower = time.perf_counter() - t0
    solver = g.CGSolver(preconditioner=g.JacobiPreconditioner())
    g.solve(pplan, solver=solver, relative_tolerance=1e-12)  # warm-up
    t0 = time.perf_counter()
    state = g.solve(pplan, solver=solver, relative_tolerance=1e-12)
    t_dev = time.perf_counter() - t0
    inv = 1.0 / pyoracle.diagonal_physical(pplan)[pplan.solve_dofs]
    oracle_iters = 100  # bounded sample: the full oracle solve takes ~35 s
    t0 = time.perf_counter()
    _, it_or = pyoracle.cg(pplan, b, inv_diag=inv, rtol=1e-12, maxiter=oracle_iters, lowered=False)
    t_or = time.perf_counter() - t0
    it_or = oracle_iters if it_or < 0 else it_or  # -1: stopped at maxiter (not converged in the sample)
    t0 = time.perf_counter()
    local = ad.reconcile(pplan, ad.local_projections(pplan, src))  # the default (no linear_solve) transfer
    t_local = time.perf_counter() - t0
    glob = state.coefficients  # full layout on the target space
    xr = np.random.default_rng(3).standard_normal(pplan.full_dof_count)
    y_dev = g.apply_physical(pplan, xr)
    y_or = pyoracle.apply(pplan, xr, physical=True)
    out = {"workload": f"3D CG p={degree}, {roots}^3 roots; corner octant merged (composite rules), opposite "
                       f"octant split along x, one layer raised to p={degree + 1}; L2 projection of a random field",
           "source_dofs": space.scalar_dof_count(), "target_dofs": pplan.full_dof_count,
           "merged_cells": len(corner), "target_space_seconds": t_space, "host_lowering_seconds": t_lower,
           "device_solve_seconds": t_dev, "device_iterations": state.info.iterations,
           "device_seconds_per_iteration": t_dev / max(state.info.iterations, 1),
           "oracle_cg_seconds_per_iteration": t_or / max(it_or, 1), "oracle_sample_iterations": it_or,
           "local_projection_host_seconds": t_local,
           "local_vs_global_projection_max_difference_scaled": float(np.abs(local - glob).max() /
                                                                     np.abs(glob).max()),
           "projection_apply_max_difference_vs_oracle_scaled": float(np.abs(y_dev - y_or).max() /
                                                                     np.abs(y_or).max())}
    log(f"transfer: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--depth", type=int, default=6, help="uniform refinements of the root (6 -> 64^3 cells)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the GMG-CG / Jacobi-CG solve leg")
    ap.add_argument("--no-adaptive", action="store_true", help="skip the adaptive hp leg")
    ap.add_argument("--no-lognormal", action="store_true", help="skip the log-normal GMG-CG leg")
    ap.add_argument("--no-sipg", action="store_true", help="skip the interior penalty DG (face kernels) leg")
    ap.add_argument("--no-newton", action="store_true", help="skip the nonlinear Bratu Newton leg")
    ap.add_argument("--no-transfer", action="store_true", help="skip the hp state-transfer leg")
    ap.add_argument("--transfer-only", action="store_true", help="run only the hp state-transfer leg and print its JSON")
    ap.add_argument("--sipg-only", action="store_true",
                    help="run only the interior penalty DG leg (face kernel profiling) and print its JSON")
    ap.add_argument("--fcm", action="store_true",
                    help="run the finite-cell sphere leg (48^3 background cells, 64 right-hand sides, block CG "
                         "with hp-multigrid to 1e-8); off by default: fitting the cut-cell rules of every level "
                         "takes minutes of host time")
    ap.add_argument("--no-fcm", action="store_true", help="(default) skip the finite-cell sphere leg")
    ap.add_argument("--fcm-only", action="store_true", help="run only the finite-cell sphere leg and print its JSON")
    ap.add_argument("--fcm-cells", type=int, default=48, help="background cells per axis of the finite-cell leg")
    ap.add_argument("--fcm-rhs", type=int, default=64, help="right-hand sides of the finite-cell leg")
    ap.add_argument("--fcm-iterations", type=int, default=2000,
                    help="CG iteration limit of the finite-cell leg's solves (1e-8)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-correction FGMRES part of the log-normal leg")
    ap.add_argument("--lognormal-only", action="store_true",
                    help="run only the log-normal leg (with the learned correction) and print its JSON")
    ap.add_argument("--lognormal-depth", type=int, default=5, help="uniform refinements of the log-normal leg (5 -> 32^3)")
    ap.add_argument("--learned-pairs", type=int, default=2000, help="training residuals of the learned leg")
    ap.add_argument("--learned-steps", type=int, default=3000, help="Adam steps of the learned leg")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="after timing, copy a 1 GiB buffer three times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (cells would shard by Morton "
              "range with an interface-dof exchange per apply)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np
    import torch

    import gricoamd as g
    from gricoamd import _devabi, _native, ops

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.sipg_only:
        print(json.dumps({"sipg": sipg_leg()}), flush=True)
        return
    if args.transfer_only:
        print(json.dumps({"transfer": transfer_leg()}), flush=True)
        return
    if args.lognormal_only:
        print(json.dumps({"lognormal": lognormal_leg(depth=args.lognormal_depth, learned=not args.no_learned,
                                                     pairs=args.learned_pairs, steps=args.learned_steps)}),
              flush=True)
        return
    if args.fcm_only:
        print(json.dumps({"fcm": fcm_leg(cells=args.fcm_cells, nrhs=args.fcm_rhs, maxiter=args.fcm_iterations)}), flush=True)
        return
    t0 = time.perf_counter()
    problem = build_problem(args.depth)
    plan = g.compile(problem)
    log(f"host compile: {time.perf_counter() - t0:.1f}s, dofs {plan.full_dof_count}, "
        f"cells {sum(b.cell_count for b in plan.buckets)}")
    t0 = time.perf_counter()
    dev = ops._ws(plan)
    log(f"device upload: {time.perf_counter() - t0:.1f}s, {dev.device_bytes() / 2**20:.0f} MiB")
    log(dev.describe().strip())

    n = plan.full_dof_count
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20241203)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    dev.set_stream(stream.cuda_stream)
    lib = _native.device()
    for _ in range(args.warmup):
        _devabi.check(lib.gmf_apply_async(dev.ws, x.data_ptr(), y.data_ptr()))
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        _devabi.check(lib.gmf_apply_async(dev.ws, x.data_ptr(), y.data_ptr()))
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps

    cell_ms = ctypes.c_double()
    total_ms = ctypes.c_double()
    _devabi.check(lib.gmf_apply_timed(dev.ws, x.data_ptr(), y.data_ptr(), max(5, args.steps // 2),
                                      ctypes.byref(cell_ms), ctypes.byref(total_ms)))
    cells = sum(b.cell_count for b in plan.buckets)
    nq = plan.buckets[0].n_points
    k = 1 if plan.buckets[0].form == 0 else 16
    alg_bytes = 8 * n + 8 * n + 8 * cells * nq * k
    achieved = alg_bytes / (cell_ms.value * 1e-3) / 1e9
    log(f"apply {ms * 1e3:.1f} us/step; cell pass {cell_ms.value * 1e3:.1f} us, timed apply "
        f"{total_ms.value * 1e3:.1f} us; algorithmic bytes {alg_bytes / 1e6:.1f} MB")

    result = {
        "metric": "MF apply GDOF/s (%HBM peak)",
        "value": n / (ms * 1e-3) / 1e9,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) input vector)",
        "config": {
            "workload": f"3D Poisson [0,1]^3, {2 ** args.depth}^3 cells (uniform refinement of one root), "
                        f"p=4 FullTensorBasis, 5^3 Gauss points, kappa=1+0.5 sin sin sin, zero Dirichlet; "
                        f"full-layout matrix-free apply",
            "full_dofs": n,
            "cells": cells,
            "kernel_tuning": dev.tuning(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    traffic = _traffic_from_profiles(n, cells)
    if traffic is not None:
        result["roofline"]["traffic"] = traffic
    if args.calibration_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            _devabi.check(lib.gmf_vector_copy(dev.ws, dst.data_ptr(), src.data_ptr(), src.numel()))
        torch.cuda.synchronize()
        del src, dst
    try:
        result["rhs_l2"] = rhs_l2_leg(plan)
    except Exception as exc:  # keep the headline line
        log(f"rhs/l2 leg failed: {exc!r}")
        result["rhs_l2"] = {"error": repr(exc)}
    if not args.no_sipg:
        try:
            result["sipg"] = sipg_leg()
        except Exception as exc:  # keep the headline line
            log(f"sipg leg failed: {exc!r}")
            result["sipg"] = {"error": repr(exc)}
    if not args.no_newton:
        try:
            result["newton"] = newton_leg()
        except Exception as exc:  # keep the headline line
            log(f"newton leg failed: {exc!r}")
            result["newton"] = {"error": repr(exc)}
    if not args.no_transfer:
        try:
            result["transfer"] = transfer_leg()
        except Exception as exc:  # keep the headline line
            log(f"transfer leg failed: {exc!r}")
            result["transfer"] = {"error": repr(exc)}
    if not args.no_solve:
        result["solve"] = solve_leg(problem, plan)
    if not args.no_adaptive:
        try:
            result["adaptive"] = adaptive_leg()
        except Exception as exc:  # keep the headline line
            log(f"adaptive leg failed: {exc!r}")
            result["adaptive"] = {"error": repr(exc)}
    if not args.no_lognormal:
        try:
            result["lognormal"] = lognormal_leg(depth=args.lognormal_depth, learned=not args.no_learned,
                                                pairs=args.learned_pairs, steps=args.learned_steps)
        except Exception as exc:  # keep the headline line
            log(f"log-normal leg failed: {exc!r}")
            result["lognormal"] = {"error": repr(exc)}
    if args.fcm and not args.no_fcm:
        try:
            result["fcm"] = fcm_leg(cells=args.fcm_cells, nrhs=args.fcm_rhs, maxiter=args.fcm_iterations)
        except Exception as exc:  # keep the headline line
            log(f"fcm leg failed: {exc!r}")
            result["fcm"] = {"error": repr(exc)}
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
