"""Linear solver policies and `solve` on the device.

Reference: src/solve.jl (CGSolver / FGMRESSolver / JacobiPreconditioner
policies 10-113; solve 219-243; _solve_affine 249-266 with defaults
relative_tolerance = sqrt(eps), absolute_tolerance = 0,
maxiter = max(1000, 2 n_red); _cg_solve 376-431; _fgmres_solve 433-542;
Jacobi 651-774). The Krylov loops run in libgrico_mf.so; only the stopping
test and the Givens rotations touch the host.
"""

import ctypes
import math

import numpy as np

from . import _devabi, _native, ops
from .plan import AssemblyPlan, compile as compile_plan
from .problem import AffineProblem


class JacobiPreconditioner:
    """Jacobi policy (src/solve.jl:651-656). precision "fp32" applies the
    inverse diagonal in single precision inside the FP64 Krylov method."""

    def __init__(self, precision="fp64"):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision


class CGSolver:
    def __init__(self, preconditioner=None):
        self.preconditioner = preconditioner


class FGMRESSolver:
    def __init__(self, preconditioner=None, restart=30):
        if int(restart) < 1:
            raise ValueError("restart must be positive")
        self.preconditioner = preconditioner
        self.restart = int(restart)


class AutoLinearSolver:
    pass


class State:
    """Full-layout coefficient vector of a solved plan (src/fields.jl:528-560)."""

    def __init__(self, plan, coefficients, info=None):
        self.plan = plan
        self.coefficients = np.asarray(coefficients, dtype=np.float64)
        self.info = info


class ConvergenceError(RuntimeError):
    pass


class DevicePreconditioner:
    """Owning wrapper of a gmf_preconditioner handle."""

    def __init__(self, plan, handle, keep=()):
        self.plan = plan
        self.handle = handle
        self._keep = keep

    def __del__(self):
        if _native is None:  # interpreter shutdown
            return
        lib = _native._device
        if lib is not None and getattr(self, "handle", None) is not None and self.handle.value:
            lib.gmf_precond_destroy(self.handle)
            self.handle = None

    def apply(self, r):
        torch = ops._torch()
        rt = ops.to_device(r, self.plan.reduced_dof_count)
        z = torch.empty_like(rt)
        _devabi.check(_native.device().gmf_precond_apply(self.handle, rt.data_ptr(), z.data_ptr()))
        return z

    def apply_block(self, R):
        """Z = M^-1 R for the columns of R (n x k): one block V-cycle for FP64
        multigrid with Jacobi or Chebyshev smoothing, else column by column
        (gmf_precond_apply_block)."""
        torch = ops._torch()
        n = self.plan.reduced_dof_count
        Rt = torch.as_tensor(R, dtype=torch.float64).to("cuda")
        if Rt.ndim != 2 or Rt.shape[0] != n:
            raise ValueError(f"expected a block of shape ({n}, k)")
        Rc = Rt.t().contiguous()  # column-major storage: one column after another
        Z = torch.empty_like(Rc)
        _devabi.check(_native.device().gmf_precond_apply_block(self.handle, Rc.data_ptr(), Z.data_ptr(),
                                                               int(Rt.shape[1])))
        return Z.t()


def reduced_diagonal_general(plan):
    """Diagonal of P^T A P when P is not a selection (src/assembly.jl:662-675
    gating; src/solve.jl:722-738 probing). P = S + sum_p e_p w_p^T with S the
    selection of the free dofs and w_p the expansion of relation / mean pivot p,
    so for a symmetric A
    diag_j = A[s_j, s_j] + 2 sum_p w_p[j] A[p, s_j] + sum_{p,q} w_p[j] A[p, q] w_q[j],
    which needs the physical diagonal and one physical apply per pivot (the
    local physical diagonal is exact on single-term tensor-cell maps). Other
    plans are probed with one reduced apply per reduced dof, as the reference
    does."""
    from .problem import SPD
    torch = ops._torch()
    n = plan.reduced_dof_count
    local_exact = all(len(b.multi_offsets) <= 1 for b in plan.buckets) and not (
        plan.cloud_buckets or plan.surface_buckets or plan.boundary_buckets or plan.interface_buckets)
    if not (isinstance(plan.problem.operator_class, SPD) and local_exact):
        d = np.zeros(n)
        e = torch.zeros(n, dtype=torch.float64, device="cuda")
        for j in range(n):
            e.zero_()
            e[j] = 1.0
            d[j] = float(ops.reduced_apply(plan, e)[j])
        return d
    rows_of = np.diff(plan.map_offsets)
    sel = plan.solve_dofs
    pivots = np.nonzero((rows_of > 0) & (plan.reduced_index < 0))[0]
    d = ops.diagonal_physical(plan)[sel].copy()
    if len(pivots) == 0:
        return d
    W = np.zeros((len(pivots), n))
    cols = np.zeros((len(pivots), plan.full_dof_count))
    for i, p in enumerate(pivots.tolist()):
        lo, hi = plan.map_offsets[p], plan.map_offsets[p + 1]
        W[i, plan.map_indices[lo:hi]] = plan.map_coefficients[lo:hi]
        e = np.zeros(plan.full_dof_count)
        e[p] = 1.0
        cols[i] = ops.apply_physical(plan, e)
    App = cols[:, pivots]  # A[pivot_i, pivot_k] = (A e_k)[pivot_i]
    d += 2.0 * np.einsum("pj,pj->j", W, cols[:, sel])
    d += np.einsum("pj,pq,qj->j", W, App.T, W)
    return d


def jacobi_inverse_diagonal(plan):
    """Inverse reduced diagonal with the reference's zero / non-finite check."""
    torch = ops._torch()
    if getattr(plan, "selection_map", True):
        d = ops.reduced_diagonal(plan, as_numpy=False)
    else:
        d = ops.to_device(reduced_diagonal_general(plan))
    inv = torch.empty_like(d)
    dev = ops._ws(plan)
    _devabi.check(_native.device().gmf_invert_diagonal(dev.ws, d.data_ptr(), inv.data_ptr(), d.numel()))
    return inv


def compile_preconditioner(policy, plan):
    if policy is None:
        return None
    if isinstance(policy, DevicePreconditioner):
        return policy
    if isinstance(policy, JacobiPreconditioner):
        inv = jacobi_inverse_diagonal(plan)
        handle = _native.c_void_p()
        dev = ops._ws(plan)
        create = _native.device().gmf_precond_create_jacobi_f32 if getattr(policy, "precision", "fp64") == "fp32" \
            else _native.device().gmf_precond_create_jacobi
        _devabi.check(create(dev.ws, inv.data_ptr(), ctypes.byref(handle)))
        return DevicePreconditioner(plan, handle, keep=(inv,))
    compile_hook = getattr(policy, "compile", None)
    if compile_hook is not None:
        return compile_hook(plan)
    raise TypeError(f"unsupported preconditioner policy {type(policy).__name__}")


def _default_solver(plan, solver):
    if solver is None or isinstance(solver, AutoLinearSolver):
        return CGSolver(preconditioner=JacobiPreconditioner())
    return solver


def solve_reduced(plan, solver, b_red, relative_tolerance=None, absolute_tolerance=0.0, maxiter=None,
                  initial_solution=None):
    """Solve P^T A P x = b on the device; returns (x torch tensor, info)."""
    torch = ops._torch()
    n = plan.reduced_dof_count
    rtol = math.sqrt(np.finfo(np.float64).eps) if relative_tolerance is None else float(relative_tolerance)
    atol = float(absolute_tolerance)
    if not (math.isfinite(rtol) and rtol >= 0 and math.isfinite(atol) and atol >= 0):
        raise ValueError("tolerances must be finite and non-negative")
    maxiter = max(1000, 2 * n) if maxiter is None else int(maxiter)
    if maxiter < 1:
        raise ValueError("maxiter must be positive")
    b = ops.to_device(b_red, n)
    x = torch.zeros(n, dtype=torch.float64, device="cuda") if initial_solution is None \
        else ops.to_device(initial_solution, n).clone()
    if n == 0:  # every dof constrained: the solution is the map's shift
        info = _devabi.SolveInfo()
        info.converged = 1
        return x, info
    M = compile_preconditioner(solver.preconditioner, plan)
    dev = ops._ws(plan)
    info = _devabi.SolveInfo()
    handle = M.handle if M is not None else None
    lib = _native.device()
    if isinstance(solver, CGSolver):
        status = lib.gmf_cg_solve(dev.ws, handle, b.data_ptr(), x.data_ptr(), int(initial_solution is not None),
                                  rtol, atol, maxiter, ctypes.byref(info))
    elif isinstance(solver, FGMRESSolver):
        status = lib.gmf_fgmres_solve(dev.ws, handle, b.data_ptr(), x.data_ptr(),
                                      int(initial_solution is not None), solver.restart, rtol, atol, maxiter,
                                      ctypes.byref(info))
    else:
        raise TypeError(f"unsupported solver {type(solver).__name__}")
    if status == -3:
        raise ConvergenceError(_native.device().gmf_last_error().decode())
    _devabi.check(status)
    return x, info


def reduced_apply_block(plan, X):
    """Y[:, j] = P^T A P X[:, j] for a (n_red, k) block (gmf_reduced_apply_block);
    point-cloud batches take every column in one launch."""
    torch = ops._torch()
    Xc = _block(plan, X)
    Y = torch.empty_like(Xc)
    _devabi.check(_native.device().gmf_reduced_apply_block(ops._ws(plan).ws, Xc.data_ptr(), Y.data_ptr(),
                                                           Xc.shape[0]))
    return Y.T


def _block(plan, B):
    """Device copy of an (n_red, k) block, stored column after column (k x n_red)."""
    torch = ops._torch()
    n = plan.reduced_dof_count
    Bt = torch.as_tensor(B, dtype=torch.float64)
    if Bt.dim() != 2 or Bt.shape[0] != n:
        raise ValueError(f"block must have shape ({n}, k)")
    if not 1 <= Bt.shape[1] <= 4096:
        raise ValueError("column count must be in 1..4096")
    return Bt.T.to("cuda").contiguous()


def solve_block(plan, B, solver=None, relative_tolerance=None, absolute_tolerance=0.0, maxiter=None,
                initial_solution=None):
    """CG on every column of an (n_red, k) block of reduced right-hand sides in
    lockstep (gmf_cg_solve_block): the k solves of solve_reduced with a shared
    operator, each column with its own scalars and stopping test. Returns
    (X as an (n_red, k) device view, list of per-column info)."""
    torch = ops._torch()
    solver = solver or CGSolver(preconditioner=JacobiPreconditioner())
    if not isinstance(solver, CGSolver):
        raise TypeError("block solves support CGSolver")
    n = plan.reduced_dof_count
    rtol = math.sqrt(np.finfo(np.float64).eps) if relative_tolerance is None else float(relative_tolerance)
    atol = float(absolute_tolerance)
    if not (math.isfinite(rtol) and rtol >= 0 and math.isfinite(atol) and atol >= 0):
        raise ValueError("tolerances must be finite and non-negative")
    maxiter = max(1000, 2 * n) if maxiter is None else int(maxiter)
    if maxiter < 1:
        raise ValueError("maxiter must be positive")
    Bd = _block(plan, B)
    k = Bd.shape[0]
    X = torch.zeros_like(Bd) if initial_solution is None else _block(plan, initial_solution).clone()
    M = compile_preconditioner(solver.preconditioner, plan)
    infos = (_devabi.SolveInfo * k)()
    status = _native.device().gmf_cg_solve_block(ops._ws(plan).ws, M.handle if M is not None else None,
                                                 Bd.data_ptr(), X.data_ptr(), k, int(initial_solution is not None),
                                                 rtol, atol, maxiter, infos)
    if status == -3:
        raise ConvergenceError(_native.device().gmf_last_error().decode())
    _devabi.check(status)
    return X.T, list(infos)


def _default_affine_preconditioner(problem):
    """GMG for SPD problems with a coarsening level, else Jacobi (src/multigrid.jl:259-266)."""
    from .multigrid import GeometricMultigridPreconditioner, has_coarsening_level
    from .problem import SPD
    candidate = GeometricMultigridPreconditioner()
    if isinstance(problem.operator_class, SPD) and has_coarsening_level(problem.field.space, candidate):
        return candidate
    return JacobiPreconditioner()


def solve(problem_or_plan, solver=None, relative_tolerance=None, absolute_tolerance=0.0, maxiter=None,
          initial_solution=None):
    """solve(problem; solver=...) / solve(plan; solver=...) -> State.

    On a problem the default (AutoLinearSolver) is CG with the hp geometric
    multigrid preconditioner when the problem is SPD and has a coarsening level,
    else CG with Jacobi (src/multigrid.jl:199-266); on a compiled plan it is CG
    with Jacobi (src/solve.jl:219-243)."""
    from .multigrid import GeometricMultigridPreconditioner, check_cg_policy, compile_geometric_multigrid
    is_plan = isinstance(problem_or_plan, AssemblyPlan)
    if not is_plan and (solver is None or isinstance(solver, AutoLinearSolver)):
        solver = CGSolver(preconditioner=_default_affine_preconditioner(problem_or_plan))
    if isinstance(getattr(solver, "preconditioner", None), GeometricMultigridPreconditioner):
        if is_plan:
            raise ValueError("GeometricMultigridPreconditioner requires the original AffineProblem so every level "
                             "can rediscretize the operators; call solve(problem, solver=CGSolver("
                             "preconditioner=GeometricMultigridPreconditioner()))")
        if isinstance(solver, CGSolver):
            check_cg_policy(problem_or_plan, solver.preconditioner)
        plan = compile_plan(problem_or_plan)
        hierarchy = compile_geometric_multigrid(problem_or_plan, solver.preconditioner, fine_plan=plan)
        M = hierarchy.device_preconditioner()
        solver = CGSolver(preconditioner=M) if isinstance(solver, CGSolver) else \
            FGMRESSolver(preconditioner=M, restart=solver.restart)
    else:
        plan = problem_or_plan if is_plan else compile_plan(problem_or_plan)
        solver = _default_solver(plan, solver)
    # Device-resident right-hand side (point loads lowered on the host) for
    # every batch the device plan holds; loads on batches it does not hold
    # (plan.device_rhs is False) take the host path.
    b = None
    if plan.device_rhs:
        b = ops.reduced_rhs(plan, as_numpy=False)
    if b is None:
        b = plan.reduced_rhs()
    x, info = solve_reduced(plan, solver, b, relative_tolerance, absolute_tolerance, maxiter, initial_solution)
    full = ops.expand_reduced(plan, x, include_shift=True).cpu().numpy()
    return State(plan, full, info)
