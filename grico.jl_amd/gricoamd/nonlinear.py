"""Nonlinear residual problems and Newton's method on the device.

Reference: ResidualProblem (src/problem.jl), cell residual / tangent
accumulators (src/accumulators.jl:609-621 contracts, 2015-2148 tensor
kernels), residual! / tangent_apply! / _reduced_residual! /
_reduced_tangent_apply! (src/assembly.jl:935-1080, constraint residual rows
1922-1940), Newton (src/solve.jl:278-339 _solve_residual) and its default
linear solve (686-711: CG for an SPD tangent, FGMRES otherwise, no
preconditioner).

Callbacks are evaluated on the host, vectorized over a bucket's points:

    cell_residual_accumulate(self, q, state, test_component)           -> TestChannels or value
    cell_tangent_accumulate(self, q, state, increment, test_component) -> TestChannels or value

with `state.value` an array and `state.gradient` a tuple of arrays (physical
gradient). Per Newton step the device interpolates the state's value and
gradient at every point (gmf_interpolate_channels), the host evaluates the
residual callback into point loads W (v, g), the device projects them
(gmf_rhs_physical) and restricts them to the free dofs; the tangent is lowered
at the state by probing the tangent callback with unit increments (it is linear
in the increment) and written into the plan's point tables
(gmf_plan_set_point_data), and the correction is solved on the device.
"""

import ctypes
import math

import numpy as np

from . import _devabi, _native, ops
from .plan import FORM_GENERAL, AssemblyPlan, _channels
from .problem import SPD, AffineProblem, Dirichlet, TrialChannels
from .solve import CGSolver, ConvergenceError, FGMRESSolver, State, solve_reduced


class StateChannels:
    """Value and physical gradient of the state at the points (src/accumulators.jl:243-270)."""

    def __init__(self, value, gradient):
        self.value = value
        self.gradient = tuple(gradient)


class ResidualProblem:
    """Nonlinear problem R(u; v) = 0 with Dirichlet constraints (src/problem.jl)."""

    def __init__(self, field, operator_class=None):
        self.field = field
        self.operator_class = operator_class or SPD()
        self.cell_operators = []  # (test_field, state_field, operator)
        self.constraints = []

    def add_cell_accumulator(self, test, state, op):
        if test is not self.field or state is not self.field:
            raise ValueError("this backend supports a single scalar field")
        for name in ("cell_residual_accumulate", "cell_tangent_accumulate"):
            if not callable(getattr(op, name, None)):
                raise TypeError(f"residual operators implement {name}")
        self.cell_operators.append((test, state, op))
        return self

    def add_constraint(self, constraint):
        if not isinstance(constraint, Dirichlet):
            raise NotImplementedError("only Dirichlet constraints are supported")
        self.constraints.append(constraint)
        return self


class _TangentAt:
    """Affine view of one residual operator: the tangent at a fixed state
    (arrays over a bucket's cells x points), for lowering by probing; the
    bucket's probing chunks select their cells with set_window."""

    def __init__(self, op, state):
        self.op = op
        self.state = state
        self.window = (0, None)

    def set_window(self, lo, hi):
        self.window = (lo, hi)

    def cell_accumulate(self, q, trial, test_component):
        lo, hi = self.window
        st = StateChannels(self.state.value[lo:hi], [g[lo:hi] for g in self.state.gradient])
        return self.op.cell_tangent_accumulate(q, st, trial, test_component)


def _as_general(bucket):
    """Store a DIFFUSION-form bucket as the GENERAL form (the tangent's form
    must not change between Newton steps)."""
    if bucket.form == FORM_GENERAL:
        return
    D = bucket.dim
    c = bucket.point_data.reshape(bucket.cell_count, bucket.n_points)
    M = np.zeros((bucket.cell_count, bucket.n_points, D + 1, D + 1))
    for a in range(D):
        M[..., a + 1, a + 1] = c
    bucket.form = FORM_GENERAL
    bucket.point_data = np.ascontiguousarray(M.reshape(bucket.cell_count, -1))


class ResidualPlan:
    """Compiled residual problem: the host tables of an AssemblyPlan whose cell
    action is the tangent, re-lowered at each state."""

    def __init__(self, problem):
        self.problem = problem
        self.operator_class = problem.operator_class
        shim = AffineProblem(problem.field, problem.operator_class)
        zero_ops = []
        for (_, _, op) in problem.cell_operators:
            zero_ops.append(_ZeroTangent(op, problem.field.space.dim))
            shim.add_cell_accumulator(problem.field, problem.field, zero_ops[-1])
        for con in problem.constraints:
            shim.add_constraint(con)
        self.plan = AssemblyPlan(shim)
        if not self.plan.tensor_only:
            raise NotImplementedError("residual problems cover tensor-product cells")
        for b in self.plan.buckets:
            _as_general(b)
        self.full_dof_count = self.plan.full_dof_count
        self.reduced_dof_count = self.plan.reduced_dof_count

    # --- state channels and lowering -------------------------------------------------
    def state_channels(self, u_full):
        """Per bucket StateChannels (arrays cells x Q) of a full-layout state (device)."""
        torch = ops._torch()
        D = self.plan.dim
        C = D + 1
        total = sum(b.cell_count * b.n_points for b in self.plan.buckets)
        out = torch.empty(total * C, dtype=torch.float64, device="cuda")
        u = ops.to_device(u_full, self.full_dof_count)
        dev = ops._ws(self.plan)
        _devabi.check(_native.device().gmf_interpolate_channels(dev.ws, u.data_ptr(), C, out.data_ptr()))
        host = out.cpu().numpy()
        res, off = [], 0
        for b in self.plan.buckets:
            n = b.cell_count * b.n_points * C
            arr = host[off:off + n].reshape(b.cell_count, b.n_points, C)
            off += n
            res.append(StateChannels(arr[..., 0], [arr[..., 1 + a] for a in range(D)]))
        return res

    def set_state(self, u_full, states=None):
        """Lower the tangent at the state into the device plan (src/assembly.jl:1050-1080)."""
        states = states or self.state_channels(u_full)
        dev = ops._ws(self.plan)
        ops_ = [op for (_, _, op) in self.problem.cell_operators]
        for i, (b, st) in enumerate(zip(self.plan.buckets, states)):
            form, data = b.lower_operators([_TangentAt(op, st) for op in ops_])
            if form != FORM_GENERAL:
                tmp = _Tmp(b, form, data)
                _as_general(tmp)
                data = tmp.point_data
            b.point_data = np.ascontiguousarray(data)
            _devabi.check(_native.device().gmf_plan_set_point_data(
                dev.handle, i, b.point_data.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return states

    # --- residual ------------------------------------------------------------------
    def residual_loads(self, states):
        """Point loads W (v, g) of the residual callbacks at the states."""
        D = self.plan.dim
        per = []
        for b, st in zip(self.plan.buckets, states):
            shape = (b.cell_count, b.n_points)
            v = np.zeros(shape)
            g = [np.zeros(shape) for _ in range(D)]
            q = b.quadrature()
            for (_, _, op) in self.problem.cell_operators:
                vv, gg = _channels(op.cell_residual_accumulate(q, st, 0), shape, D)
                v = v + vv
                for a in range(D):
                    g[a] = g[a] + gg[a]
            per.append(np.stack([b.weights * v] + [b.weights * ga for ga in g], axis=-1).reshape(-1))
        return np.concatenate(per) if per else np.zeros(0)

    def residual_physical(self, u_full, states=None):
        """sum over cells of the projected residual loads (full layout, device tensor)."""
        torch = ops._torch()
        states = states or self.state_channels(u_full)
        loads = ops.to_device(self.residual_loads(states))
        y = torch.empty(self.full_dof_count, dtype=torch.float64, device="cuda")
        dev = ops._ws(self.plan)
        _devabi.check(_native.device().gmf_rhs_physical(dev.ws, loads.data_ptr(), self.plan.dim + 1, y.data_ptr()))
        return y

    def reduced_residual(self, u_full, states=None):
        return ops.project_reduced(self.plan, self.residual_physical(u_full, states))

    def expand(self, x_red):
        return ops.expand_reduced(self.plan, x_red, include_shift=True)


class _Tmp:
    def __init__(self, b, form, data):
        self.form, self.point_data = form, data
        self.dim, self.cell_count, self.n_points = b.dim, b.cell_count, b.n_points


class _ZeroTangent:
    """Tangent at the zero state, used once to build the plan's tables."""

    def __init__(self, op, dim):
        self.op = op
        self.dim = dim

    def cell_accumulate(self, q, trial, test_component):
        z = np.zeros_like(np.asarray(trial.value, dtype=np.float64))
        return self.op.cell_tangent_accumulate(q, StateChannels(z, [z] * self.dim), trial, test_component)


def compile_residual(problem):
    return ResidualPlan(problem)


def residual(rplan, state):
    """Full-layout residual with constraint rows u[d] - g_d (src/assembly.jl:997-1026, 1922-1940)."""
    u = np.asarray(state.coefficients if isinstance(state, State) else state, dtype=np.float64)
    r = rplan.residual_physical(u).cpu().numpy()
    p = rplan.plan
    r[p.fixed_dofs] = u[p.fixed_dofs] - p.fixed_values
    return r


def tangent_apply(rplan, state, increment):
    """Tangent at `state` applied to a full-layout increment (fixed rows copy the increment)."""
    u = np.asarray(state.coefficients if isinstance(state, State) else state, dtype=np.float64)
    rplan.set_state(u)
    return ops.apply(rplan.plan, np.asarray(increment, dtype=np.float64))


def newton_solve(rplan, initial_state=None, relative_tolerance=None, absolute_tolerance=None, maxiter=20,
                 damping=1.0, linear_relative_tolerance=None, linear_absolute_tolerance=0.0, linear_maxiter=None,
                 linear_solve=None, preconditioner=None):
    """Newton's method (src/solve.jl:278-339): stop when the reduced residual
    norm is at most max(atol, rtol * initial norm); the correction solves the
    tangent system with CG (SPD problem) or FGMRES and no preconditioner
    (src/solve.jl:686-711), or with `linear_solve(rplan, state, rhs, ...)`.
    Returns a State; `state.info` holds the residual norm history."""
    eps = math.sqrt(np.finfo(np.float64).eps)
    rtol = eps if relative_tolerance is None else float(relative_tolerance)
    atol = eps if absolute_tolerance is None else float(absolute_tolerance)
    lrtol = eps if linear_relative_tolerance is None else float(linear_relative_tolerance)
    if not (damping > 0 and math.isfinite(damping)):
        raise ValueError("damping must be positive and finite")
    if maxiter < 1:
        raise ValueError("maxiter must be positive")
    if preconditioner is not None and linear_solve is None:
        raise ValueError("matrix-free tangent preconditioning is not implemented; pass preconditioner=None "
                         "or a custom linear_solve")
    torch = ops._torch()
    p = rplan.plan
    n = p.reduced_dof_count
    if initial_state is None:
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
    else:
        u0 = np.asarray(initial_state.coefficients if isinstance(initial_state, State) else initial_state)
        x = ops.to_device(np.ascontiguousarray(u0[p.solve_dofs], dtype=np.float64))
    solver = CGSolver(preconditioner=None) if isinstance(rplan.operator_class, SPD) else \
        FGMRESSolver(preconditioner=None)
    history = []
    r0 = None
    for _ in range(int(maxiter)):
        u = rplan.expand(x)
        states = rplan.state_channels(u)
        r = rplan.reduced_residual(u, states)
        norm = float(torch.linalg.vector_norm(r))
        history.append(norm)
        r0 = norm if r0 is None else r0
        if norm <= max(atol, rtol * r0):
            return _state(rplan, x, history)
        if linear_solve is not None:
            corr = ops.to_device(linear_solve(rplan, u, -r, relative_tolerance=lrtol,
                                              absolute_tolerance=linear_absolute_tolerance,
                                              maxiter=linear_maxiter), n)
        else:
            rplan.set_state(u, states)
            corr, _ = solve_reduced(p, solver, -r, lrtol, linear_absolute_tolerance, linear_maxiter)
        x = x + damping * corr
    u = rplan.expand(x)
    r = rplan.reduced_residual(u)
    norm = float(torch.linalg.vector_norm(r))
    history.append(norm)
    if norm <= max(atol, rtol * (r0 if r0 is not None else 0.0)):
        return _state(rplan, x, history)
    raise ConvergenceError(f"Newton solve did not converge in {int(maxiter)} iterations")


def _state(rplan, x, history):
    full = rplan.expand(x).cpu().numpy()
    return State(rplan.plan, full, info={"residual_norms": history, "iterations": len(history) - 1})
