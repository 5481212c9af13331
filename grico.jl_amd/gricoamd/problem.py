"""Problem description: fields, accumulator operators and constraints.

Reference: src/problem.jl (AffineProblem 372, SPD 234, operator classes),
src/fields.jl (ScalarField), src/accumulators.jl (TrialChannels 118-122,
TestChannels 148-151, cell_accumulate contract 523-539, the quadrature-point
handle 647-676). In this backend the callbacks are evaluated on the host only,
vectorized over every quadrature point of a cell batch, to lower each
accumulator to per-point coefficient tables; the device never calls user code.

Operators are plain Python objects that implement some of:

    cell_accumulate(self, q, trial, test_component) -> TestChannels or value
    cell_rhs_accumulate(self, q, test_component)    -> TestChannels or value
    boundary_accumulate(self, q, trial, test_component) / boundary_rhs_accumulate
    surface_accumulate(self, q, trial, test_component) / surface_rhs_accumulate
    interface_accumulate(self, q, trace_trial, test_component) -> TraceTestChannels
    interface_rhs_accumulate(self, q, test_component)           -> TraceTestChannels

(src/accumulators.jl:396-520 registration, 559-641 callback contracts).

`q.point` is a tuple of D coordinate arrays, `q.weight` the weights,
`q.cell_size` a tuple of D arrays; `trial.value` is an array and
`trial.gradient` a tuple of D arrays (physical gradient). Callbacks must be
written with numpy operations so they broadcast over points.
"""

import numpy as np

LOWER = 0
UPPER = 1


class SPD:
    """Declares the affine operator symmetric positive definite."""


class NonsymmetricOperator:
    pass


class ScalarField:
    def __init__(self, space, name="u"):
        self.space = space
        self.name = name
        self.components = 1


class TrialChannels:
    def __init__(self, component, value, gradient):
        self.component = component
        self.value = value
        self.gradient = gradient


class TestChannels:
    def __init__(self, value, gradient):
        self.value = value
        self.gradient = tuple(gradient)


class TracePair:
    """Oriented (minus, plus) pair (src/accumulators.jl:153-166): `jump` is
    plus - minus, `average` the arithmetic mean."""

    def __init__(self, minus, plus):
        self.minus = minus
        self.plus = plus


class TraceTrialChannels:
    """Two-sided trial data of an interface callback (src/accumulators.jl:168-180)."""

    def __init__(self, component, minus, plus):
        self.component = component
        self.minus = minus
        self.plus = plus

    @property
    def value(self):
        return TracePair(self.minus.value, self.plus.value)

    @property
    def gradient(self):
        return TracePair(tuple(self.minus.gradient), tuple(self.plus.gradient))


class TraceTestChannels:
    """Coefficients of the minus and plus test traces (src/accumulators.jl:196-209)."""

    def __init__(self, minus, plus):
        self.minus = minus
        self.plus = plus


def value(channels):
    return channels.value


def gradient(channels):
    return channels.gradient


def component(channels):
    return channels.component


def minus(pair):
    return pair.minus


def plus(pair):
    return pair.plus


def jump(a, b=None):
    """plus - minus of a TracePair, or jump(minus, plus) (src/integration.jl:375-380)."""
    if b is None:
        a, b = a.minus, a.plus
    if isinstance(a, tuple):
        return tuple(y - x for x, y in zip(a, b))
    return b - a


def average(a, b=None):
    if b is None:
        a, b = a.minus, a.plus
    if isinstance(a, tuple):
        return tuple(0.5 * (x + y) for x, y in zip(a, b))
    return 0.5 * (a + b)


def inner(a, b):
    if isinstance(a, TracePair):
        return TracePair(inner(a.minus, b.minus if isinstance(b, TracePair) else b),
                         inner(a.plus, b.plus if isinstance(b, TracePair) else b))
    if isinstance(a, tuple):
        out = a[0] * b[0]
        for x, y in zip(a[1:], b[1:]):
            out = out + x * y
        return out
    return a * b


def normal_component(vector, normal):
    """Contraction with the normal, side by side for a TracePair (src/accumulators.jl:295-302)."""
    if isinstance(vector, TracePair):
        return TracePair(normal_component(vector.minus, normal), normal_component(vector.plus, normal))
    return inner(tuple(vector), tuple(normal))


def normal(q):
    return q.normal


def weight(q):
    return q.weight


def cell_size(q, *_field):
    return q.cell_size


class QuadraturePoints:
    """Vectorized stand-in for the reference's per-point handle (surface
    points also carry the physical unit `normal`)."""

    def __init__(self, points, weights, cell_sizes, normal=None, face_axis=None, face_side=None):
        self.point = tuple(points)
        self.weight = weights
        self.cell_size = tuple(cell_sizes)
        self.normal = None if normal is None else tuple(normal)
        self.face_axis = face_axis
        self.face_side = face_side

    def coordinate(self, axis):
        return self.point[axis]


class BoundaryFace:
    def __init__(self, axis, side):
        self.axis = int(axis)
        self.side = int(side)


class Dirichlet:
    """Strong Dirichlet data g on one boundary face (src/plans.jl:486-514).

    `data` is a number or a callable of the point tuple (numpy arrays)."""

    def __init__(self, field, boundary, data=0.0):
        self.field = field
        self.boundary = boundary
        self.data = data


class MeanValue:
    """Integral-mean constraint of a field (src/problem.jl:190-195): the
    integral of u over the compiled cells equals `target` times their measure.
    Compiled as one explicit linear row on a free pivot dof
    (src/plans.jl:750-805)."""

    def __init__(self, field, target):
        t = np.asarray(target, dtype=np.float64)
        if t.size != 1 or not np.isfinite(t).all():
            raise ValueError("mean-value target of a scalar field must be one finite number")
        self.field = field
        self.target = float(t.reshape(-1)[0])


class PointQuadrature:
    """Explicit quadrature on the biunit reference cell [-1, 1]^D (src/quadrature.jl
    PointQuadrature): points (n x D) and weights (n)."""

    def __init__(self, points, weights):
        self.points = np.ascontiguousarray(np.atleast_2d(np.asarray(points, dtype=np.float64)))
        self.weights = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1))
        if self.points.shape[0] != self.weights.shape[0]:
            raise ValueError("point and weight data must match")
        if not (np.all(np.isfinite(self.points)) and np.all(np.isfinite(self.weights))):
            raise ValueError("quadrature data must be finite")

    def point_count(self):
        return self.weights.shape[0]


class SurfaceQuadrature:
    """Embedded-surface piece inside one leaf (src/embedded.jl:33-48): reference
    points and reference-surface weights plus one reference-space unit normal per
    point."""

    def __init__(self, leaf, quadrature, normals):
        if int(leaf) < 0:
            raise ValueError("leaf must be a cell id")
        self.leaf = int(leaf)
        self.quadrature = quadrature
        n = np.atleast_2d(np.asarray(normals, dtype=np.float64))
        if n.shape[0] == 1 and quadrature.point_count() > 1:
            n = np.repeat(n, quadrature.point_count(), axis=0)
        if n.shape != quadrature.points.shape:
            raise ValueError("one normal per surface quadrature point is required")
        norms = np.linalg.norm(n, axis=1)
        if np.any(norms <= 0) or not np.all(np.isfinite(norms)):
            raise ValueError("surface normals must be finite and nonzero")
        self.normals = np.ascontiguousarray(n / norms[:, None])


class AffineProblem:
    def __init__(self, field, operator_class=None):
        self.field = field
        self.operator_class = operator_class or SPD()
        self.cell_operators = []   # (test_field, trial_field, operator)
        self.cell_rhs = []         # (test_field, operator)
        self.constraints = []
        self.mean_constraints = []
        self.cell_quadratures = {}     # leaf -> PointQuadrature
        self.surface_quadratures = []  # SurfaceQuadrature
        self.embedded_surfaces = []    # geometry converted to SurfaceQuadrature pieces at compile time
        self.surface_operators = []    # (test_field, trial_field, operator)
        self.surface_rhs = []          # (test_field, operator)
        self.boundary_operators = []   # (BoundaryFace, test_field, trial_field, operator)
        self.boundary_rhs = []         # (BoundaryFace, test_field, operator)
        self.interface_operators = []  # (test_field, trial_field, operator)
        self.interface_rhs = []        # (test_field, operator)

    def _fields(self, args, kind):
        if len(args) == 3:
            test, trial, op = args
            if test is not self.field or trial is not self.field:
                raise ValueError("this backend supports a single scalar field")
            return (test, trial, op), True
        if len(args) == 2:
            test, op = args
            if test is not self.field:
                raise ValueError("this backend supports a single scalar field")
            return (test, op), False
        raise TypeError(f"add_{kind}_accumulator(..., test, trial, op) or (..., test, op)")

    def add_boundary_accumulator(self, boundary, *args):
        """boundary_accumulate / boundary_rhs_accumulate operators on one
        boundary side (src/accumulators.jl:396-421)."""
        if not isinstance(boundary, BoundaryFace):
            raise TypeError("add_boundary_accumulator expects a BoundaryFace")
        if not 0 <= boundary.axis < self.field.space.dim or boundary.side not in (LOWER, UPPER):
            raise ValueError("boundary face axis or side out of range")
        entry, bilinear = self._fields(args, "boundary")
        (self.boundary_operators if bilinear else self.boundary_rhs).append((boundary,) + entry)
        return self

    def add_interface_accumulator(self, *args):
        """interface_accumulate / interface_rhs_accumulate operators on every
        interior interface (src/accumulators.jl:488-520)."""
        entry, bilinear = self._fields(args, "interface")
        (self.interface_operators if bilinear else self.interface_rhs).append(entry)
        return self

    def add_cell_quadrature(self, leaf, quadrature):
        """Replace the tensor rule of one leaf (src/embedded.jl:203-208)."""
        if quadrature.points.shape[1] != self.field.space.dim:
            raise ValueError("quadrature dimension must match the space")
        self.cell_quadratures[int(leaf)] = quadrature
        return self

    def add_surface_quadrature(self, surface):
        """Attach one embedded-surface piece (src/embedded.jl:181-186)."""
        if surface.quadrature.points.shape[1] != self.field.space.dim:
            raise ValueError("surface quadrature dimension must match the space")
        self.surface_quadratures.append(surface)
        return self

    def add_embedded_surface(self, surface):
        """Attach surface geometry that compile() turns into SurfaceQuadrature
        pieces on the active leaves of the problem's space
        (src/embedded.jl:140-163); it must provide surface_quadratures(space)."""
        if not hasattr(surface, "surface_quadratures"):
            raise TypeError("embedded surfaces must provide surface_quadratures(space)")
        self.embedded_surfaces.append(surface)
        return self

    def add_surface_accumulator(self, *args):
        """surface_accumulate(q, trial, test_component) / surface_rhs_accumulate
        operators (src/accumulators.jl:435-466)."""
        if len(args) == 3:
            test, trial, op = args
            if test is not self.field or trial is not self.field:
                raise ValueError("this backend supports a single scalar field")
            self.surface_operators.append((test, trial, op))
        elif len(args) == 2:
            test, op = args
            if test is not self.field:
                raise ValueError("this backend supports a single scalar field")
            self.surface_rhs.append((test, op))
        else:
            raise TypeError("add_surface_accumulator(test, trial, op) or (test, op)")
        return self

    def add_cell_accumulator(self, *args):
        if len(args) == 3:
            test, trial, op = args
            if test is not self.field or trial is not self.field:
                raise ValueError("this backend supports a single scalar field")
            self.cell_operators.append((test, trial, op))
        elif len(args) == 2:
            test, op = args
            if test is not self.field:
                raise ValueError("this backend supports a single scalar field")
            self.cell_rhs.append((test, op))
        else:
            raise TypeError("add_cell_accumulator(test, trial, op) or (test, op)")
        return self

    def add_constraint(self, constraint):
        if isinstance(constraint, MeanValue):
            if constraint.field is not self.field:
                raise ValueError("this backend supports a single scalar field")
            self.mean_constraints.append(constraint)
            return self
        if not isinstance(constraint, Dirichlet):
            raise NotImplementedError("only Dirichlet and MeanValue constraints are supported")
        self.constraints.append(constraint)
        return self


def add_cell_accumulator(problem, *args):
    return problem.add_cell_accumulator(*args)


def add_constraint(problem, constraint):
    return problem.add_constraint(constraint)


def add_cell_quadrature(problem, leaf, quadrature):
    return problem.add_cell_quadrature(leaf, quadrature)


def add_surface_quadrature(problem, surface):
    return problem.add_surface_quadrature(surface)


def add_embedded_surface(problem, surface):
    return problem.add_embedded_surface(surface)


def add_surface_accumulator(problem, *args):
    return problem.add_surface_accumulator(*args)


def add_boundary_accumulator(problem, boundary, *args):
    return problem.add_boundary_accumulator(boundary, *args)


def add_interface_accumulator(problem, *args):
    return problem.add_interface_accumulator(*args)
