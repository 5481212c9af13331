"""Compiled hp spaces (`HpSpace`, `SpaceOptions` and the degree/quadrature policies).

Reference: src/space.jl (SpaceOptions 234-245, HpSpace 281-326,
DegreePlusQuadrature 165-171 / 714-725, UniformDegree / AxisDegrees /
ByLeafDegrees / StoredDegrees policies) and src/basis.jl (FullTensorBasis,
TrunkBasis). The continuity algebra runs in the native host library.
"""

import ctypes

import numpy as np

from . import _native
from .mesh import Domain


class FullTensorBasis:
    kind = 0


class TrunkBasis:
    kind = 1


class UniformDegree:
    def __init__(self, degree):
        self.degree = int(degree)

    def leaf_degrees(self, domain, leaves):
        return np.full((len(leaves), domain.dim), self.degree, dtype=np.int32)


class AxisDegrees:
    def __init__(self, degrees):
        self.degrees = tuple(int(d) for d in degrees)

    def leaf_degrees(self, domain, leaves):
        if len(self.degrees) != domain.dim:
            raise ValueError("AxisDegrees length must match the dimension")
        return np.tile(np.asarray(self.degrees, dtype=np.int32), (len(leaves), 1))


class ByLeafDegrees:
    """Per-leaf degrees from a callable `f(domain, leaf) -> int or tuple`."""

    def __init__(self, fn):
        self.fn = fn

    def leaf_degrees(self, domain, leaves):
        out = np.zeros((len(leaves), domain.dim), dtype=np.int32)
        for i, leaf in enumerate(leaves):
            d = self.fn(domain, int(leaf))
            out[i] = d if np.ndim(d) else [d] * domain.dim
        return out


class StoredDegrees:
    """Explicit degree tuples for a given list of leaves (cell ids)."""

    def __init__(self, leaves, degrees):
        self.map = {int(l): tuple(int(v) for v in d) for l, d in zip(leaves, np.asarray(degrees))}

    def leaf_degrees(self, domain, leaves):
        return np.asarray([self.map[int(l)] for l in leaves], dtype=np.int32).reshape(len(leaves), domain.dim)


class DegreePlusQuadrature:
    def __init__(self, extra_points=1):
        if int(extra_points) < 0:
            raise ValueError("extra_points must be non-negative")
        self.extra_points = int(extra_points)

    def shape(self, degrees):
        return tuple(max(int(p) + self.extra_points, 1) for p in degrees)


class SpaceOptions:
    def __init__(self, basis=None, degree=None, quadrature=None, continuity="cg"):
        self.basis = basis if basis is not None else TrunkBasis()
        self.degree = degree if degree is not None else UniformDegree(1)
        self.quadrature = quadrature if quadrature is not None else DegreePlusQuadrature(1)
        self.continuity = continuity


def _cg_mask(continuity, dim):
    if isinstance(continuity, str):
        kinds = (continuity,) * dim
    else:
        kinds = tuple(continuity)
        if len(kinds) != dim:
            raise ValueError("continuity tuple length must match the space dimension")
    mask = 0
    for a, k in enumerate(kinds):
        if k not in ("cg", "dg"):
            raise ValueError("continuity policy axes must be either 'cg' or 'dg'")
        if k == "cg":
            mask |= 1 << a
    return mask, kinds


class HpSpace:
    """A compiled hp space on the active leaves of `domain` (or an explicit leaf set)."""

    def __init__(self, domain: Domain, options: SpaceOptions = None, leaves=None, degrees=None):
        options = options or SpaceOptions()
        self.domain = domain
        self.options = options
        self.dim = domain.dim
        self.basis = options.basis
        self.quadrature_policy = options.quadrature
        self.cg_mask, self.continuity = _cg_mask(options.continuity, domain.dim)
        given = np.asarray(domain.active_leaves() if leaves is None else leaves, dtype=np.int64)
        if degrees is None:
            degrees = options.degree.leaf_degrees(domain, given)
        degrees = _native.as_i32(np.asarray(degrees).reshape(len(given), self.dim))
        handle = _native.c_void_p()
        _native.check_host(_native.host().gh_space_compile(
            domain.grid.handle, len(given), _native.ptr(_native.as_i64(given), _native.c_int64),
            _native.ptr(degrees, _native.c_int32), self.basis.kind, self.cg_mask, ctypes.byref(handle)))
        try:
            self._read(handle)
        finally:
            _native.host().gh_space_destroy(handle)

    def _read(self, handle):
        class Sizes(ctypes.Structure):
            _fields_ = [(n, ctypes.c_int64) for n in (
                "leaf_count", "dof_count", "mode_count", "term_count", "boundary_count",
                "interface_count", "boundary_variable_count", "constraint_row_count")]

        sz = Sizes()
        lib = _native.host()
        _native.check_host(lib.gh_space_sizes_get(handle, ctypes.byref(sz)))
        d = self.dim
        self.leaves = np.zeros(sz.leaf_count, dtype=np.int64)
        self.degrees = np.zeros((sz.leaf_count, d), dtype=np.int32)
        self.mode_offsets = np.zeros(sz.leaf_count + 1, dtype=np.int64)
        self.mode_box = np.zeros(sz.mode_count, dtype=np.int32)
        self.term_offsets = np.zeros(sz.mode_count + 1, dtype=np.int64)
        self.term_indices = np.zeros(sz.term_count, dtype=np.int64)
        self.term_coefficients = np.zeros(sz.term_count, dtype=np.float64)
        P = _native.ptr
        _native.check_host(lib.gh_space_export(
            handle, P(self.leaves, _native.c_int64), P(self.degrees, _native.c_int32),
            P(self.mode_offsets, _native.c_int64), P(self.mode_box, _native.c_int32),
            P(self.term_offsets, _native.c_int64), P(self.term_indices, _native.c_int64),
            P(self.term_coefficients, _native.c_double)))
        nb, ni = sz.boundary_count, sz.interface_count
        self.boundary_leaf = np.zeros(nb, dtype=np.int64)
        self.boundary_axis = np.zeros(nb, dtype=np.int32)
        self.boundary_side = np.zeros(nb, dtype=np.int32)
        self.interface_minus = np.zeros(ni, dtype=np.int64)
        self.interface_plus = np.zeros(ni, dtype=np.int64)
        self.interface_axis = np.zeros(ni, dtype=np.int32)
        _native.check_host(lib.gh_space_faces(
            handle, P(self.boundary_leaf, _native.c_int64), P(self.boundary_axis, _native.c_int32),
            P(self.boundary_side, _native.c_int32), P(self.interface_minus, _native.c_int64),
            P(self.interface_plus, _native.c_int64), P(self.interface_axis, _native.c_int32)))
        self.dof_count = int(sz.dof_count)
        self.boundary_variable_count = int(sz.boundary_variable_count)
        self.constraint_row_count = int(sz.constraint_row_count)
        self.leaf_index = {int(l): i for i, l in enumerate(self.leaves)}
        self.quadrature_shapes = np.asarray(
            [self.quadrature_policy.shape(deg) for deg in self.degrees], dtype=np.int32).reshape(-1, d)

    # Query API mirroring the reference (cell ids, 0-based dofs).
    def scalar_dof_count(self):
        return self.dof_count

    def active_leaf_count(self):
        return len(self.leaves)

    def cell_degrees(self, leaf):
        return tuple(int(v) for v in self.degrees[self.leaf_index[int(leaf)]])

    def local_modes(self, leaf):
        li = self.leaf_index[int(leaf)]
        shape = self.degrees[li] + 1
        boxes = self.mode_box[self.mode_offsets[li]:self.mode_offsets[li + 1]]
        return [tuple(int(v) for v in np.unravel_index(b, tuple(shape[::-1]))[::-1]) for b in boxes]

    def local_mode_count(self, leaf):
        li = self.leaf_index[int(leaf)]
        return int(self.mode_offsets[li + 1] - self.mode_offsets[li])

    def mode_terms(self, leaf, mode):
        """[(dof, coefficient), ...] of one local mode (empty if the mode is inactive)."""
        li = self.leaf_index[int(leaf)]
        shape = self.degrees[li] + 1
        box = 0
        stride = 1
        for a in range(self.dim):
            if not 0 <= mode[a] < shape[a]:
                return []
            box += int(mode[a]) * stride
            stride *= int(shape[a])
        lo, hi = self.mode_offsets[li], self.mode_offsets[li + 1]
        hits = np.nonzero(self.mode_box[lo:hi] == box)[0]
        if hits.size == 0:
            return []
        m = lo + int(hits[0])
        t0, t1 = self.term_offsets[m], self.term_offsets[m + 1]
        return [(int(i), float(c)) for i, c in zip(self.term_indices[t0:t1], self.term_coefficients[t0:t1])]

    def is_mode_active(self, leaf, mode):
        return len(self.mode_terms(leaf, mode)) > 0

    def cell_quadrature_shape(self, leaf):
        return tuple(int(v) for v in self.quadrature_shapes[self.leaf_index[int(leaf)]])


def scalar_dof_count(space):
    return space.scalar_dof_count()
