"""Cartesian domains on a binary refinement tree.

Reference: src/geometry.jl (Domain 114-141; cell size = ldexp(extent/root, -level)
at 689-695, lower = origin + size*coord at 675-680, biunit map center + xi*half
at 560-569) and src/topology.jl / src/refinement.jl for the tree. Cell ids and
axes are 0-based here (the reference is 1-based).
"""

import ctypes

import numpy as np

from . import _native

LOWER = 0
UPPER = 1


class CartesianGrid:
    """Owning wrapper of the native refinement tree."""

    def __init__(self, root_counts):
        self.root_counts = tuple(int(c) for c in root_counts)
        self.dim = len(self.root_counts)
        handle = _native.c_void_p()
        rc = _native.as_i64(self.root_counts)
        _native.check_host(_native.host().gh_grid_create(self.dim, _native.ptr(rc, _native.c_int64),
                                                          ctypes.byref(handle)))
        self._handle = handle
        self._revision = 0
        self._cache = None

    def __del__(self):
        h = getattr(self, "_handle", None)
        if _native is None:  # interpreter shutdown
            return
        if h is not None and h.value:
            _native.host().gh_grid_destroy(h)
            self._handle = None

    @property
    def handle(self):
        return self._handle

    def refine(self, cell, axis):
        """Bisect active leaf `cell` normal to `axis`; returns the first child id."""
        first = _native.c_int64()
        _native.check_host(_native.host().gh_grid_refine(self._handle, int(cell), int(axis),
                                                          ctypes.byref(first)))
        self._revision += 1
        self._cache = None
        return first.value

    def cell_count(self):
        return int(_native.host().gh_grid_cell_count(self._handle))

    def _export(self):
        if self._cache is None:
            n = self.cell_count()
            d = self.dim
            levels = np.zeros((n, d), dtype=np.int32)
            coords = np.zeros((n, d), dtype=np.int64)
            parent = np.zeros(n, dtype=np.int64)
            first_child = np.zeros(n, dtype=np.int64)
            split = np.zeros(n, dtype=np.int32)
            active = np.zeros(n, dtype=np.uint8)
            _native.check_host(_native.host().gh_grid_export(
                self._handle, _native.ptr(levels, _native.c_int32), _native.ptr(coords, _native.c_int64),
                _native.ptr(parent, _native.c_int64), _native.ptr(first_child, _native.c_int64),
                _native.ptr(split, _native.c_int32), _native.ptr(active, ctypes.c_uint8)))
            self._cache = dict(levels=levels, coords=coords, parent=parent, first_child=first_child,
                               split=split, active=active.astype(bool))
        return self._cache

    def levels(self):
        return self._export()["levels"]

    def coords(self):
        return self._export()["coords"]

    def parent(self):
        return self._export()["parent"]

    def first_child(self):
        return self._export()["first_child"]

    def split_axis(self):
        return self._export()["split"]

    def active_leaves(self):
        return np.nonzero(self._export()["active"])[0].astype(np.int64)


class Domain:
    """Axis-aligned box [origin, origin+extent] on a root grid (`Domain` in the reference)."""

    def __init__(self, origin, extent, root_counts):
        origin = tuple(float(v) for v in origin)
        extent = tuple(float(v) for v in extent)
        if not (len(origin) == len(extent) == len(root_counts)):
            raise ValueError("origin, extent and root_counts must have the same dimension")
        if any(e <= 0 for e in extent):
            raise ValueError("extent must be positive")
        self.origin = origin
        self.extent = extent
        self.grid = CartesianGrid(root_counts)
        self.dim = len(origin)

    def refine(self, cell, axis):
        return self.grid.refine(cell, axis)

    def active_leaves(self):
        return self.grid.active_leaves()

    # Geometry, vectorized over cell ids.
    def cell_sizes(self, cells):
        cells = np.asarray(cells, dtype=np.int64)
        lev = self.grid.levels()[cells]
        out = np.empty(lev.shape)
        for a in range(self.dim):
            out[:, a] = np.ldexp(self.extent[a] / self.grid.root_counts[a], -lev[:, a])
        return out

    def cell_lowers(self, cells):
        cells = np.asarray(cells, dtype=np.int64)
        size = self.cell_sizes(cells)
        crd = self.grid.coords()[cells]
        return np.asarray(self.origin)[None, :] + size * crd

    def cell_centers_and_half_sizes(self, cells):
        size = self.cell_sizes(cells)
        lower = self.cell_lowers(cells)
        half = 0.5 * size
        return lower + half, half


def uniform_refine(domain, depth):
    """Refine every active leaf along every axis `depth` times (axis 1 first).

    Mirrors benchmark/hpmg_scaling.jl:121-150 (_refine_leaf_all_axes!,
    _uniform_refine!).
    """
    for _ in range(depth):
        for leaf in domain.active_leaves().tolist():
            leaves = [leaf]
            for axis in range(domain.dim):
                nxt = []
                for cur in leaves:
                    first = domain.refine(cur, axis)
                    nxt.extend((first, first + 1))
                leaves = nxt
    return domain
