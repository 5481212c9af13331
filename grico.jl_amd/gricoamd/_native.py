"""ctypes bindings of the two native libraries.

The device library is mandatory for every operator/solver entry point: there
is no CPU fallback for the matrix-free action or the solvers, and a missing or
unloadable library raises immediately.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.join(os.path.dirname(_HERE), "lib")

c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_f64 = ctypes.POINTER(ctypes.c_double)
P_u8 = ctypes.POINTER(ctypes.c_uint8)

_host = None
_device = None


class NativeLibraryMissing(RuntimeError):
    pass


def library_path(name):
    return os.path.join(_LIBDIR, name)


def host():
    """The host plan-compiler library (libgrico_host.so)."""
    global _host
    if _host is None:
        path = library_path("libgrico_host.so")
        if not os.path.exists(path):
            raise NativeLibraryMissing(f"{path} is not built; run grico.jl_amd/build.py")
        lib = ctypes.CDLL(path)
        lib.gh_last_error.restype = ctypes.c_char_p
        lib.gh_gauss_legendre.argtypes = [c_int, P_f64, P_f64]
        lib.gh_fe_basis.argtypes = [c_double, c_int, P_f64, P_f64]
        lib.gh_restriction.argtypes = [c_int, c_int, c_int, c_int64, P_f64]
        lib.gh_grid_create.argtypes = [c_int, P_i64, ctypes.POINTER(c_void_p)]
        lib.gh_grid_destroy.argtypes = [c_void_p]
        lib.gh_grid_cell_count.argtypes = [c_void_p]
        lib.gh_grid_cell_count.restype = c_int64
        lib.gh_grid_refine.argtypes = [c_void_p, c_int64, c_int, P_i64]
        lib.gh_grid_export.argtypes = [c_void_p, P_i32, P_i64, P_i64, P_i64, P_i32, P_u8]
        lib.gh_space_compile.argtypes = [c_void_p, c_int64, P_i64, P_i32, c_int, c_int,
                                         ctypes.POINTER(c_void_p)]
        lib.gh_space_destroy.argtypes = [c_void_p]
        lib.gh_space_sizes_get.argtypes = [c_void_p, c_void_p]
        lib.gh_space_export.argtypes = [c_void_p, P_i64, P_i32, P_i64, P_i32, P_i64, P_i64, P_f64]
        lib.gh_space_faces.argtypes = [c_void_p, P_i64, P_i32, P_i32, P_i64, P_i64, P_i32]
        _host = lib
    return _host


def device():
    """The HIP device library (libgrico_mf.so); raises if it is not built."""
    global _device
    if _device is None:
        path = library_path("libgrico_mf.so")
        if not os.path.exists(path):
            raise NativeLibraryMissing(
                f"{path} is not built; the matrix-free operator has no CPU fallback")
        lib = ctypes.CDLL(path)
        from . import _devabi
        _devabi.declare(lib)
        _device = lib
    return _device


def check_host(status):
    if status != 0:
        msg = host().gh_last_error().decode()
        raise ValueError(msg)


def ptr(array, ctype):
    """Pointer to a contiguous numpy array (None -> NULL)."""
    if array is None:
        return ctypes.cast(None, ctypes.POINTER(ctype))
    assert array.flags["C_CONTIGUOUS"]
    return array.ctypes.data_as(ctypes.POINTER(ctype))


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
