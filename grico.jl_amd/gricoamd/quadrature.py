"""One-dimensional polynomials and Gauss-Legendre rules (host lib wrappers).

Reference: src/polynomials.jl (integrated Legendre family 67-115, degree-zero
constant 238-247) and src/quadrature.jl:342-365 (gauss_legendre_rule).
"""

import functools

import numpy as np

from . import _native


@functools.lru_cache(maxsize=None)
def _gauss(n):
    pts = np.zeros(n)
    wts = np.zeros(n)
    _native.check_host(_native.host().gh_gauss_legendre(
        n, _native.ptr(pts, _native.c_double), _native.ptr(wts, _native.c_double)))
    pts.setflags(write=False)
    wts.setflags(write=False)
    return pts, wts


def gauss_legendre_rule(point_count):
    """(points, weights) of the n-point rule on [-1, 1], points ascending."""
    if int(point_count) < 1:
        raise ValueError("point_count must be positive")
    return _gauss(int(point_count))


def fe_basis_values_and_derivatives(x, degree):
    """psi_0..psi_p and derivatives at each x (arrays of shape (len(x), p+1))."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    vals = np.zeros((x.size, degree + 1))
    ders = np.zeros((x.size, degree + 1))
    v = np.zeros(degree + 1)
    d = np.zeros(degree + 1)
    lib = _native.host()
    for i, xi in enumerate(x):
        _native.check_host(lib.gh_fe_basis(float(xi), int(degree), _native.ptr(v, _native.c_double),
                                           _native.ptr(d, _native.c_double)))
        vals[i] = v
        ders[i] = d
    return vals, ders


@functools.lru_cache(maxsize=None)
def basis_tables(degree, point_count):
    """B[i, q] = psi_i(xi_q), G[i, q] = psi_i'(xi_q) on the Gauss points."""
    pts, _ = _gauss(point_count)
    vals, ders = fe_basis_values_and_derivatives(pts, degree)
    B = np.ascontiguousarray(vals.T)
    G = np.ascontiguousarray(ders.T)
    B.setflags(write=False)
    G.setflags(write=False)
    return B, G


def restriction_matrix(source_degree, target_degree, delta, relative):
    out = np.zeros((target_degree + 1, source_degree + 1))
    _native.check_host(_native.host().gh_restriction(int(source_degree), int(target_degree),
                                                     int(delta), int(relative),
                                                     _native.ptr(out, _native.c_double)))
    return out
