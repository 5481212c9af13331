"""Pins the MLP oracle (oracle/mlp_oracle.py): BF16 rounding known answers and
backpropagation against central finite differences of the float64 network."""

import numpy as np
import pytest

import mlp_oracle as M
from gricoamd.learned import init_parameters


def test_bf16_rounding_known_answers():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 spacing at 1
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38, 0.0],
                    dtype=np.float32)
    out = M.to_bf16(vals)
    assert out[0] == one
    assert out[1] == one                      # tie -> even (1.0)
    assert out[2] == one + 2 * ulp            # tie -> even (1 + 2^-6)
    assert out[3] == one                      # below half ulp
    assert out[4] == np.float32(-2.5)         # exactly representable
    assert np.isfinite(out[5])
    assert out[6] == 0.0


def test_backprop_matches_finite_differences():
    widths = (6, 8, 7, 5)
    rng = np.random.default_rng(0)
    p = init_parameters(widths, seed=3).astype(np.float64)
    p[p.size - 5:] = rng.standard_normal(5) * 0.1  # nonzero output bias
    x = rng.standard_normal((9, 6))
    t = rng.standard_normal((9, 5))
    _, g = M.loss_and_gradient(p, widths, x, t, bf16=False)
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = M.loss_and_gradient(p + e, widths, x, t, bf16=False)
        lm, _ = M.loss_and_gradient(p - e, widths, x, t, bf16=False)
        fd[i] = (lp - lm) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_bf16_oracle_close_to_float64():
    widths = (64, 256, 256, 64)
    rng = np.random.default_rng(1)
    p = init_parameters(widths, seed=1)
    x = rng.standard_normal((32, 64)).astype(np.float32)
    y16, _ = M.forward(p, widths, x, bf16=True)
    y64, _ = M.forward(p, widths, x, bf16=False)
    assert np.abs(y16 - y64).max() <= 3e-2 * np.abs(y64).max()
