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


# ---- energy-norm training objective of the learned block correction ----------

@pytest.fixture(scope="module")
def small_lognormal():
    """2^3 cells, p = 3 (64 local modes), log-normal coefficient: dense reduced
    operator, reduced patch maps and cell weights."""
    import gricoamd as g
    import pyoracle
    import problems as P
    from gricoamd.learned import cell_weights
    plan = g.compile(P.lognormal_poisson(depth=1, degree=3, seed=3))
    b = plan.buckets[0]
    red = plan.reduced_index
    dofs = np.asarray(b.local_index)
    ridx = np.where(dofs >= 0, red[np.maximum(dofs, 0)], -1)
    A = pyoracle.reduced_matrix(plan)
    return A, ridx, cell_weights(plan).astype(np.float64)


def test_patches_and_share_weights_by_loops(small_lognormal):
    A, ridx, _ = small_lognormal
    n = A.shape[0]
    r = np.random.default_rng(4).standard_normal(n)
    X, scale = M.patches(r, ridx)
    pu = M.share_weights(ridx, n)
    for c in range(ridx.shape[0]):
        loc = [r[j] if j >= 0 else 0.0 for j in ridx[c]]
        nrm = np.sqrt(sum(v * v for v in loc))
        assert scale[c] == pytest.approx(nrm, rel=1e-14)
        assert np.allclose(X[c], np.array(loc) / nrm, rtol=1e-14, atol=0)
    for j in range(n):
        holders = sum(int(j in row) for row in ridx.tolist())
        assert pu[j] == pytest.approx(1.0 / holders)


def test_energy_loss_is_energy_error_ratio(small_lognormal):
    """The restated objective equals ||e' - c||_A^2 / ||e'||_A^2 with e' solved
    from A e' = r' (an independent route to the same number)."""
    A, ridx, cw = small_lognormal
    widths = (64, 16, 12, 64)
    rng = np.random.default_rng(5)
    p = init_parameters(widths, seed=6).astype(np.float64)
    p[p.size - 64:] = rng.standard_normal(64) * 0.05
    r = rng.standard_normal(A.shape[0])
    e = np.linalg.solve(A, r)
    energy = e @ A @ e
    loss, _ = M.energy_loss_and_gradient(p, widths, A, r, energy, ridx, cw, bf16=False)
    X, scale = M.patches(r, ridx)
    Y, _ = M.forward(p, widths, X, bf16=False)
    c = M.assemble(Y, scale, cw, ridx, len(r))
    d = e - c
    assert loss == pytest.approx(d @ A @ d / energy, rel=1e-12)
    # a zero network output leaves the whole error: ratio 1
    z = p.copy()
    z[:] = 0.0
    l0, _ = M.energy_loss_and_gradient(z, widths, A, r, energy, ridx, cw, bf16=False)
    assert l0 == pytest.approx(1.0, abs=1e-14)


def test_energy_gradient_matches_finite_differences(small_lognormal):
    A, ridx, cw = small_lognormal
    widths = (64, 16, 12, 64)
    rng = np.random.default_rng(7)
    p = init_parameters(widths, seed=8).astype(np.float64)
    p[p.size - 64:] = rng.standard_normal(64) * 0.05
    r = rng.standard_normal(A.shape[0])
    e = np.linalg.solve(A, r)
    energy = e @ A @ e
    _, gr = M.energy_loss_and_gradient(p, widths, A, r, energy, ridx, cw, bf16=False)
    h = 1e-6
    idx = np.unique(np.concatenate([np.arange(0, p.size, 7), np.arange(p.size - 64, p.size)]))
    for i in idx:
        d = np.zeros_like(p)
        d[i] = h
        lp, _ = M.energy_loss_and_gradient(p + d, widths, A, r, energy, ridx, cw, bf16=False)
        lm, _ = M.energy_loss_and_gradient(p - d, widths, A, r, energy, ridx, cw, bf16=False)
        assert gr[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-8 * np.abs(gr).max())


def test_assemble_by_loops(small_lognormal):
    A, ridx, cw = small_lognormal
    n = A.shape[0]
    rng = np.random.default_rng(9)
    Y = rng.standard_normal(ridx.shape)
    scale = rng.uniform(0.5, 2.0, ridx.shape[0])
    c = M.assemble(Y, scale, cw, ridx, n)
    ref = np.zeros(n)
    for cell in range(ridx.shape[0]):
        for i, j in enumerate(ridx[cell]):
            if j >= 0:
                holders = int((ridx == j).any(axis=1).sum())
                ref[j] += Y[cell, i] * scale[cell] / cw[cell] / holders
    assert np.allclose(c, ref, rtol=1e-14, atol=0)
