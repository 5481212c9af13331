"""ORACLE (test infrastructure only): numpy restatement of the per-cell MLP of
libgrico_mf.so (forward, mean-squared-error loss, backpropagation), with the
device's BF16 rounding points emulated: inputs, weights and hidden activations
rounded to BF16 (round to nearest even), products accumulated in float64 here
(FP32 on the device), biases and the output in FP32, the loss gradient dy
rounded to BF16 before it enters the backward products.

energy_loss_and_gradient restates the device's training objective of the
learned block correction (gmf_learned_train / gmf_learned_gradient): the
residual's per-cell patches scaled to unit norm, the network, the outputs
rescaled per cell, averaged on shared dofs and assembled into a correction c,
and the energy-norm error ratio ||e' - c||_A^2 / ||e'||_A^2 with its gradient.

Only tests/ use this module. With bf16=False it is the exact float64 network,
pinned in the tests by central finite differences of its own loss.
"""

import numpy as np


def to_bf16(a):
    """Round float32 values to the nearest BF16 (ties to even); returns float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def unpack(params, widths):
    out, o = [], 0
    for k, n in zip(widths[:-1], widths[1:]):
        W = params[o:o + k * n].reshape(k, n)
        o += k * n
        b = params[o:o + n]
        o += n
        out.append((W, b))
    return out


def forward(params, widths, x, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if bf16 else np.float64), widths)
    acts = [r(x)]
    h = acts[0].astype(np.float64)
    for i, (W, b) in enumerate(layers):
        z = h @ r(W).astype(np.float64) + np.asarray(b, dtype=np.float64)
        if i < len(layers) - 1:
            h = r(np.maximum(z, 0.0).astype(np.float32 if bf16 else np.float64)).astype(np.float64)
            acts.append(h)
        else:
            y = z.astype(np.float32) if bf16 else z
    return y, acts


def _backprop(layers, acts, dz, bf16):
    """Parameter gradient for the output gradient dz (already rounded as the
    device rounds it) through the activations of a forward pass."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        a = acts[i].astype(np.float64)
        grads.append((a.T @ dz, dz.sum(axis=0)))
        if i > 0:
            dh = dz @ r(W).astype(np.float64).T
            dh = np.where(acts[i] > 0, dh, 0.0)
            dz = r(dh.astype(np.float32 if bf16 else np.float64)).astype(np.float64)
    grads.reverse()
    return np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])


def loss_and_gradient(params, widths, x, t, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if bf16 else np.float64), widths)
    y, acts = forward(params, widths, x, bf16)
    d = np.asarray(y, dtype=np.float64) - np.asarray(t, dtype=np.float64)
    count = d.size
    loss = float((d * d).sum() / count)
    dz = r((2.0 / count) * d.astype(np.float32 if bf16 else np.float64)).astype(np.float64)
    return loss, _backprop(layers, acts, dz, bf16)


def patches(residual, ridx):
    """Per-cell patches of a reduced vector (ridx: cells x width reduced dof
    indices, -1 for none) scaled to unit Euclidean norm, and the norms."""
    loc = np.where(ridx >= 0, residual[np.maximum(ridx, 0)], 0.0)
    scale = np.sqrt((loc * loc).sum(axis=1))
    inv = np.where(scale > 0, 1.0 / np.where(scale > 0, scale, 1.0), 0.0)
    return loc * inv[:, None], scale


def share_weights(ridx, n):
    """1 / (number of cells holding the dof) per reduced dof."""
    count = np.zeros(n)
    np.add.at(count, ridx[ridx >= 0], 1.0)
    return np.where(count > 0, 1.0 / np.where(count > 0, count, 1.0), 0.0)


def assemble(Y, scale, cw, ridx, n):
    """c = sum over cells of P_c^T (w * Y_c * scale_c / cw_c), w the share weights."""
    pu = share_weights(ridx, n)
    vals = Y * (scale / cw)[:, None]
    c = np.zeros(n)
    m = ridx >= 0
    np.add.at(c, ridx[m], (pu[np.maximum(ridx, 0)] * vals)[m])
    return c


def energy_loss_and_gradient(params, widths, A, residual, energy, ridx, cw, bf16=True):
    """Training objective of the learned block correction on one residual r'
    with energy E = e'^T r' (A: dense reduced operator, A e' = r'):
    loss = 1 + (c^T A c - 2 c^T r') / E = ||e' - c||_A^2 / E, and its gradient
    over the packed parameters. With bf16 the device's rounding points are
    emulated (FP32 patches, norms and outputs, BF16 network and output
    gradient)."""
    n = len(residual)
    X, scale = patches(residual, ridx)
    cw = np.asarray(cw, dtype=np.float64)
    if bf16:
        X = X.astype(np.float32)
        scale = scale.astype(np.float32).astype(np.float64)
    layers = unpack(np.asarray(params, dtype=np.float32 if bf16 else np.float64), widths)
    Y, acts = forward(params, widths, X, bf16)
    Y = np.asarray(Y, dtype=np.float64)
    c = assemble(Y, scale, cw, ridx, n)
    q = A @ c
    loss = 1.0 + (c @ q - 2.0 * (c @ residual)) / energy
    g = 2.0 * (q - residual) / energy
    pu = share_weights(ridx, n)
    dY = np.where(ridx >= 0, pu[np.maximum(ridx, 0)] * g[np.maximum(ridx, 0)], 0.0) * (scale / cw)[:, None]
    if bf16:
        dY = to_bf16(dY.astype(np.float32)).astype(np.float64)
    return float(loss), _backprop(layers, acts, dY, bf16)
