"""ORACLE (test infrastructure only): numpy restatement of the per-cell MLP of
libgrico_mf.so (forward, mean-squared-error loss, backpropagation), with the
device's BF16 rounding points emulated: inputs, weights and hidden activations
rounded to BF16 (round to nearest even), products accumulated in float64 here
(FP32 on the device), biases and the output in FP32, the loss gradient dy
rounded to BF16 before it enters the backward products.

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


def loss_and_gradient(params, widths, x, t, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if bf16 else np.float64), widths)
    y, acts = forward(params, widths, x, bf16)
    d = np.asarray(y, dtype=np.float64) - np.asarray(t, dtype=np.float64)
    count = d.size
    loss = float((d * d).sum() / count)
    dz = r((2.0 / count) * d.astype(np.float32 if bf16 else np.float64)).astype(np.float64)
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
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat
