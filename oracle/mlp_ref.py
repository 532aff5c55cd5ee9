"""ORACLE — test infrastructure only.

FP64 numpy forward and backward of the thermal-scattering surrogate MLP
(widths w, SiLU between layers, loss L = sum (f(x) - y)^2 / (2 n n_out)),
in the parameter layout of the C ABI: per layer W [out x in] row-major then
b [out].  The network has no counterpart in the reference; this restatement
is pinned by central finite differences in tests/test_oracle_mlp.py.
"""
import numpy as np


def unpack(params, widths):
    layers, off = [], 0
    for i in range(len(widths) - 1):
        n_in, n_out = widths[i], widths[i + 1]
        W = params[off:off + n_in * n_out].reshape(n_out, n_in)
        off += n_in * n_out
        b = params[off:off + n_out]
        off += n_out
        layers.append((W, b))
    return layers


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(params, widths, x):
    h = np.asarray(x, dtype=np.float64)
    layers = unpack(np.asarray(params, dtype=np.float64), widths)
    for i, (W, b) in enumerate(layers):
        z = h @ W.T + b
        h = silu(z) if i + 1 < len(layers) else z
    return h


def loss_and_grad(params, widths, x, y):
    params = np.asarray(params, dtype=np.float64)
    layers = unpack(params, widths)
    hs, zs = [np.asarray(x, dtype=np.float64)], []
    for i, (W, b) in enumerate(layers):
        z = hs[-1] @ W.T + b
        zs.append(z)
        hs.append(silu(z) if i + 1 < len(layers) else z)
    out = hs[-1]
    n, no = out.shape
    d = out - y
    loss = float((d * d).sum() / (2 * n * no))
    dz = d / (n * no)
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        grads.append((dz.T @ hs[i], dz.sum(0)))
        if i > 0:
            dh = dz @ W
            z = zs[i - 1]
            s = 1.0 / (1.0 + np.exp(-z))
            dz = dh * s * (1 + z * (1 - s))
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
