"""Reference MLP for the surrogate's device kernels.  TEST INFRASTRUCTURE ONLY.

Float64 numpy forward and backward of the same network as casim_mlp_*:
layers W_l [in x out] + b_l, SiLU on hidden layers, identity output, mean
squared error in normalised units.  With bf16=True the weights and the
activations entering each GEMM are rounded to BF16 (round to nearest even)
exactly where the device rounds them, so the two differ only by FP32
accumulation order.  The backward pass is pinned by finite differences
(tests/test_mlp.py).
"""
import numpy as np


def to_bf16(a):
    """Round float values to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(n_in, width, depth):
    fin = [n_in] + [width] * depth
    fout = [width] * depth + [1]
    off, spans = 0, []
    for a, b in zip(fin, fout):
        spans.append((off, off + a * b, off + a * b + b, a, b))
        off += a * b + b
    return spans, off


def unpack(params, n_in, width, depth):
    spans, _ = layout(n_in, width, depth)
    return [(params[w0:w1].reshape(a, b).astype(np.float64), params[w1:b1].astype(np.float64))
            for w0, w1, b1, a, b in spans]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, n_in, width, depth, bf16=False):
    layers = unpack(params, n_in, width, depth)
    r = to_bf16 if bf16 else (lambda v: v)
    h = r(x)
    hs, zs = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ r(W) + b
        zs.append(z)
        if l < depth:
            h = r(silu(z))
            hs.append(h)
    return zs[-1][:, 0], hs, zs, layers


def loss_and_grad(params, x, y, n_in, width, depth, bf16=False):
    """MSE and its gradient in the packed parameter layout."""
    r = to_bf16 if bf16 else (lambda v: v)
    pred, hs, zs, layers = forward(params, x, n_in, width, depth, bf16)
    n = len(y)
    e = pred - y
    loss = float(np.mean(e * e))
    dz = (2.0 * e / n)[:, None]
    grads = [None] * len(layers)
    for l in range(depth, -1, -1):
        W, b = layers[l]
        dzr = r(dz)
        grads[l] = (hs[l].T @ dzr, dz.sum(axis=0))
        if l == 0:
            break
        dh = dzr @ r(W).T
        dz = dh * dsilu(zs[l - 1])
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return loss, flat
