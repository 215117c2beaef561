"""FP64 numpy restatement of the warm-start network (test infrastructure only).

Only tests/ use this module, as the checker of the device MLP kernels
(iceicle_amd/csrc/mlp.hip).  The network has no reference counterpart; the
mathematics it restates is the textbook fully connected network
H_{l+1} = act(H_l W_l + b_l) (tanh on hidden layers, identity on the
output), the mean squared error loss and its gradient by backpropagation,
with parameters packed layer by layer as W_0 [w0][w1], b_0, W_1, b_1, ...
The gradient is pinned by central finite differences in tests/test_mlp.py.
"""
import numpy as np


def unpack(widths, p):
    """Views (W_l, b_l) of the packed parameter vector."""
    out, off = [], 0
    for l in range(len(widths) - 1):
        n = widths[l] * widths[l + 1]
        W = p[off:off + n].reshape(widths[l], widths[l + 1])
        off += n
        b = p[off:off + widths[l + 1]]
        off += widths[l + 1]
        out.append((W, b))
    if off != p.size:
        raise ValueError("parameter count does not match the widths")
    return out


def forward(widths, p, x, keep=False):
    layers = unpack(widths, np.asarray(p, dtype=np.float64))
    h = np.asarray(x, dtype=np.float64)
    hs = [h]
    for l, (W, b) in enumerate(layers):
        z = h @ W + b
        h = np.tanh(z) if l < len(layers) - 1 else z
        hs.append(h)
    return (h, hs) if keep else h


def loss(widths, p, x, t):
    y = forward(widths, p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(widths, p, x, t):
    """(loss, d loss / d p) for the mean squared error over all outputs."""
    p = np.asarray(p, dtype=np.float64)
    layers = unpack(widths, p)
    y, hs = forward(widths, p, x, keep=True)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    d = 2.0 * (y - t) / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (hs[l].T @ d, d.sum(axis=0))
        if l > 0:
            d = (d @ W.T) * (1.0 - hs[l] ** 2)
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return float(np.mean((y - t) ** 2)), g
