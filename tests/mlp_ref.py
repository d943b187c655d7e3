"""numpy restatement of the learned start's network (bcd_mlp_*), the checker
of its HIP forward and backward kernels.

The network is n_in -> n_hid -> n_hid -> n_out with ReLU between layers;
parameters are laid out as W1 [n_hid x n_in], b1, W2 [n_hid x n_hid], b2,
W3 [n_out x n_hid], b3.  With bf16=True every product operand is rounded to
BF16 (round to nearest even) as the matrix-core kernels round it, and the
sums are taken in float64 (the kernels accumulate in FP32, so the two agree
to FP32 rounding).  With bf16=False the arithmetic is exact float64, which
central differences check."""
import numpy as np


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p, n_in, n_hid, n_out):
    sizes = [n_hid * n_in, n_hid, n_hid * n_hid, n_hid, n_out * n_hid, n_out]
    out, o = [], 0
    for s in sizes:
        out.append(np.asarray(p[o:o + s]))
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(n_hid, n_in), b1, W2.reshape(n_hid, n_hid), b2, W3.reshape(n_out, n_hid), b3


def forward(p, x, n_in, n_hid, n_out, rnd=True):
    r = bf16 if rnd else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, n_hid, n_out)
    h1 = np.maximum(r(x) @ r(W1).T + b1, 0.0)
    h2 = np.maximum(r(h1) @ r(W2).T + b2, 0.0)
    y = r(h2) @ r(W3).T + b3
    return h1, h2, y


def loss_grad(p, x, t, n_in, n_hid, n_out, rnd=True):
    """Mean squared loss over all outputs and its gradient, as bcd_mlp_loss_grad."""
    r = bf16 if rnd else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, n_hid, n_out)
    h1, h2, y = forward(p, x, n_in, n_hid, n_out, rnd)
    if rnd:   # the kernels keep activations in FP32 between layers
        h1, h2, y = h1.astype(np.float32), h2.astype(np.float32), y.astype(np.float32)
    d = y - t
    loss = float(np.mean(np.asarray(d, np.float64) ** 2))
    dy = 2.0 * d / d.size
    if rnd:
        dy = dy.astype(np.float32)
    gW3 = r(dy).T @ r(h2)
    gb3 = dy.sum(0)
    dh2 = (r(dy) @ r(W3)) * (h2 > 0)
    if rnd:
        dh2 = dh2.astype(np.float32)
    gW2 = r(dh2).T @ r(h1)
    gb2 = dh2.sum(0)
    dh1 = (r(dh2) @ r(W2)) * (h1 > 0)
    if rnd:
        dh1 = dh1.astype(np.float32)
    gW1 = r(dh1).T @ r(x)
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
