"""The learned start's network kernels (bcd_mlp_*) against the numpy
restatement (tests/mlp_ref.py): forward values, loss and every parameter
gradient, at the network's real shape (64 -> 256 -> 256 -> 32) and a row
count that is not a multiple of the 64-row tiles; then training on the
device.

Bar: the kernels round every product operand to BF16 as mlp_ref does and
accumulate in FP32 where mlp_ref accumulates in float64, so values agree to
FP32 rounding of the sums: forward outputs and the loss to 1e-4 relative, the
gradient to 1e-3 relative in norm (a pre-activation within FP32 rounding of
zero can take the other ReLU branch, or an activation the other BF16
neighbour)."""
import numpy as np
import pytest

from bcdor_amd import capi
from mlp_ref import forward, loss_grad

pytestmark = pytest.mark.gpu

SHAPE = (64, 256, 32)


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, SHAPE[0])).astype(np.float32)
    A = rng.standard_normal((SHAPE[0], SHAPE[2])) / 8.0
    t = np.tanh(x @ A).astype(np.float32)
    return x, t


def test_forward_and_gradient_match_numpy():
    print("mlp forward/backward n=300", flush=True)
    net = capi.Mlp(*SHAPE, seed=3)
    p = net.params()
    rng = np.random.default_rng(4)
    p = (p + 0.05 * rng.standard_normal(p.size)).astype(np.float32)   # non-zero biases too
    net.set_params(p)
    x, t = data(300, 5)
    y = net.forward(x)
    _, _, y_ref = forward(p, x, *SHAPE)
    assert np.max(np.abs(y - y_ref)) < 1e-4 * np.max(np.abs(y_ref)), np.max(np.abs(y - y_ref))
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref = loss_grad(p, x, t, *SHAPE)
    assert abs(loss - loss_ref) < 1e-4 * loss_ref, (loss, loss_ref)
    n_in, n_hid, n_out = SHAPE
    sizes = [n_hid * n_in, n_hid, n_hid * n_hid, n_hid, n_out * n_hid, n_out]
    o = 0
    for name, s in zip(["W1", "b1", "W2", "b2", "W3", "b3"], sizes):
        a, b = g[o:o + s], g_ref[o:o + s]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"  {name}: relative gradient difference {rel:.2e}", flush=True)
        assert rel < 1e-3, (name, rel)
        o += s


def test_fit_on_device():
    """Adam on the device lowers the loss of a smooth regression tenfold, and
    the same seed gives the same parameters bit for bit."""
    print("mlp fit", flush=True)
    x, t = data(4096, 6)
    runs = []
    for _ in range(2):
        net = capi.Mlp(*SHAPE, seed=7)
        hist = net.fit(x, t, epochs=40, batch=512, lr=1e-3, seed=8)
        runs.append(net.params())
        print(f"  loss {hist[0]:.4e} -> {hist[-1]:.4e}", flush=True)
        assert hist[-1] < 0.1 * hist[0]
    assert np.array_equal(runs[0], runs[1])
