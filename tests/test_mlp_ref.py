"""The numpy network (the checker of the HIP network kernels) against central
differences, and its BF16 rounding against exact cases."""
import numpy as np

from mlp_ref import bf16, loss_grad


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8, -2.5, 0.0, 3e38], np.float32)
    got = bf16(x)
    # BF16 keeps 7 fraction bits: 1 + 2^-7 is exact; 1 + 3*2^-8 is a tie that
    # rounds up to the even 1 + 2^-6; 1 + 2^-8 is a tie that rounds down to 1
    assert got[0] == 1.0 and got[1] == 1.0 + 2 ** -7 and got[2] == 1.0 + 2 ** -6 and got[3] == 1.0
    assert got[4] == -2.5 and got[5] == 0.0
    r = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    b = bf16(r)
    assert np.all(np.abs(b - r) <= np.abs(r) * 2.0 ** -8)
    assert np.all((b.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)


def test_gradient_matches_central_differences():
    n_in, n_hid, n_out, n = 7, 9, 5, 6
    rng = np.random.default_rng(1)
    npar = n_hid * n_in + n_hid + n_hid * n_hid + n_hid + n_out * n_hid + n_out
    p = rng.standard_normal(npar) * 0.5
    x = rng.standard_normal((n, n_in))
    t = rng.standard_normal((n, n_out))
    _, g = loss_grad(p, x, t, n_in, n_hid, n_out, rnd=False)
    h = 1e-6
    num = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        num[i] = (loss_grad(p + e, x, t, n_in, n_hid, n_out, rnd=False)[0] -
                  loss_grad(p - e, x, t, n_in, n_hid, n_out, rnd=False)[0]) / (2 * h)
    assert np.max(np.abs(num - g)) < 1e-7 * max(1.0, np.max(np.abs(g))), np.max(np.abs(num - g))
