"""bcd_reduce_mtx_lncz_noalloc (the Lanczos reduction without device
allocation, buffers in the library's fixed per-device scratch) against
bcd_reduce_mtx_lncz on the same inputs: the same kernels on the same data,
so every output bit, the return value, i_piv and the in-place pivot
normalisation must match.  Two calls in a row, the second smaller (the
scratch keeps the first call's contents), then a call beyond the scratch."""
import numpy as np
import pytest

import pyoracle as po
from bcdor_amd import capi
from helpers import load_golden

pytestmark = pytest.mark.gpu


def case(z, ms, i, side):
    rows = po.sort_rows(z[f"{side}_{i}"], side == "bk")
    piv = z[f"piv_vec_{i}"]
    itr = z[f"piv_{side}_{i}"].astype(np.int32)
    tot = min(int(itr[itr > 0].sum()), rows.shape[0])
    return rows, piv, itr, tot, int(ms.no[i])


def test_noalloc_equals_allocating_entry():
    ms, vpp, g, z = load_golden("c2_hf_split50")
    ctx = capi.Context(ms, vpp)
    # the largest channel first, then a smaller one: the second call runs in
    # a scratch region that still holds the first call's data
    sizes = sorted(((z[f"fw_{int(i)}"].shape[0], int(i)) for i in z["channels"]), reverse=True)
    calls = [(sizes[0][1], "fw"), (sizes[-1][1], "bk"), (sizes[0][1], "fw")]
    assert z[f"bk_{calls[1][0]}"].shape[0] < z[f"fw_{calls[0][0]}"].shape[0]
    for i, side in calls:
        rows, piv, itr, tot, n = case(z, ms, i, side)
        na_out = max(2, len(itr) + 1)
        print(f"noalloc channel {i} {side}: {rows.shape[0]} poles, {tot} vectors", flush=True)
        a = ctx.reduce_mtx_lncz(na_out, n, tot, rows, 2, n, piv.copy(), itr.copy())
        b = ctx.reduce_mtx_lncz(na_out, n, tot, rows, 2, n, piv.copy(), itr.copy(), noalloc=True)
        assert a[0] == b[0] and a[0] > 0
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2][:a[2][0] + 2], b[2][:b[2][0] + 2])
        assert np.array_equal(a[3], b[3])


def test_noalloc_rejects_what_exceeds_its_scratch():
    ms, vpp, g, z = load_golden("c1_hf_split50")
    ctx = capi.Context(ms, vpp)
    cap = capi.lib().bcd_noalloc_scratch_bytes()
    # a pole count whose Krylov workspace alone outgrows the scratch: the
    # entry refuses it on the host, before any device work
    n = 2
    M = 40000
    rows = np.zeros((M, 2 + n))
    rows[:, 0] = np.arange(M, dtype=np.float64)
    rows[:, 2:] = 1.0
    itr = np.array([cap // (8 * M)], np.int32)
    assert 8 * M * int(itr[0]) > cap - 8 * M * (n + 8)
    r = ctx.reduce_mtx_lncz(2, n, int(itr[0]), rows, 2, n, np.ones((1, n)), itr, noalloc=True)
    assert r[0] == -1003
