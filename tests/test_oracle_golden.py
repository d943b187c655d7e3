"""The CPU oracle against golden vectors of the compiled reference.

Golden vectors: tests/golden/*.npz, written by tests/golden/make_golden.py from
oracle/_ref/ref_harness (reference sources compiled as they are).  Bar:
bit-exact.  The oracle follows the reference's operation order and both are
built for x86-64 without FMA contraction, so every double must match.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import load_golden, rel_err

CASES = ["c1_hf_split50", "c1_hf_piv10", "c1_dressed", "c2_hf_split50"]


@pytest.fixture(scope="module", params=CASES)
def case(request):
    ms, vpp, g, z = load_golden(request.param)
    S = po.System(ms, vpp)
    P = po.Prop.from_propagator(S, g)
    return request.param, ms, S, P, z


def test_interaction_count(case):
    name, ms, S, P, z = case
    assert S.n_me() == int(z["nme_ref"])


def test_pole_counts(case):
    name, ms, S, P, z = case
    for i in z["channels"]:
        assert po.count_poles(S, int(i), P) == tuple(z[f"count_{i}"])


def test_sigma_inf(case):
    name, ms, S, P, z = case
    for i in z["channels"]:
        u = po.hf_pot(S, int(i), P)
        assert np.array_equal(u, z[f"hf_{i}"]), (name, i, rel_err(u, z[f"hf_{i}"]))


def test_sigma2_rows(case):
    name, ms, S, P, z = case
    for i in z["channels"]:
        fw, bk = po.build_sigma2(S, int(i), P)
        assert fw.shape == z[f"fw_{i}"].shape and bk.shape == z[f"bk_{i}"].shape
        assert np.array_equal(fw, z[f"fw_{i}"]), (name, i, "fw", rel_err(fw, z[f"fw_{i}"]))
        assert np.array_equal(bk, z[f"bk_{i}"]), (name, i, "bk", rel_err(bk, z[f"bk_{i}"]))


def test_lanczos_rows(case):
    name, ms, S, P, z = case
    n_checked = 0
    for i in z["channels"]:
        if f"lfw_{i}" not in z.files:
            continue
        fw, bk = z[f"fw_{i}"], z[f"bk_{i}"]
        ofw, pfw, obk, pbk, den, ifw, ibk = po.reduce_selfen_lncz(
            int(ms.no[i]), fw, bk, z[f"piv_vec_{i}"], z[f"piv_fw_{i}"], z[f"piv_bk_{i}"])
        assert np.array_equal(np.stack([ifw, ibk]), z[f"itr_{i}"]), (name, i, "rescaled n_itr")
        assert ofw.shape == z[f"lfw_{i}"].shape and obk.shape == z[f"lbk_{i}"].shape, (name, i)
        assert np.array_equal(ofw, z[f"lfw_{i}"]), (name, i, "fw", rel_err(ofw, z[f"lfw_{i}"]))
        assert np.array_equal(obk, z[f"lbk_{i}"]), (name, i, "bk", rel_err(obk, z[f"lbk_{i}"]))
        if ofw.shape[0]:
            assert np.array_equal(pfw[:pfw[0] + 2], z[f"ipfw_{i}"]), (name, i)
        if obk.shape[0]:
            assert np.array_equal(pbk[:pbk[0] + 2], z[f"ipbk_{i}"]), (name, i)
        n_checked += 1
    assert n_checked > 0
