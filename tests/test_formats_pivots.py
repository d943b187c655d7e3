"""Pivot files: the DysPivots writer and the Make_LancPivots reader against
files the reference wrote.

* tests/golden/ref_examples/Pivts.dat is the reference's own example, written
  by MakePivots=50 from examples/sp_prop_O16_HF (the manual's tutorial);
* the sc_* fixtures hold the pivots files the compiled reference wrote from
  synthetic HF propagators (tests/golden/make_golden_sc.py).
"""
import os

import numpy as np
import pytest

from bcdor_amd import formats, synth
from helpers import GOLDEN

EX = os.path.join(GOLDEN, "ref_examples")


def o16():
    ms = formats.read_model_space(os.path.join(EX, "input_msp_N7"))
    g = formats.read_propagator(os.path.join(EX, "sp_prop_O16_HF"), ms)
    return ms, g


def test_dys_pivots_reproduces_reference_example(tmp_path):
    ms, g = o16()
    out = tmp_path / "Pivts.dat"
    formats.dys_pivots(str(out), ms, g, 50, 50)
    assert out.read_bytes() == open(os.path.join(EX, "Pivts.dat"), "rb").read()


def test_read_reference_example():
    ms, g = o16()
    piv = formats.read_pivots(os.path.join(EX, "Pivts.dat"), ms)
    assert sorted(piv) == list(range(ms.nsubsh))
    for i, (vec, ifw, ibk) in piv.items():
        assert vec.shape == (ms.no[i], ms.no[i])
        assert np.all(ifw == 50) and np.all(ibk == 50)
        # quasiholes first, then quasiparticles by ascending energy: the HF
        # orbitals of the subshell, to the printed precision
        ep, _, fqp, eh, _, fqh = g.subshell(i)
        order = [fqh[j, :ms.no[i]] for j in np.argsort(eh, kind="stable")] + \
                [fqp[j, :ms.no[i]] for j in np.argsort(ep, kind="stable")]
        assert np.array_equal(vec, np.array(order))


@pytest.mark.parametrize("name", ["c1_lanc10", "c2_split50"])
def test_dys_pivots_reproduces_reference_synthetic(tmp_path, name):
    z = np.load(f"{GOLDEN}/sc_{name}.npz", allow_pickle=False)
    ms, _ = synth.make_system(int(z["nmax"]), int(z["seed"]), hw=float(z["hw"]))
    g = formats.Propagator(z["hf_np"], z["hf_nh"], z["hf_ep"], z["hf_zp"], z["hf_fqp"], z["hf_eh"], z["hf_zh"],
                           z["hf_fqh"])
    n = int(z["lanc_n"]) if int(z["lanc_kind"]) == 0 else 1
    out = tmp_path / "Pivts.dat"
    formats.dys_pivots(str(out), ms, g, n, n)
    assert out.read_bytes() == z["pivots_file_ref"].tobytes()


def test_write_read_round_trip(tmp_path):
    ms, g = o16()
    piv = formats.read_pivots(os.path.join(EX, "Pivts.dat"), ms)
    rng = np.random.default_rng(3)
    piv2 = {i: (v + rng.standard_normal(v.shape) * 1e-3, f - 7 * (i % 3), b + (i % 5)) for i, (v, f, b) in piv.items()}
    for full in (True, False):
        p = tmp_path / f"p{int(full)}"
        formats.write_pivots(str(p), ms, piv2, full_precision=full)
        back = formats.read_pivots(str(p), ms)
        for i in piv2:
            v, f, b = piv2[i]
            want = v if full else np.vectorize(lambda x: float("%18.6e" % x))(v)
            assert np.array_equal(back[i][0], want) and np.array_equal(back[i][1], f) and np.array_equal(back[i][2], b)


def test_reader_skips_lines_like_the_reference(tmp_path):
    """A header that announces more pivots than lines follow: the reference
    reads that many lines and keeps only those starting with two integers,
    so it runs into the next subshell's block; the reader does the same."""
    ms, _ = o16()
    text = ("#\n\nshell,npvts:   0   4  # a \n  5  6   1.0 2.0 3.0 4.0\n\nshell,npvts:   1   1  # b \n"
            "  7  8   5.0 6.0 7.0 8.0\n")
    p = tmp_path / "p"
    p.write_text(text)
    piv = formats.read_pivots(str(p), ms)
    assert piv[0][1].tolist() == [5, 7] and piv[0][0].shape == (2, 4)
    assert piv[1][1].tolist() == [7]
