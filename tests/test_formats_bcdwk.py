"""bcdwk self-energy files (Wide_self_energy_t::save_SelfEn_bin and
load_SelfEn_bin, Self_energy_ext_Op_d.cpp:43-300).

Pinned against files the compiled reference wrote with its own save_SelfEn_bin
(tests/golden/bcdwk/<case>/, written by tests/golden/make_golden.py from the
same rows the case's .npz holds):
* write_bcdwk on the reference's rows gives the reference's file byte for byte,
  for the unreduced and the Lanczos-reduced Sigma(2), both sides and each side
  alone (i_FwBk = 0, +1, -1), including channels with an empty side;
* selfenergy_filename gives the reference's file names;
* read_bcdwk returns those rows and pivots, and write(read(file)) == file for
  the two-side files;
* the split -Fw/-Bk pair reads back as the combined file;
* the older layout of the reference's examples/bcdwk still reads (read_bcdwk_any
  tells the two apart).
"""
import glob
import os

import numpy as np
import pytest

from bcdor_amd import formats
from helpers import GOLDEN, load_golden

WK = os.path.join(GOLDEN, "bcdwk")
SIDES = (("", 0), ("Fw", 1), ("Bk", -1))
# case -> channels whose files the reference wrote (see make_golden.py)
CASES = {"c1_hf_split50": ((0, 9), ("Unred", "Lanc")), "c1_dressed": ((11, 12), ("Lanc",))}


def _rows(z, key, i, width):
    a = np.asarray(z[f"{key}_{i}"], np.float64)
    return a.reshape(-1, width)


def channel_file(ms, z, i, kind):
    """The SelfEnergyFile of channel i built from the golden case's rows."""
    no = int(ms.no[i])
    if kind == "Unred":
        fw, bk = z[f"fw_{i}"], z[f"bk_{i}"]
        width = fw.shape[-1] if fw.size else bk.shape[-1]
        return formats.selfenergy_for_channel(ms, i, _rows(z, "fw", i, width), _rows(z, "bk", i, width), width - no)
    fw, bk = z[f"lfw_{i}"], z[f"lbk_{i}"]
    width = fw.shape[-1] if fw.size else bk.shape[-1]
    ipf = z[f"ipfw_{i}"] if fw.size else None
    ipb = z[f"ipbk_{i}"] if bk.size else None
    return formats.selfenergy_for_channel(ms, i, _rows(z, "lfw", i, width), _rows(z, "lbk", i, width),
                                          width - no, ipf, ipb)


def ref_cases():
    for case, (chans, kinds) in CASES.items():
        for i in chans:
            for kind in kinds:
                for fb, side in SIDES:
                    yield pytest.param(case, i, kind, fb, side, id=f"{case}-{i}-{kind}{fb}")


@pytest.mark.parametrize("case,i,kind,fb,side", list(ref_cases()))
def test_writer_matches_reference_bytes(tmp_path, case, i, kind, fb, side):
    ms, _vpp, _g, z = load_golden(case)
    ref = formats.selfenergy_filename(os.path.join(WK, case), kind + fb, int(ms.j2[i]), int(ms.ip[i]),
                                      int(ms.ch[i]))
    assert os.path.exists(ref), f"reference file name differs: {ref}"
    sef = channel_file(ms, z, i, kind)
    out = str(tmp_path / "mine.bin")
    formats.write_bcdwk(out, sef, side)
    with open(ref, "rb") as f:
        want = f.read()
    with open(out, "rb") as f:
        got = f.read()
    assert got == want

    back = formats.read_bcdwk(ref)
    assert (back.ish, back.ntot_den, back.ntot_orb) == (i, sef.ntot_den, sef.ntot_orb)
    assert (back.j2, back.ip, back.ch) == (int(ms.j2[i]), int(ms.ip[i]), int(ms.ch[i]))
    if side >= 0:
        assert np.array_equal(back.fw, sef.fw) and np.array_equal(back.i_piv_fw, sef.i_piv_fw)
    else:
        assert back.fw.shape[0] == 0
    if side <= 0:
        assert np.array_equal(back.bk, sef.bk) and np.array_equal(back.i_piv_bk, sef.i_piv_bk)
    else:
        assert back.bk.shape[0] == 0
    if side == 0:
        # a one-side file's Lanczos descriptor still describes the side left out,
        # so only two-side files come back byte for byte
        again = str(tmp_path / "again.bin")
        formats.write_bcdwk(again, back, side)
        with open(again, "rb") as f:
            assert f.read() == want


@pytest.mark.parametrize("case", sorted(CASES))
def test_split_pair_reads_as_combined(case):
    chans, kinds = CASES[case]
    ms, _vpp, _g, _z = load_golden(case)
    for i in chans:
        for kind in kinds:
            name = lambda fb: formats.selfenergy_filename(os.path.join(WK, case), kind + fb, int(ms.j2[i]),
                                                          int(ms.ip[i]), int(ms.ch[i]))
            whole = formats.read_bcdwk(name(""))
            split = formats.read_bcdwk_split(name("Fw"), name("Bk"))
            for a in ("fw", "bk", "i_piv_fw", "i_piv_bk"):
                assert np.array_equal(getattr(whole, a), getattr(split, a)), (i, kind, a)


def test_old_layout_examples():
    """The examples/bcdwk files (older layout: no Lanczos descriptor, a full
    NTOT_DIM pivot array after each side) are refused by the current-layout
    reader and read by read_bcdwk_any."""
    files = sorted(glob.glob(os.path.join(GOLDEN, "ref_examples", "bcdwk", "*.bin")))
    assert files
    ms = formats.read_model_space(os.path.join(GOLDEN, "ref_examples", "input_msp_N7"))
    for path in files:
        with pytest.raises(ValueError):
            formats.read_bcdwk(path)
        s = formats.read_bcdwk_any(path)
        old = formats.read_bcdwk_old(path)
        assert np.array_equal(s.fw, old.fw) and np.array_equal(s.bk, old.bk)
        i = s.ish
        assert os.path.basename(path) == f"SelfEn_J2pi={ms.j2[i]}{'+-'[ms.ip[i]]}_dt={ms.ch[i]}.bin"
        assert s.isp_msp == tuple(range(int(ms.orb0[i]), int(ms.orb0[i]) + int(ms.no[i])))
        assert s.mor_n == tuple(int(ms.orb_n[k]) for k in s.isp_msp)


def test_any_reads_current_layout():
    for path in sorted(glob.glob(os.path.join(WK, "*", "*.bin"))):
        a, b = formats.read_bcdwk_any(path), formats.read_bcdwk(path)
        assert np.array_equal(a.fw, b.fw) and np.array_equal(a.bk, b.bk)


def test_filename_symmetry_labels():
    """Lambda, parity or charge != 0 add the _T=<lam><+|-><charge> tag
    (get_filename, Self_energy_ext_Op_d.cpp:43-52)."""
    assert formats.selfenergy_filename("d", "SelfEn", 3, 1, 0) == os.path.join("d", "SelfEn_Dyn_J2pi=3-_dt=0.bin")
    assert formats.selfenergy_filename("d", "SelfEn", 3, 0, 1, fb="-Fw", lam=2, parity=1, charge=-1) == \
        os.path.join("d", "SelfEn_Dyn_T=2-_-1_J2pi=3+_dt=1-Fw.bin")
