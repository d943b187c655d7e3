"""GPU path against the compiled reference's own runs and the boundary seams
the reference's drivers call.

* The whole solve (bcd_batch_*) against tests/golden/sc_*.npz, the records of
  the reference program itself (Run_itrs with LAPACK, tests/golden/
  make_golden_sc.py): HF=-1, then ExtSE with the full Sigma(2) or with the
  Lanczos-reduced one from the reference's own pivots file.  The GPU derives
  its pivots from its own HF result and rounds them as that file carries
  them (piv_round), as the reference's MakePivots -> LanDysPiv workflow does.
* The HF stage with the kinetic term (Trel=1: U1body_fac = (A-1)/A) against
  the reference's HF run.
* Set_external_EFermi (bcd_batch_set_efermi) against the oracle with the same
  forced Fermi energies.
* A Lanczos length per instance (bcd_batch_create_ex) against instances
  solved alone.
* Reduce_SelfEn_Lncz's in-place heapsort (bcd_reduce_selfen_lncz) bit for bit
  on rows with J-degenerate ties, and Reduce_Mtx_Lncz's own entry point
  (bcd_reduce_mtx_lncz) with its error codes.
* The pivots-file rounding on the device against the C library's round trip.
"""
import numpy as np
import pytest

import pyoracle as po
from bcdor_amd import capi, formats, synth
from helpers import GOLDEN, ZSIG_BY_NMAX, load_golden, prop_records, spectral_parity

pytestmark = pytest.mark.gpu


def fixture(name):
    z = np.load(f"{GOLDEN}/sc_{name}.npz", allow_pickle=False)
    ms, vpp = synth.make_system(int(z["nmax"]), int(z["seed"]), hw=float(z["hw"]))
    return z, ms, vpp


def prop(z, pre):
    return formats.Propagator(z[pre + "_np"], z[pre + "_nh"], z[pre + "_ep"], z[pre + "_zp"], z[pre + "_fqp"],
                              z[pre + "_eh"], z[pre + "_zh"], z[pre + "_fqh"])


def worst(ga, gb, ms, zsig):
    e = zz = m = 0.0
    for i in range(ms.nsubsh):
        assert (ga.np_[i], ga.nh[i]) == (gb.np_[i], gb.nh[i]), i
        for sa, sb in ((0, 1), (3, 4)):
            a = np.stack([ga.subshell(i)[sa], ga.subshell(i)[sb]], 1)
            b = np.stack([gb.subshell(i)[sa], gb.subshell(i)[sb]], 1)
            x, y, w = spectral_parity(a, b, zsig=zsig)
            e, zz, m = max(e, x), max(zz, y), max(m, w)
    return e, zz, m


@pytest.mark.parametrize("name,zsig,tol", [("c1_full", 0.0, 1e-10), ("c1_lanc10", ZSIG_BY_NMAX[3], 1e-9),
                                           ("c2_split50", ZSIG_BY_NMAX[7], 1e-9)])
def test_solve_equals_reference_run(name, zsig, tol):
    z, ms, vpp = fixture(name)
    lanc = int(z["lanc_kind"]) >= 0
    par = capi.solve_params(mode=2, hf_itr_max=-1, sc_itr_max=int(z["sc_itr"]), lanczos=lanc,
                            lanc_kind=max(int(z["lanc_kind"]), 0), lanc_n=max(int(z["lanc_n"]), 1), conv_kind=0,
                            hf_conv=1e-7, sc_conv=1e-7, piv_round=1)
    print(f"reference loop {name}", flush=True)
    b = capi.Batch([(ms, vpp)], par)
    b.solve()
    st = b.stats(0)
    assert st["hf_its"] == len(z["hf_check"]) and st["sc_its"] == len(z["sc_check"]), st
    e, zz, m = worst(b.result(0), prop(z, "sc"), ms, zsig)
    print(f"  {name}: energies {e:.2e}, Z {zz:.2e}, moments {m:.2e} (Z >= {zsig:g})", flush=True)
    assert e < tol and zz < tol and m < tol
    last = z["sc_dyson"][-ms.nsubsh:]
    e_ref = float(np.sum((last[:, 4] + last[:, 5]) / 2.0))
    assert abs(st["e_kolt"] - e_ref) < 1e-9 * abs(e_ref)


def test_hf_with_kinetic_term_equals_reference_run():
    z, ms, vpp = fixture("c1_hf_trel1")
    par = capi.solve_params(mode=0, hf_itr_max=-1, lanczos=False, conv_kind=0, hf_conv=1e-7,
                            t_fac=15.0 / 16.0, hw=float(z["hw"]))
    b = capi.Batch([(ms, vpp)], par)
    b.solve()
    assert b.stats(0)["hf_its"] == len(z["hf_check"])
    e, zz, _ = worst(b.result(0), prop(z, "hf"), ms, 0.0)
    assert e < 1e-12 and zz < 1e-12, (e, zz)


def test_forced_fermi_energy():
    """SetEf on two subshells: the first moves one quasiparticle below E_F
    (a changed occupation), the second only moves the split point; four HF
    sweeps and then sc0 sweeps with the full Sigma(2), against the oracle."""
    ms, vpp = synth.make_system(3, 0)
    S = po.System(ms, vpp)
    ghf, _ = po.run(S, po.Prop.bare(S), "hf", itr_max=-1, conv_kind=0, conv=1e-7)
    g = ghf.to_propagator(ms.nsubsh)
    ep, _, _, eh, _, _ = g.subshell(0)
    forced = {0: float(np.max(ep) + 30.0), 3: -5.0}
    for mode, itr in ((0, 4), (2, 3)):
        par = capi.solve_params(mode=mode, hf_itr_max=4, sc_itr_max=itr, lanczos=False, conv_kind=0,
                                hf_conv=1e-7, sc_conv=1e-7)
        b = capi.Batch([(ms, vpp)], par)
        for ish, ef in forced.items():
            b.set_efermi(0, ish, ef)
            b.set_efermi(0, ish, ef + 100.0)     # later values are ignored, as in the reference
        b.solve()
        o, _ = po.run(S, po.Prop.bare(S), "hf", itr_max=4, conv_kind=0, conv=1e-7, ef_force=forced)
        if mode == 2:
            o, _ = po.run(S, o, "sc0", itr_max=itr, conv_kind=0, conv=1e-7, ef_force=forced)
        ref = o.to_propagator(ms.nsubsh)
        assert ref.np_[0] != g.np_[0] or ref.nh[0] != g.nh[0], "the forced E_F changes the occupation"
        e, zz, m = worst(b.result(0), ref, ms, 0.0)
        assert e < 1e-10 and zz < 1e-10, (mode, e, zz)


def test_lanczos_length_per_instance():
    systems = [synth.make_system(nm, s) for nm, s in ((3, 1), (5, 2), (3, 3))]
    lens = [20, 44, 31]
    par = capi.solve_params(mode=2, lanczos=True, lanc_kind=1, lanc_n=50, conv_kind=1, hf_conv=1e-8, sc_conv=1e-8)
    b = capi.Batch(systems, par, lanc_n=lens)
    b.solve()
    for i, sysm in enumerate(systems):
        one = capi.Batch([sysm], capi.solve_params(mode=2, lanczos=True, lanc_kind=1, lanc_n=lens[i], conv_kind=1,
                                                   hf_conv=1e-8, sc_conv=1e-8))
        one.solve()
        ga, gb = b.result(i), one.result(0)
        assert np.array_equal(ga.ep, gb.ep) and np.array_equal(ga.eh, gb.eh), i
        assert b.stats(i) == one.stats(0)


def test_reduce_selfen_sorts_like_the_reference():
    """Rows with J-degenerate energies: the in-place sort leaves them exactly
    as Sort_u (fw) / Sort_d (bk) do; the reduced rows then match the oracle's
    (which follows the reference bit for bit) on the rows the reference
    reproduces under rounding noise."""
    ms, vpp, g, z = load_golden("c2_hf_split50")
    ctx = capi.Context(ms, vpp)
    n_ties = 0
    for i in z["channels"]:
        i = int(i)
        fw, bk = z[f"fw_{i}"].copy(), z[f"bk_{i}"].copy()
        n_ties += len(fw) - len(np.unique(fw[:, 0]))
        out = ctx.reduce_selfen_lncz(int(ms.no[i]), fw, bk, z[f"piv_vec_{i}"], z[f"piv_fw_{i}"], z[f"piv_bk_{i}"])
        sorted_fw, sorted_bk = out[-2], out[-1]
        assert np.array_equal(sorted_fw, po.sort_rows(z[f"fw_{i}"], False)), i
        assert np.array_equal(sorted_bk, po.sort_rows(z[f"bk_{i}"], True)), i
    assert n_ties > 0


def test_reduce_mtx_lncz_entry_and_codes():
    ms, vpp, g, z = load_golden("c1_hf_split50")
    ctx = capi.Context(ms, vpp)
    i = 0
    n = int(ms.no[i])
    rows = po.sort_rows(z[f"fw_{i}"], False)
    piv = z[f"piv_vec_{i}"]
    itr = z[f"piv_fw_{i}"].astype(np.int32)
    tot = min(int(itr[itr > 0].sum()), rows.shape[0])
    na_out = len(itr) + 1
    L, out, ipv, piv_after = ctx.reduce_mtx_lncz(na_out, n, tot, rows, 2, n, piv.copy(), itr.copy())
    oref = np.zeros(tot * (na_out + n)); iref = np.zeros(na_out + n, np.int32)
    pref = piv.copy().reshape(-1); itr2 = itr.copy()
    Lr = po.lib().orc_reduce_mtx_lncz(na_out, n, tot, oref, iref, 2, n, rows.shape[0], rows.reshape(-1).copy(),
                                      len(itr), pref, itr2)
    assert L == Lr and np.array_equal(ipv[:ipv[0] + 2], iref[:iref[0] + 2])
    assert np.array_equal(piv_after.reshape(-1), pref)            # pivots normalised in place, same bits
    assert np.max(np.abs(out[:4] - oref.reshape(tot, -1)[:4])) < 1e-12 * np.max(np.abs(oref))
    # the reference's pass-through codes: empty input, no iterations requested
    assert ctx.reduce_mtx_lncz(na_out, n, tot, rows[:0], 2, n, piv.copy(), itr.copy())[0] == -1
    assert ctx.reduce_mtx_lncz(na_out, n, 0, rows, 2, n, piv.copy(), itr.copy())[0] == -1
    assert ctx.reduce_mtx_lncz(na_out, n, tot, rows, 2, n, piv.copy(), np.zeros_like(itr))[0] == -2


def test_pivot_round_on_device():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(100000) * 10.0 ** rng.uniform(-15, 2, 100000)
    x = np.concatenate([x, [33 / 256, -0.12890625, 9.9999995, 0.99999995, 0.0, 1.0]])
    got = capi.pivot_round(x)
    want = np.array([float("%18.6e" % v) for v in x])
    big = np.abs(x) >= 1e-16
    assert np.array_equal(got[big], want[big])
