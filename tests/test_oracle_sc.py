"""The oracle's Dyson stage, Fermi energy, pivots and iteration loop against
the compiled reference's own run (Run_itrs, Main.cpp:105-479).

Fixtures: tests/golden/sc_*.npz, written by tests/golden/make_golden_sc.py from
oracle/_ref/ref_run -- the whole reference program with LAPACK from MKL,
recording every DysonMatrix call and convergence check at %.17g.  The
reference and the oracle build bit-identical Sigma^inf, Sigma(2) and Lanczos
rows (tests/test_oracle_golden.py); they differ in the eigensolver (MKL's
dsyevd against orc_eigh), so Dyson outputs agree to rounding of the
eigenproblem, which these tests bound at 1e-11 relative on every fragment
(no spectroscopic-factor cut).
"""
import numpy as np
import pytest

import pyoracle as po
from bcdor_amd import formats, synth
from helpers import GOLDEN, prop_records, spectral_parity, v_digest

CASES = ["c1_full", "c1_lanc10", "c1_hf_trel1", "c2_split50"]
SC_CASES = ["c1_full", "c1_lanc10", "c2_split50"]
A_NUC = 16


def load(name):
    z = np.load(f"{GOLDEN}/sc_{name}.npz", allow_pickle=False)
    ms, vpp = synth.make_system(int(z["nmax"]), int(z["seed"]), hw=float(z["hw"]))
    assert v_digest(vpp) == str(z["v_digest"]), "interaction generator drifted from the golden digest"
    return z, ms, vpp


def prop(z, pre):
    return formats.Propagator(z[pre + "_np"], z[pre + "_nh"], z[pre + "_ep"], z[pre + "_zp"], z[pre + "_fqp"],
                              z[pre + "_eh"], z[pre + "_zh"], z[pre + "_fqh"])


def u1_fac(trel):
    """U1body_fac of Set_nucleus_data (Initialize.cpp:65-74)."""
    return {0: 1.0, 1: (A_NUC - 1) / A_NUC, 2: 0.0, 7: 0.0}[trel]


def pivots(z, ms):
    return {i: (z[f"piv_vec_{i}"], z[f"piv_fw_{i}"], z[f"piv_bk_{i}"]) for i in range(ms.nsubsh)}


def worst(ga, gb, ms, zsig=0.0):
    e = zz = m = 0.0
    for i in range(ms.nsubsh):
        a, b, c = spectral_parity(prop_records(ga, i), prop_records(gb, i), zsig=zsig)
        e, zz, m = max(e, a), max(zz, b), max(m, c)
    return e, zz, m


@pytest.fixture(scope="module")
def golden_sys():
    cache = {}

    def get(name):
        if name not in cache:
            z, ms, vpp = load(name)
            cache[name] = (z, ms, po.System(ms, vpp))
        return cache[name]
    return get


@pytest.mark.parametrize("name", CASES)
def test_hf_loop(golden_sys, name):
    """HF=-1 from the bare propagator: same number of sweeps, the same last
    convergence measure, every fragment of the HF propagator to 1e-12."""
    z, ms, S = golden_sys(name)
    g, st = po.run(S, po.Prop.bare(S), "hf", itr_max=-1, conv_kind=0, conv=1e-7,
                   t_fac=u1_fac(int(z["trel"])), hw=float(z["hw"]))
    chk = z["hf_check"]
    assert st.iterations == len(chk)
    assert abs(st.last_diff - chk[-1, 1]) < 1e-12
    e, zz, _ = worst(g.to_propagator(ms.nsubsh), prop(z, "hf"), ms)
    assert e < 1e-12 and zz < 1e-12, (e, zz)
    # Koltun energy of the last sweep: sum over channels of (x1 + x2) / 2
    last = z["hf_dyson"][-ms.nsubsh:]
    e_ref = float(np.sum((last[:, 4] + last[:, 5]) / 2.0))
    assert abs(st.e_kolt - e_ref) < 1e-10 * max(1.0, abs(e_ref))


@pytest.mark.parametrize("name", [c for c in SC_CASES if "lanc" in c or "split" in c])
def test_make_pivots_equal_reference_file(golden_sys, name):
    """DysPivots from the reference's HF propagator, rounded as the reference
    prints and reads them: every component equal to the reference's file."""
    z, ms, S = golden_sys(name)
    ghf = po.Prop.from_propagator(S, prop(z, "hf"))
    for i in range(ms.nsubsh):
        vec = po.make_pivots(S, ghf, i, round_print=True)
        assert np.array_equal(vec, z[f"piv_vec_{i}"]), i
    # unrounded, they differ by the print rounding only
    vec = po.make_pivots(S, ghf, 0, round_print=False)
    assert np.max(np.abs(vec - z["piv_vec_0"])) < 5e-7 * max(1.0, np.max(np.abs(vec)))


def test_print_rounding_of_pivots():
    """The rounding matches the reference's text round trip on values that
    exercise it: ties of the 7th digit, values that carry into the next
    decade, tiny and negative values."""
    z, ms, vpp = load("c1_lanc10")
    S = po.System(ms, vpp)
    g = formats.bare_propagator(ms)
    vals = [0.12890625, -0.12890625, 9.9999995, 0.99999995e-3, 1.2345675e-20, -7.77777777e-9, 0.0, 1.0]
    for v in vals:
        g2 = formats.Propagator(g.np_, g.nh, g.ep, g.zp, g.fqp.copy(), g.eh, g.zh, g.fqh.copy())
        g2.fqh[0, 0] = v
        P = po.Prop.from_propagator(S, g2)
        got = po.make_pivots(S, P, 0, z_cut=0.0, round_print=True)[0, 0]
        assert got == float(f"{v:18.6e}"), (v, got)


@pytest.mark.parametrize("name", SC_CASES)
def test_sc0_loop(golden_sys, name):
    """ExtSE from the reference's HF propagator with its Sigma(2) (full or
    Lanczos-reduced from the reference's pivots file): the same number of
    sweeps, the same last measure, and every fragment of the final
    propagator to 1e-11 (energies relative, Z absolute), no Z cut."""
    z, ms, S = golden_sys(name)
    ghf = po.Prop.from_propagator(S, prop(z, "hf"))
    lanc = int(z["lanc_kind"]) >= 0
    g, st = po.run(S, ghf, "sc0", itr_max=int(z["sc_itr"]), lanczos=lanc, conv_kind=0, conv=1e-7,
                   pivots=pivots(z, ms) if lanc else None)
    chk = z["sc_check"]
    assert st.iterations == len(chk)
    assert abs(st.last_diff - chk[-1, 1]) < 1e-9 * max(1.0, chk[-1, 1]) + 1e-10
    e, zz, m = worst(g.to_propagator(ms.nsubsh), prop(z, "sc"), ms)
    assert e < 1e-11 and zz < 1e-11 and m < 1e-13, (e, zz, m)
    last = z["sc_dyson"][-ms.nsubsh:]
    e_ref = float(np.sum((last[:, 4] + last[:, 5]) / 2.0))
    assert abs(st.e_kolt - e_ref) < 1e-10 * max(1.0, abs(e_ref))


def channel_sigma(S, ms, ghf, z, i, lanc):
    fw, bk = po.build_sigma2(S, i, ghf)
    if not lanc:
        return dict(fw=fw, bk=bk, ipf=None, ipb=None, npf=0, npb=0, den=2)
    vec, ifw, ibk = z[f"piv_vec_{i}"], z[f"piv_fw_{i}"], z[f"piv_bk_{i}"]
    ofw, pfw, obk, pbk, den, _, _ = po.reduce_selfen_lncz(int(ms.no[i]), fw, bk, vec, ifw, ibk)
    return dict(fw=ofw, bk=obk, ipf=pfw, ipb=pbk, npf=int(pfw[0]) if ofw.shape[0] else 0,
                npb=int(pbk[0]) if obk.shape[0] else 0, den=den)


@pytest.mark.parametrize("name", SC_CASES)
def test_dyson_and_fermi_energy(golden_sys, name):
    """The last sc0 sweep channel by channel: from the reference's input
    propagator of that sweep, Seek_EFermi equals the reference's value bit
    for bit, and orc_dyson (assembly, eigensolve, split at E_F, sort,
    Calc_En) reproduces the fragments the reference appended and its Calc_En
    sums to 1e-11."""
    z, ms, S = golden_sys(name)
    lanc = int(z["lanc_kind"]) >= 0
    ghf = po.Prop.from_propagator(S, prop(z, "hf"))
    gin = po.Prop.from_propagator(S, prop(z, "scprev"))
    gout = prop(z, "sc")
    last = z["sc_dyson"][-ms.nsubsh:]
    for i in range(ms.nsubsh):
        rec = last[i]
        assert int(rec[1]) == i
        ef = po.seek_efermi(S, gin, i)
        assert ef == rec[2], (i, ef, rec[2])
        sg = channel_sigma(S, ms, ghf, z, i, lanc)
        h = po.hf_pot(S, i, gin)
        n = int(ms.no[i])
        u1 = np.diag(ms.orb_e[ms.orb0[i]:ms.orb0[i] + n])
        sq, sh, k3 = po.dyson(n, int(ms.j2[i]), h, u1, sg["fw"], sg["ipf"], sg["npf"], sg["bk"], sg["ipb"], sg["npb"],
                              ef, den_fw=sg["den"], den_bk=sg["den"])
        assert (sq.shape[0], sh.shape[0]) == (int(rec[7]), int(rec[8])), i
        ref = prop_records(gout, i)
        ours = np.concatenate([sq[:, :2], sh[:, :2]])
        e, zz, _ = spectral_parity(ours, ref, zsig=0.0)
        assert e < 1e-11 and zz < 1e-11, (i, e, zz)
        scale = np.maximum(1.0, np.abs(rec[4:7]))
        assert np.all(np.abs(k3 - rec[4:7]) / scale < 1e-11), (i, k3, rec[4:7])


@pytest.mark.parametrize("name", SC_CASES)
def test_sc0_fixed_point(golden_sys, name):
    """Invariant of the converged sc0 solution (reference's and oracle's):
    one more Dyson sweep with h = U1 + Sigma^inf[g] and the fixed Sigma(2)
    gives back g to the convergence tolerance (summed |de| of the sweep
    below the reference's own stopping value scale)."""
    z, ms, S = golden_sys(name)
    lanc = int(z["lanc_kind"]) >= 0
    ghf = po.Prop.from_propagator(S, prop(z, "hf"))
    g_ref = prop(z, "sc")
    o, _ = po.run(S, ghf, "sc0", itr_max=int(z["sc_itr"]), lanczos=lanc, conv_kind=0, conv=1e-7,
                  pivots=pivots(z, ms) if lanc else None)
    for g in (g_ref, o.to_propagator(ms.nsubsh)):
        G = po.Prop.from_propagator(S, g)
        tot = 0.0
        for i in range(ms.nsubsh):
            sg = channel_sigma(S, ms, ghf, z, i, lanc)
            n = int(ms.no[i])
            u1 = np.diag(ms.orb_e[ms.orb0[i]:ms.orb0[i] + n])
            sq, sh, _ = po.dyson(n, int(ms.j2[i]), po.hf_pot(S, i, G), u1, sg["fw"], sg["ipf"], sg["npf"], sg["bk"],
                                 sg["ipb"], sg["npb"], po.seek_efermi(S, G, i), den_fw=sg["den"], den_bk=sg["den"])
            ep, _, _, eh, _, _ = g.subshell(i)
            assert (sq.shape[0], sh.shape[0]) == (len(ep), len(eh)), i
            tot += np.sum(np.abs(sq[:, 0] - ep)) + np.sum(np.abs(sh[:, 0] - eh))
        # the loop stopped once the summed change fell below 1e-7; the next
        # sweep changes less (the map contracts)
        assert tot < 1e-7, (name, tot)
