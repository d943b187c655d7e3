"""GPU parity of the whole SCGF solve (bcd_batch_*) against the oracle's
restatement of Run_itrs on the same inputs.

Bar: quasiparticle/quasihole energies 1e-10 relative (floor 1 MeV),
spectroscopic factors 1e-10 absolute, the same fragment counts per subshell,
the same number of sweeps, and instances of a batch identical to the same
instances solved alone.

Fragments with Z below helpers.ZSIG are compared through the Z-weighted
spectral moments only: when a Lanczos chain runs past Krylov breakdown their
energies are rounding-determined in the reference itself (see helpers.py).
"""
import numpy as np
import pytest

import pyoracle as po
from bcdor_amd import capi, synth
from helpers import ZSIG_BY_NMAX, spectral_parity

pytestmark = pytest.mark.gpu


def oracle_solve(ms, vpp, lanczos, conv_kind, hf_conv, sc_conv, sc_itr_max=-1, lanc_n=50):
    S = po.System(ms, vpp)
    ghf, st_hf = po.run(S, po.Prop.bare(S), "hf", itr_max=-1, conv_kind=conv_kind, conv=hf_conv)
    g, st = po.run(S, ghf, "sc0", itr_max=sc_itr_max, lanczos=lanczos, lanc_kind=1, lanc_n=lanc_n,
                   conv_kind=conv_kind, conv=sc_conv)
    return g.to_propagator(ms.nsubsh), st_hf, st


def check_prop(g, ref, tol=1e-10, zsig=1e-6):
    assert np.array_equal(g.np_, ref.np_) and np.array_equal(g.nh, ref.nh), "fragment counts differ"
    for i in range(len(g.np_)):
        ep, zp, _, eh, zh, _ = g.subshell(i)
        rp, rz, _, rh, rzh, _ = ref.subshell(i)
        # quasiparticles and quasiholes separately, so the split at E_F is checked too
        for a, b in ((np.stack([ep, zp], 1), np.stack([rp, rz], 1)), (np.stack([eh, zh], 1), np.stack([rh, rzh], 1))):
            e_err, z_err, m_err = spectral_parity(a, b, tol, zsig=zsig)
            assert e_err < tol and z_err < tol and m_err < tol, (i, e_err, z_err, m_err)


def test_config1_full_sigma_sc0():
    """Plain FP64 config 1: HF start, full (unreduced) Sigma(2), sc0 with the
    reference's ItrMax=20 and sum|de| < 1e-7 criterion."""
    ms, vpp = synth.make_system(3, 0)
    par = capi.solve_params(mode=2, hf_itr_max=-1, sc_itr_max=20, lanczos=False, conv_kind=0, hf_conv=1e-7,
                            sc_conv=1e-7)
    b = capi.Batch([(ms, vpp)], par)
    b.solve()
    st = b.stats(0)
    print("config1 full", st, b.timings(), flush=True)
    ref, st_hf, st_sc = oracle_solve(ms, vpp, False, 0, 1e-7, 1e-7, sc_itr_max=20)
    assert st["hf_its"] == st_hf.iterations and st["sc_its"] == st_sc.iterations
    check_prop(b.result(0), ref)
    assert abs(st["e_kolt"] - st_sc.e_kolt) < 1e-9 * max(1.0, abs(st_sc.e_kolt))


def test_nmax5_lanczos_sc0():
    """Lanczos sc0 at Nmax=5, where the reference is self-consistent under
    pole reordering (tests/test_oracle_dyson.py::test_lanczos_breakdown_sensitivity)."""
    ms, vpp = synth.make_system(5, 0)
    par = capi.solve_params(mode=2, lanczos=True, lanc_kind=1, lanc_n=50, conv_kind=1, hf_conv=1e-8, sc_conv=1e-8)
    b = capi.Batch([(ms, vpp)], par)
    b.solve()
    st = b.stats(0)
    print("nmax5 lanczos", st, b.timings(), flush=True)
    ref, st_hf, st_sc = oracle_solve(ms, vpp, True, 1, 1e-8, 1e-8)
    assert st["hf_its"] == st_hf.iterations and st["sc_its"] == st_sc.iterations
    check_prop(b.result(0), ref, zsig=ZSIG_BY_NMAX[5])


def test_config2_lanczos_sc0():
    """Plain FP64 config 2: Nmax=7, 50 Krylov vectors per side, sc0 to
    max|de| < 1e-8."""
    ms, vpp = synth.make_system(7, 0)
    par = capi.solve_params(mode=2, lanczos=True, lanc_kind=1, lanc_n=50, conv_kind=1, hf_conv=1e-8, sc_conv=1e-8)
    b = capi.Batch([(ms, vpp)], par)
    b.solve()
    st = b.stats(0)
    w = b.work()
    print("config2", st, b.timings(), w["sigma2_poles"], w["lanczos_poles"], flush=True)
    # The Lanczos ran on J-merged pole groups, so this also holds the merge
    # to the bar.
    assert 0 < w["lanczos_poles"] < w["sigma2_poles"]
    ref, st_hf, st_sc = oracle_solve(ms, vpp, True, 1, 1e-8, 1e-8)
    assert st["hf_its"] == st_hf.iterations and st["sc_its"] == st_sc.iterations
    check_prop(b.result(0), ref, zsig=ZSIG_BY_NMAX[7])


def test_batch_equals_single():
    """Instances of different sizes in one batch give exactly what they give alone."""
    systems = [synth.make_system(nm, s) for nm, s in ((3, 1), (5, 2), (3, 3))]
    par = capi.solve_params(mode=2, lanczos=True, lanc_kind=1, lanc_n=30, conv_kind=1, hf_conv=1e-8, sc_conv=1e-8)
    b = capi.Batch(systems, par)
    b.solve()
    for i, sysm in enumerate(systems):
        one = capi.Batch([sysm], par)
        one.solve()
        ga, gb = b.result(i), one.result(0)
        assert np.array_equal(ga.ep, gb.ep) and np.array_equal(ga.eh, gb.eh), i
        assert np.array_equal(ga.fqp, gb.fqp) and np.array_equal(ga.fqh, gb.fqh), i
        assert b.stats(i) == one.stats(0)


@pytest.mark.parametrize("nmax,lanczos,lanc_n", [(3, False, 50), (5, True, 50), (7, True, 50), (7, True, 70)])
def test_arrowhead_equals_dense(nmax, lanczos, lanc_n):
    """The arrowhead Dyson solver (default) against the dense Householder + QL
    solver on the whole solve: same sweep counts and fragments to 1e-10.  With
    the full Sigma every fragment is compared; with Lanczos the HF stage's
    last-bit differences reach the Lanczos input and move the small-Z
    fragments as rounding does in the reference, so the Z cut applies.  With
    70 vectors per side the reduced sides exceed 64 rows: the one-wave pole
    blocks then hold two columns per lane.  (Not 100: several 2h1p sides at
    Nmax 7 hold fewer distinct pole energies than that, e.g. 61 among 96
    poles with one coupling column, so the Krylov chain runs past breakdown
    and the solutions it feeds are rounding-determined in either solver.)"""
    ms, vpp = synth.make_system(nmax, 0)
    out = []
    for solver in (0, 1):
        par = capi.solve_params(mode=2, lanczos=lanczos, lanc_kind=1, lanc_n=lanc_n, conv_kind=1, hf_conv=1e-8,
                                sc_conv=1e-8, sc_itr_max=40, dyson_solver=solver)
        b = capi.Batch([(ms, vpp)], par)
        b.solve()
        print(f"nmax {nmax} solver {solver}", b.stats(0), b.timings(), flush=True)
        out.append((b.result(0), b.stats(0)))
    (ga, sa), (gb, sb) = out
    assert sa["hf_its"] == sb["hf_its"] and sa["sc_its"] == sb["sc_its"]
    check_prop(ga, gb, tol=1e-10, zsig=ZSIG_BY_NMAX[nmax] if lanczos else 0.0)
    assert abs(sa["e_kolt"] - sb["e_kolt"]) < 1e-9 * max(1.0, abs(sb["e_kolt"]))


@pytest.mark.parametrize("nmax", [5, 7])
def test_mixed_precision_error_vs_fp64(nmax):
    """Mixed mode (FP32 operands and products in the Sigma(2) coupling
    contraction, FP64 sums; Lanczos, eigensolves and the sc loop in FP64)
    against the FP64 GPU solve of the same batch.  The FP32 rounding of V and
    of the fragment amplitudes is amplified by the 50-vector Lanczos
    truncation and the sc0 feedback: measured worst errors are 3.9e-6 (Nmax 5)
    and 8.4e-6 (Nmax 7) relative in energy, 3.2e-6 absolute in Z, 1.5e-7 in the
    Z-weighted moments.  The bounds below are those with a margin of 2-3x."""
    systems = [synth.make_system(nmax, s) for s in (0, 1, 2)]
    out = {}
    for prec in (0, 1):
        par = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                                conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, precision=prec)
        b = capi.Batch(systems, par)
        b.solve()
        out[prec] = [b.result(i) for i in range(len(systems))]
        conv = [b.converged(i) for i in range(len(systems))]
        b.close()
    zsig = ZSIG_BY_NMAX[nmax]
    we = wz = wm = 0.0
    for g, ref in zip(out[1], out[0]):
        assert np.array_equal(g.np_, ref.np_) and np.array_equal(g.nh, ref.nh)
        for i in range(len(g.np_)):
            ep, zp, _, eh, zh, _ = g.subshell(i)
            rp, rz, _, rh, rzh, _ = ref.subshell(i)
            for a, b_ in ((np.stack([ep, zp], 1), np.stack([rp, rz], 1)),
                          (np.stack([eh, zh], 1), np.stack([rh, rzh], 1))):
                e_err, z_err, m_err = spectral_parity(a, b_, 1e-6, zsig=zsig)
                we, wz, wm = max(we, e_err), max(wz, z_err), max(wm, m_err)
    print(f"mixed vs fp64, nmax {nmax}: energy {we:.2e} Z {wz:.2e} moments {wm:.2e}; converged {conv}", flush=True)
    assert we < 3e-5 and wz < 1e-5 and wm < 1e-6
    assert we > 0.0   # the mixed path really ran


@pytest.mark.parametrize("nmax", [5, 7])
def test_anderson_mixing_same_fixed_point(nmax):
    """Anderson mixing of the density (accel = 5) against the reference's plain
    iteration on the same batch: both stop at max|de| < 1e-8, so they sit on
    the same fixed point up to what that tolerance leaves open.  Solutions
    above the Nmax Z cut agree to 1e-6 relative in energy and 1e-6 in Z, the
    Z-weighted moments to 1e-7; the mixed iteration takes fewer sweeps."""
    systems = [synth.make_system(nmax, s) for s in (0, 1, 2)]
    out, its = {}, {}
    for accel in (0, 5):
        par = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                                conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, accel=accel)
        b = capi.Batch(systems, par)
        b.solve()
        assert all(b.converged(i) for i in range(len(systems))), accel
        out[accel] = [b.result(i) for i in range(len(systems))]
        its[accel] = sum(b.stats(i)["hf_its"] + b.stats(i)["sc_its"] for i in range(len(systems)))
        b.close()
    zsig = ZSIG_BY_NMAX[nmax]
    we = wz = wm = 0.0
    for g, ref in zip(out[5], out[0]):
        assert np.array_equal(g.np_, ref.np_) and np.array_equal(g.nh, ref.nh)
        for i in range(len(g.np_)):
            ep, zp, _, eh, zh, _ = g.subshell(i)
            rp, rz, _, rh, rzh, _ = ref.subshell(i)
            for a, b_ in ((np.stack([ep, zp], 1), np.stack([rp, rz], 1)),
                          (np.stack([eh, zh], 1), np.stack([rh, rzh], 1))):
                e_err, z_err, m_err = spectral_parity(a, b_, 1e-6, zsig=zsig)
                we, wz, wm = max(we, e_err), max(wz, z_err), max(wm, m_err)
    print(f"anderson vs plain, nmax {nmax}: energy {we:.2e} Z {wz:.2e} moments {wm:.2e}; sweeps {its}", flush=True)
    assert we < 1e-6 and wz < 1e-6 and wm < 1e-7
    assert its[5] < its[0]


def test_lanczos_autotune():
    """The tile-shape autotuner times every shape that fits each pole-count
    bucket on the batch's own Lanczos jobs during one solve and keeps the
    fastest, or the shape of the next larger occupied bucket when that one is
    within LANC_MERGE_TOL of the fastest.  Shapes only regroup the block
    sums, so the tuned solve agrees with the default-shape solve to rounding
    (spectral parity at 1e-10), and later solves with the chosen shapes
    repeat it bit for bit."""
    systems = [synth.make_system(5, s) for s in (0, 1, 2)]
    par = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                            conv_kind=1, hf_conv=1e-8, sc_conv=1e-8)
    b = capi.Batch(systems, par)
    b.solve()
    ref = [b.result(i) for i in range(len(systems))]
    default, _ = b.lanczos_shapes()
    b.autotune()
    b.solve()
    shapes, ms = b.lanczos_shapes()
    tuned = [b.result(i) for i in range(len(systems))]
    b.solve()
    again = [b.result(i) for i in range(len(systems))]
    b.close()
    print("default", default.tolist(), "tuned", shapes.tolist(), flush=True)
    timed = np.flatnonzero(ms.max(axis=1) > 0)
    assert timed.size > 0
    for bk in range(capi.LANC_BUCKETS):
        nt, per = capi.LANC_SHAPES[shapes[bk]]
        if bk in timed:
            best = ms[bk][ms[bk] > 0].min()
            larger = timed[timed > bk]
            prev = shapes[larger.min()] if larger.size else -1
            assert ms[bk, shapes[bk]] == best or (
                shapes[bk] == prev and ms[bk, prev] <= np.float32(1.0 + capi.LANC_MERGE_TOL) * best), bk
        else:
            assert shapes[bk] == default[bk]
        assert nt * per >= (capi.LANC_BUCKET_MAX[bk - 1] + 1 if bk else 1)
    zsig = ZSIG_BY_NMAX[5]
    for g, r, a in zip(tuned, ref, again):
        assert np.array_equal(g.ep, a.ep) and np.array_equal(g.fqp, a.fqp) and np.array_equal(g.eh, a.eh)
        for i in range(len(r.np_)):
            ep, zp, _, eh, zh, _ = g.subshell(i)
            rp, rz, _, rh, rzh, _ = r.subshell(i)
            for x, y in ((np.stack([ep, zp], 1), np.stack([rp, rz], 1)), (np.stack([eh, zh], 1), np.stack([rh, rzh], 1))):
                e_err, z_err, m_err = spectral_parity(x, y, 1e-10, zsig=zsig)
                assert e_err < 1e-10 and z_err < 1e-10 and m_err < 1e-10, (e_err, z_err, m_err)


def test_profiled_solve_equals_timed_solve():
    """The timed solve runs each pole-count bucket's Lanczos jobs and then its
    pole blocks on one of four streams; a profiled solve runs the pole blocks
    after every bucket, one kernel group at a time.  Each side's work is the
    same either way, so both give the same propagators bit for bit."""
    systems = [synth.make_system(nm, s) for nm, s in ((7, 0), (5, 1), (5, 2))]
    par = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                            conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, accel=5)
    b = capi.Batch(systems, par)
    b.solve()
    timed = [b.result(i) for i in range(len(systems))]
    b.set_profile(True)
    b.solve()
    prof = [b.result(i) for i in range(len(systems))]
    kt = b.kernel_times()
    b.close()
    assert kt["poles"][1] > 0 and kt["merge"][1] > 0
    for g, r in zip(prof, timed):
        assert np.array_equal(g.np_, r.np_) and np.array_equal(g.nh, r.nh)
        assert np.array_equal(g.ep, r.ep) and np.array_equal(g.eh, r.eh)
        assert np.array_equal(g.fqp, r.fqp) and np.array_equal(g.fqh, r.fqh)
