"""Benchmark: converged SCGF solves per second on one MI355X.

Default workload (--workload batched): 256 independent synthetic closed-shell
systems in one batch, hw uniform in [10, 30] MeV, interaction scale sigma
uniform in [1, 3], Nmax uniform over {5, 7, 9, 11} (22 to 46 channels, up to
~0.6 GB of interaction per instance) and a Lanczos length uniform in [30, 80]
Krylov vectors per side, one seed per instance (bcdor_amd.synth.batched_specs).
One solve = HF from the bare propagator, Sigma(2) from the HF propagator,
Lanczos reduction with full FP64 reorthogonalisation, then sc0 sweeps of
Sigma^inf + Dyson until max|de| < 1e-8 (HF to the same criterion).  A step is
one bcd_batch_solve over the whole batch, already resident in HBM.

The metric is batched (and approximate where that stays within 1e-6) against
plain FP64 on the GPU.  The batched modes are timed on the same instances
(bcd_batch_set_params switches between them): FP64 with the reference's
plain iteration; FP64 with Anderson mixing of the density in the sc0 stage
(fewer sweeps); FP64 with the learned start (a network trained on the
device, before the timing, on the trajectories of 256 other instances,
seeding the second sc0 sweep); and mixed precision in the Sigma(2) coupling
contraction.  Mixed and seeded modes check every result with one plain sweep
and redo an instance plainly when its residual exceeds 1e-6.
value = converged solves/s of the fastest mode, counting only solves within
1e-6 of the plain FP64 solve of the same instance.  Beside it: every mode's
rate, error and sweeps, and the plain FP64 rate of one instance per solve.

--workload nmax7 keeps the single-model-space workload (Nmax 7, hw 20,
sigma 2, 50 vectors per side); --workload twobody the ladder and ring
matrices.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "boccadorata-public_amd"))

import numpy as np  # noqa: E402

from bcdor_amd import capi, learn, synth  # noqa: E402

NMAX = 7
LANC_N = 50
CONV = 1e-8
HF_CAP = 100        # ItrMax of the HF stage
SC_CAP = 60         # ItrMax of the sc0 stage
FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E


PARAMS = dict(mode=2, hf_itr_max=HF_CAP, sc_itr_max=SC_CAP, lanczos=True, lanc_kind=1, lanc_n=LANC_N, conv_kind=1,
              hf_conv=CONV, sc_conv=CONV, t_fac=0.0, ext_u1=1, dyson_solver=0)


ZSIG = 1e-6   # Z below which Nmax=7 solutions are rounding-determined in the reference itself
ERR_BOUND = 1e-6
SINGLE_INSTANCES = 8   # instances solved one per batch for the plain single-instance rate


# batched modes timed on the same instances: (precision, accel)
MODES = {"fp64": (0, 0), "fp64_anderson": (0, 5), "mixed_anderson": (1, 5)}


def params(precision=0, accel=0):
    return capi.solve_params(**PARAMS, precision=precision, accel=accel)


def screened_seeds(n):
    """The first n seeds whose solve converges (HF and sc0 within the caps).
    Random interactions of this scale make a few systems oscillate at the Fermi
    surface (the reference algorithm does the same on them); they are
    replaced by the next seeds, as the generator spec prescribes.  Screening
    is one untimed GPU solve of a candidate batch."""
    cand = list(range(int(1.3 * n) + 8))
    systems = [synth.make_system(NMAX, s) for s in cand]
    b = capi.Batch(systems, params())
    b.solve()
    good = [k for k in range(len(cand)) if b.converged(k)]
    b.close()
    if len(good) < n:
        raise SystemExit(f"only {len(good)} of {len(cand)} candidate seeds converge")
    good = good[:n]
    # candidates skipped before the n-th converging one
    return [cand[k] for k in good], [systems[k] for k in good], good[-1] + 1 - n


def timed(batch, n, steps, warmup, label, tune=True):
    """warmup untimed solves (with tune, the first one also picks the Lanczos
    tile shapes with the batch autotuner), then steps timed ones; returns
    (seconds, per-instance converged counts)."""
    if warmup > 0 and tune:
        batch.autotune()
    for w in range(warmup):
        batch.solve()
        print(f"# {label} warmup {w} {batch.timings()}", file=sys.stderr, flush=True)
    conv = np.zeros(n, np.int64)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.solve()
        conv += np.array([batch.converged(i) for i in range(n)], np.int64)
    return time.perf_counter() - t0, conv


def spectral_error(g, ref, zsig=ZSIG):
    """Largest error of propagator g against ref, per subshell over both
    spectra (quasiparticles and quasiholes together): every solution with
    Z >= zsig in either one is matched to the nearest solution in energy of
    the other; the error is the relative energy difference (denominator at
    least 1 MeV) and the absolute Z difference of the match.  Also the
    relative error of the Z-weighted moments sum Z e^k, k = 0..3."""
    worst = 0.0
    for i in range(len(ref.np_)):
        ep, zp, _, eh, zh, _ = g.subshell(i)
        rp, rz, _, rh, rzh, _ = ref.subshell(i)
        a = (np.concatenate([ep, eh]), np.concatenate([zp, zh]))
        b = (np.concatenate([rp, rh]), np.concatenate([rz, rzh]))
        for (e1, z1), (e2, z2) in ((a, b), (b, a)):
            sel = z1 >= zsig
            if not np.any(sel):
                continue
            if e2.size == 0:
                return float("inf")
            k = np.argmin(np.abs(e1[sel][:, None] - e2[None, :]), axis=1)
            de = np.abs(e1[sel] - e2[k]) / np.maximum(np.abs(e1[sel]), 1.0)
            worst = max(worst, float(de.max()), float(np.abs(z1[sel] - z2[k]).max()))
        for k in range(4):
            ma, mb = np.sum(a[1] * a[0] ** k), np.sum(b[1] * b[0] ** k)
            worst = max(worst, abs(ma - mb) / (np.sum(np.abs(b[1] * b[0] ** k)) + 1.0))
    return worst


def strength_below(g, zsig):
    """Share of the spectral strength (sum of Z over both sides and every
    subshell) carried by solutions with Z below zsig: what the Z-floor
    comparisons leave out."""
    z = np.concatenate([g.zp, g.zh])
    tot = float(np.sum(z))
    return float(np.sum(z[z < zsig])) / tot if tot > 0.0 else 0.0


def moment_error(g, ref):
    """Largest relative error of the Z-weighted moments sum Z e^k, k = 0..3,
    per subshell over both spectra."""
    worst = 0.0
    for i in range(len(ref.np_)):
        ep, zp, _, eh, zh, _ = g.subshell(i)
        rp, rz, _, rh, rzh, _ = ref.subshell(i)
        a = (np.concatenate([ep, eh]), np.concatenate([zp, zh]))
        b = (np.concatenate([rp, rh]), np.concatenate([rz, rzh]))
        for k in range(4):
            ma, mb = np.sum(a[1] * a[0] ** k), np.sum(b[1] * b[0] ** k)
            worst = max(worst, abs(ma - mb) / (np.sum(np.abs(b[1] * b[0] ** k)) + 1.0))
    return worst


# HBM bytes per launch of the dominant kernel family, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE passes of this workload (calibrated against a 1 GiB
# copy): (kernel family, instances per step) -> (bytes, summary file)
MEASURED_TRAFFIC = {("nmax7", "lanczos", 256): (225.7e9, "profiles/scgf_config2_nmax7_b256_lanczos_traffic.md"),
                    ("batched", "lanczos", 256): (864.2e9, "profiles/scgf_batched256_lanczos_traffic.md")}


def roofline(batch, workload):
    """Per-kernel average launch time (hipEvents around every launch of one
    profiled solve) and the dominant kernel's achieved rate."""
    batch.set_profile(True)
    batch.solve()
    kt = batch.kernel_times()
    work = batch.work()
    batch.set_profile(False)
    totals = {k: v[0] * v[1] for k, v in kt.items()}
    dom = max(totals, key=totals.get)
    avg_ms, launches = kt[dom]
    # algorithmic work of each family over the solve (divided by its launches
    # below):
    #   arrow:   arrowhead Dyson sweeps, FLOPs as counted by the kernel (9 per
    #            pole per secular evaluation, 4 per Loewner factor, 5+2*rows
    #            per pole per eigenvector), FP64
    #   poles:   diagonalisation of the reduced sides, 9 L^3 per side, FP64
    #   tridiag: dense solver's Householder reduction, 8/3 n^3 per eigensolve
    #   ql:      dense solver's implicit QL with vector updates, rest of 9 n^3
    #   lanczos: 8*M*sum_j(2j+5+n+npiv) bytes per side
    #   sigma2:  2*poles*K*n FLOPs
    dense = PARAMS.get("dyson_solver", 0) == 1
    per_launch = {
        "arrow": (work["arrow_flops"], "mfma"),
        "poles": (work["poles_flops"], "mfma"),
        "tridiag": (work["sum_n3"] * 8.0 / 3.0 if dense else 0.0, "mfma"),
        "ql": (work["sum_n3"] * (9.0 - 8.0 / 3.0) if dense else 0.0, "mfma"),
        "lanczos": (work["lanczos_bytes"], "hbm"),
        "sigma2": (work["coupling_flops"], "mfma"),
    }
    entry = None
    if dom in per_launch and launches:
        w, bound = per_launch[dom]
        w /= launches
        if bound == "hbm":
            achieved = w / (avg_ms * 1e-3) / 1e9
            entry = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None}
        else:
            achieved = w / (avg_ms * 1e-3) / 1e12
            entry = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None}
        entry["kernel"] = dom
        entry["avg_launch_ms"] = avg_ms
        entry["launches"] = launches
        meas = MEASURED_TRAFFIC.get((workload, dom, batch.n))
        if meas:
            entry["traffic"] = meas[0]
            entry["traffic_note"] = f"HBM bytes per launch, rocprofv3 FETCH_SIZE/WRITE_SIZE, {meas[1]}"
        # the other kernel families with an algorithmic count, each against
        # its own roofline (their summed time over the solve)
        others = {}
        for k, (w, bound) in per_launch.items():
            if k == dom or w <= 0.0 or k not in totals or totals[k] <= 0.0:
                continue
            rate = w / (totals[k] * 1e-3)
            if bound == "hbm":
                others[k] = {"achieved": rate / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": rate / 1e9 / HBM_PEAK_GBS, "ms": totals[k]}
            else:
                others[k] = {"achieved": rate / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": rate / 1e12 / FP64_PEAK_TFLOPS, "ms": totals[k]}
        entry["other_kernels"] = others
    return entry, {k: {"avg_ms": v[0], "launches": v[1]} for k, v in kt.items()}, work


def cpu_baseline(seeds, seconds=15.0, threads=None):
    """The oracle (FP64 C restatement of the reference algorithm) solving the
    same workload, one instance per host thread; counts solves finished in a
    bounded window."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = threads or min(16, os.cpu_count() or 1)
    done = []
    lock = threading.Lock()
    t_end = time.perf_counter() + seconds

    def worker(k):
        j = k
        while time.perf_counter() < t_end:
            ms, vpp = synth.make_system(NMAX, seeds[j % len(seeds)])
            S = po.System(ms, vpp)
            t0 = time.perf_counter()
            ghf, st_hf = po.run(S, po.Prop.bare(S), "hf", itr_max=HF_CAP, conv_kind=1, conv=CONV)
            g, st = po.run(S, ghf, "sc0", itr_max=SC_CAP, lanczos=True, lanc_kind=1, lanc_n=LANC_N, conv_kind=1,
                           conv=CONV)
            with lock:
                if st_hf.last_diff < CONV and st.last_diff < CONV:
                    done.append(time.perf_counter() - t0)
            j += threads

    t0 = time.perf_counter()
    pool = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    wall = time.perf_counter() - t0
    return {"value": len(done) / wall, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{len(done)} Nmax={NMAX} solves (HF + Sigma2 + Lanczos{LANC_N} + sc0), one per thread, "
                      f"{wall:.1f} s wall; mean {np.mean(done):.2f} s per solve"}


# ----------------------------------------------------------- two-body matrices
T2_NMAX = 7
T2_FRAGS = 4          # quasiparticle and quasihole fragments per subshell


def t2_propagator(ms, seed=1):
    """A dressed propagator: T2_FRAGS quasiparticles (e in [20, 120]) and
    quasiholes (e in [-60, 0]) per subshell with N(0, 0.25) overlaps."""
    from bcdor_amd import formats
    rng = np.random.default_rng(seed)
    nspb = ms.nspb
    ep, eh, fp, fh = [], [], [], []
    for i in range(ms.nsubsh):
        for kind in (0, 1):
            for _ in range(T2_FRAGS):
                x = np.zeros(nspb)
                x[:ms.no[i]] = rng.standard_normal(ms.no[i]) * 0.5
                (fp if kind == 0 else fh).append(x)
                (ep if kind == 0 else eh).append(rng.uniform(20, 120) if kind == 0 else rng.uniform(-60, 0))
    fp, fh = np.array(fp), np.array(fh)
    n = np.full(ms.nsubsh, T2_FRAGS, np.int32)
    return formats.Propagator(n, n.copy(), np.array(ep), (fp ** 2).sum(1), fp, np.array(eh), (fh ** 2).sum(1), fh)


def t2_orbit_terms(ms, g, t):
    """Orbit terms of the reference's fill per matrix (entries x no(a) no(b)
    no(g) no(d)), keyed by (kind, J, pi, dT): the work measure used to carry
    a timed sample of channels over to the whole build."""
    ms_no = np.asarray(ms.no)
    sub_p = np.repeat(np.arange(ms.nsubsh), np.asarray(g.np_))
    sub_h = np.repeat(np.arange(ms.nsubsh), np.asarray(g.nh))
    out = {}
    for k in range(t.n):
        inf = t.info(k)
        b = t.basis(k)
        if inf["kind"] == 0:
            wa = ms_no[sub_p[b["n1_pp"]]] * ms_no[sub_p[b["n2_pp"]]]
            wb = ms_no[sub_h[b["k1_hh"]]] * ms_no[sub_h[b["k2_hh"]]]
        else:
            wa = ms_no[sub_p[b["n_ph"]]] * ms_no[sub_h[b["k_ph"]]]
            wb = ms_no[sub_p[b["n_hp"]]] * ms_no[sub_h[b["k_hp"]]]
        w = float(np.concatenate([wa, wb]).astype(np.float64).sum())
        out[(inf["kind"], inf["J"], inf["pi"], inf["dT"])] = w * w
    return out


def t2_cpu_baseline(ms, vpp, g, t, seconds):
    """CPU baseline of the two-body build: the compiled reference
    (oracle/_ref/ref_harness_2b, gII_b.cpp and PolProp_b.cpp as they are)
    fills the mode-3 matrices channel by channel in CoupledCluster_t's loop
    order for a bounded time, single-threaded (the reference keeps its state
    in file-scope statics).  Its rate in orbit terms per second is carried
    over to the whole build.  Without the harness the oracle's restatement is
    timed the same way (kind "port")."""
    import subprocess
    import tempfile
    from bcdor_amd import formats
    terms = t2_orbit_terms(ms, g, t)
    total = sum(terms.values())
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness_2b")
    done, wall, kind = [], 0.0, "reference"
    if os.path.exists(harness):
        with tempfile.TemporaryDirectory() as td:
            formats.write_model_space(os.path.join(td, "msp"), ms)
            formats.write_vpp_bcd(os.path.join(td, "v.bcd"), ms, vpp)
            formats.write_propagator(os.path.join(td, "prop"), ms, g)
            subprocess.run([harness, os.path.join(td, "msp"), os.path.join(td, "v.bcd"), os.path.join(td, "prop"),
                            os.path.join(td, "times"), "time", str(seconds)], check=True, cwd=td,
                           stdout=subprocess.DEVNULL)
            with open(os.path.join(td, "times")) as f:
                for line in f:
                    kd, J, pi, dT, _n, sec = line.split()
                    done.append(terms[(int(kd), int(J), int(pi), int(dT))])
                    wall += float(sec)
    else:
        kind = "port"
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        S = po.System(ms, vpp)
        P = po.Prop.from_propagator(S, g)
        t_end = time.perf_counter() + seconds
        for k in range(t.n):
            if time.perf_counter() > t_end:
                break
            inf = t.info(k)
            b = t.basis(k)
            t0 = time.perf_counter()
            if inf["kind"] == 0:
                po.ladd_fill(S, P, inf["J"], b, 3, False)
            else:
                po.ph_fill(S, P, inf["J"], inf["dT"], b, 3, False)
            wall += time.perf_counter() - t0
            done.append(terms[(inf["kind"], inf["J"], inf["pi"], inf["dT"])])
    rate = sum(done) / wall if wall > 0 else 0.0
    return {"value": rate / total, "unit": "builds/s", "cores": 1, "kind": kind,
            "sample": f"{len(done)} of {t.n} channel matrices in loop order ({sum(done):.3g} of {total:.3g} "
                      f"orbit terms) in {wall:.1f} s of fills; builds/s = (terms/s) / terms per build"}


def main_twobody(args):
    ms, vpp = synth.make_system(T2_NMAX, 0)
    g = t2_propagator(ms)
    ctx = capi.Context(ms, vpp)
    t0 = time.perf_counter()
    t = capi.TwoBody(ctx, g, ladd_mode=3, ring_mode=3, add_free_poles=False)
    print(f"# setup {time.perf_counter() - t0:.1f} s: {t.n} matrices", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        t.build()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t.build()
    dt = time.perf_counter() - t0
    t.set_profile(True)
    t.build()
    kt = t.kernel_times()
    t.set_profile(False)
    work = t.work()
    print(f"# kernels {json.dumps(kt)}", file=sys.stderr, flush=True)
    dom = max(kt, key=lambda k: kt[k][0])
    avg_ms = kt[dom][0]
    flops = {"t2_gemm1": work["flops_gemm1"], "t2_gemm2": work["flops_gemm2"]}.get(dom)
    rl = None
    if flops:
        achieved = flops / (avg_ms * 1e-3) / 1e12
        rl = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
              "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": dom, "avg_launch_ms": avg_ms,
              "launches": kt[dom][1]}
    cpu = None if args.no_cpu else t2_cpu_baseline(ms, vpp, g, t, args.cpu_seconds)
    line = {
        "metric": "ladder + ring (D)RPA matrix sets built per second (every two-body channel, i_RPAcalc=3)",
        "value": args.steps / dt,
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (Nmax={T2_NMAX} seed 0 interaction; {T2_FRAGS} qp + {T2_FRAGS} qh random fragments "
                f"per subshell)",
        "config": {"workload": f"twobody_ladder_ring_nmax{T2_NMAX}_f{T2_FRAGS}", "matrices": t.n,
                   "gemm_flops_per_build": work["flops_gemm1"] + work["flops_gemm2"],
                   "matrix_bytes_per_build": work["bytes_out"]},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main_nmax7(args):
    """The single-model-space workload (Nmax 7, hw 20, sigma 2, 50 vectors per side)."""
    t0 = time.perf_counter()
    seeds, systems, dropped = screened_seeds(args.batch)
    n = args.batch
    batches = {label: capi.Batch(systems, params(prec, accel)) for label, (prec, accel) in MODES.items()}
    plain = batches["fp64"]
    setup_s = time.perf_counter() - t0
    print(f"# setup {setup_s:.1f} s for {n} instances ({dropped} non-converging candidate seeds dropped)",
          file=sys.stderr, flush=True)
    # every batched mode on the same instances
    runs = {}
    for label, b in batches.items():
        dt, conv = timed(b, n, args.steps, args.warmup, label)
        runs[label] = {"dt": dt, "conv": conv, "stages": b.timings(),
                       "sweeps": float(np.mean([b.stats(i)["hf_its"] + b.stats(i)["sc_its"] for i in range(n)]))}
        print(f"# {label}: timed {args.steps} steps in {dt:.3f} s; last stages {runs[label]['stages']}",
              file=sys.stderr, flush=True)
    ref = [plain.result(i) for i in range(n)]
    # plain FP64 GPU: one instance per solve
    n_single = min(n, SINGLE_INSTANCES)
    t_single, err_single = 0.0, 0.0
    for i in range(n_single):
        b = capi.Batch([systems[i]], params(0, 0))
        b.solve()
        t1 = time.perf_counter()
        b.solve()
        t_single += time.perf_counter() - t1
        if not b.converged(0):
            raise SystemExit(f"single-instance solve of seed {seeds[i]} did not converge")
        err_single = max(err_single, spectral_error(b.result(0), ref[i]))
        b.close()
    single_rate = n_single / t_single
    print(f"# single-instance fp64: {n_single} solves in {t_single:.3f} s; max error vs batched {err_single:.2e}",
          file=sys.stderr, flush=True)
    for label, r in runs.items():
        errs = np.zeros(n) if label == "fp64" else \
            np.array([spectral_error(batches[label].result(i), ref[i]) for i in range(n)])
        r["ok"] = errs <= ERR_BOUND
        r["rate"] = float(np.sum(r["conv"] * r["ok"])) / r["dt"]
        r["max_err"] = float(errs.max())
        print(f"# {label}: {r['rate']:.1f} solves/s within {ERR_BOUND}; max error vs fp64 {r['max_err']:.2e}; "
              f"mean sweeps {r['sweeps']:.1f}", file=sys.stderr, flush=True)
    head = max(runs, key=lambda k: runs[k]["rate"])
    batch = batches[head]
    rl, kernels, work = roofline(batch, "nmax7")
    print(f"# kernels {json.dumps(kernels)}", file=sys.stderr, flush=True)
    print(f"# arrowhead phases (block-ms) {json.dumps(work['arrow_phase_ms'])}", file=sys.stderr, flush=True)
    its = [batch.stats(i) for i in range(n)]
    sh, tms = batch.lanczos_shapes()
    shapes = {str(capi.LANC_BUCKET_MAX[k]): list(capi.LANC_SHAPES[sh[k]]) for k in range(capi.LANC_BUCKETS)
              if tms[k].max() > 0}
    cpu = None if args.no_cpu else cpu_baseline(seeds, args.cpu_seconds)
    value = runs[head]["rate"]
    line = {
        "metric": "converged SCGF solves/sec: mixed+batched vs plain FP64 GPU at <=1e-6 rel. err "
                  "(Nmax=7, Lanczos 50/side, sc0 max|de|<1e-8)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * runs[head]["dt"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if MODES[head][0] == 0 else "f64 (f32 Sigma(2) coupling operands)",
        "data": "synthetic (seeded generator: HO j-scheme Nmax=7, hw=20, sigma=2, spe noise 0.5)",
        "config": {"workload": "config2_nmax7_lanczos50_sc0_batched", "instances_per_step": n,
                   "nmax": NMAX, "lanczos_vectors_per_side": LANC_N, "conv_max_de": CONV,
                   "itr_max_hf": HF_CAP, "itr_max_sc0": SC_CAP, "seeds_dropped_in_screening": dropped,
                   "headline_mode": head,
                   "headline_rule": "the fastest batched mode, counting only converged solves within 1e-6 of the "
                                    "FP64 solve of the same instance (plain iteration); accelerated and seeded "
                                    "modes check every result's residual with one plain sweep and redo the "
                                    "instance plainly above 1e-6",
                   "modes": {k: {"precision": "FP32 Sigma(2) coupling operands" if MODES[k][0] else "FP64",
                                 "density_mixing": f"Anderson, {MODES[k][1]} iterates" if MODES[k][1] else "none",
                                 "solves_per_s": r["rate"], "ms_per_step": 1e3 * r["dt"] / args.steps,
                                 "max_error_vs_fp64": r["max_err"], "instances_within_bound": int(r["ok"].sum()),
                                 "mean_sweeps": r["sweeps"]} for k, r in runs.items()},
                   "learned_start": None if net is None else dict(
                       learn_info, mean_sc0_sweeps_saved_vs_plain=float(
                           np.mean(runs["fp64"]["sc_sweeps"] - runs["fp64_learned"]["sc_sweeps"])),
                       note="sweeps include the residual-check sweep; saved < 0 means the seeded "
                            "iteration took more sweeps than the plain one"),
                   "plain_fp64_single_instance_solves_per_s": single_rate,
                   "plain_fp64_single_instance_sample": n_single,
                   "single_vs_batched_max_error": err_single,
                   "batched_over_plain_single": value / single_rate,
                   "error_measure": "per subshell, every Z>=1e-6 solution matched to the nearest in energy of the "
                                    "other solve: max rel. energy and abs. Z difference, and rel. error of the "
                                    "Z-weighted moments; against the FP64 plain-iteration GPU solve of the same "
                                    "instance",
                   "lanczos_tile_shapes": shapes,
                   "lanczos_tile_shapes_note": "autotuned per pole-count bucket (max M: [threads, elements per "
                                               "thread]) during warmup",
                   "mean_hf_sweeps": float(np.mean([s["hf_its"] for s in its])),
                   "mean_sc0_sweeps": float(np.mean([s["sc_its"] for s in its]))},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# ----------------------------------------------------- heterogeneous batch
BATCH_SEED = 4                # batched_specs(n, BATCH_SEED)
ZSIG_BY_NMAX = {5: 1e-5, 7: 1e-6, 9: 1e-5, 11: 1e-5}   # rounding-determined Z below these (Nmax 9, 11: 1e-5)
SINGLE_PER_NMAX = 2           # instances per Nmax solved one per batch for the plain single-instance rate


def build_batched(specs, par, threads):
    return capi.Batch.streamed(len(specs), lambda i: specs[i].make(), par, lanc_n=[s.lanc_n for s in specs],
                               threads=threads)


def screened_batch(n, threads):
    """The batch of the first n instance specs, every instance converging
    within the caps.  Random interactions of these scales can blur the shell
    gap so that HF or sc0 oscillates at the Fermi surface (the reference
    algorithm does the same on them); such an instance gets a new interaction
    seed, keeping its hw, sigma, Nmax and Lanczos length, and after 6 such
    rounds also a new hw and sigma drawn from the workload's ranges.  The candidates are
    screened in one untimed solve of the whole batch, replacements in small
    batches of their own, then the final batch is built."""
    import dataclasses
    specs = synth.batched_specs(n, BATCH_SEED)
    t0 = time.perf_counter()
    b = build_batched(specs, params(), threads)
    b.solve()
    bad = [i for i in range(n) if not b.converged(i)]
    why = {"hf": sum(1 for i in bad if b.stats(i)["hf_its"] > HF_CAP),
           "dropped": sum(1 for i in bad if b.stats(i)["hf_its"] <= HF_CAP and b.stats(i)["sc_its"] == 0)}
    b.close()
    print(f"# screening {time.perf_counter() - t0:.1f} s: {len(bad)} of {n} not converged ({why})",
          file=sys.stderr, flush=True)
    reseeded, attempt = 0, 0
    while bad and attempt < 24:
        attempt += 1
        # a new interaction seed; after 6 rounds also a new draw of hw and
        # sigma from the workload's ranges (Nmax and Lanczos length kept)
        trial = []
        for i in bad:
            sp = dataclasses.replace(specs[i], seed=specs[i].seed + 1000003 * attempt)
            if attempt > 6:
                r = np.random.default_rng(7919 * specs[i].seed + attempt)
                sp = dataclasses.replace(sp, hw=float(r.uniform(10.0, 30.0)), sigma=float(r.uniform(1.0, 3.0)))
            trial.append(sp)
        t = build_batched(trial, params(), threads)
        t.solve()
        still = []
        for k, i in enumerate(bad):
            if t.converged(k):
                specs[i] = trial[k]
                reseeded += 1
            else:
                still.append(i)
        t.close()
        bad = still
    print(f"# {reseeded} instances reseeded in {attempt} rounds, {len(bad)} left", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = build_batched(specs, params(), threads)
    print(f"# final build {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return specs, b, reseeded


def cpu_baseline_batched(specs, seconds, threads=None):
    """The oracle (FP64 C restatement of the reference algorithm) on the
    batched workload's instances, one per host thread, for a bounded time:
    threads are split evenly over the Nmax values and each works through that
    Nmax's instances in order.  The mean time per solve of each Nmax gives
    the throughput of the whole mix (uniform over Nmax) on these threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = threads or min(16, os.cpu_count() or 1)
    by_nmax = {}
    for s in specs:
        by_nmax.setdefault(s.nmax, []).append(s)
    keys = sorted(by_nmax)
    times = {k: [] for k in keys}
    lock = threading.Lock()
    t_end = time.perf_counter() + seconds

    def worker(k):
        nm = keys[k % len(keys)]
        lst = by_nmax[nm]
        j = k // len(keys)
        while time.perf_counter() < t_end and j < len(lst):
            sp = lst[j]
            ms, vpp = sp.make()
            S = po.System(ms, vpp)
            t0 = time.perf_counter()
            ghf, st_hf = po.run(S, po.Prop.bare(S), "hf", itr_max=HF_CAP, conv_kind=1, conv=CONV)
            g, st = po.run(S, ghf, "sc0", itr_max=SC_CAP, lanczos=True, lanc_kind=1, lanc_n=sp.lanc_n,
                           conv_kind=1, conv=CONV)
            dt = time.perf_counter() - t0
            with lock:
                times[nm].append(dt)
            j += max(1, threads // len(keys))

    t0 = time.perf_counter()
    pool = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    wall = time.perf_counter() - t0
    if any(not v for v in times.values()):
        return None
    frac = {k: sum(1 for s in specs if s.nmax == k) / len(specs) for k in keys}
    mean_t = sum(frac[k] * float(np.mean(times[k])) for k in keys)
    done = sum(len(v) for v in times.values())
    per = ", ".join(f"Nmax {k}: {len(times[k])} in {np.mean(times[k]):.2f} s mean" for k in keys)
    return {"value": threads / mean_t, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{done} solves of the batched workload's instances ({per}) on {threads} threads in "
                      f"{wall:.1f} s wall; value = threads / mean time per solve of the workload's Nmax mix"}


LEARN_TRAIN = 256     # instances of another generator seed whose trajectories train the learned start
LEARN_EPOCHS = 40


def train_learned_start(n, threads):
    """The learned start's network, trained on the device on the sc0
    trajectories of n instances drawn with another generator seed (held
    out from the benchmark's instances)."""
    t0 = time.perf_counter()
    specs = synth.batched_specs(n, BATCH_SEED + 1)
    tb = build_batched(specs, params(), threads)
    x, y, _ = learn.collect(tb)
    tb.close()
    t1 = time.perf_counter()
    net, hist = learn.train(x, y, epochs=LEARN_EPOCHS)
    t2 = time.perf_counter()
    info = {"train_instances": n, "generator_seed": BATCH_SEED + 1, "train_channels": int(len(x)),
            "epochs": LEARN_EPOCHS, "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1]),
            "data_s": t1 - t0, "train_s": t2 - t1}
    print(f"# learned start: {json.dumps(info)}", file=sys.stderr, flush=True)
    return net, info


def main_batched(args):
    n = args.batch
    if args.no_screen:
        specs = synth.batched_specs(n, BATCH_SEED)
        plain, reseeded = build_batched(specs, params(), args.threads), 0
    else:
        specs, plain, reseeded = screened_batch(n, args.threads)
    by_nmax = {k: sum(1 for s in specs if s.nmax == k) for k in sorted({s.nmax for s in specs})}
    wanted = {"fp64"} | set(args.modes.split(","))
    net, learn_info = (train_learned_start(LEARN_TRAIN, args.threads) if "fp64_learned" in wanted
                       else (None, None))
    # mode: (parameters, learned start)
    modes = {"fp64": (params(0, 0), False), "fp64_anderson": (params(0, 5), False),
             "fp64_learned": (params(0, 0), True), "mixed_anderson": (params(1, 5), False)}
    modes = {k: v for k, v in modes.items() if k in wanted}
    runs, results = {}, {}
    for q, (label, (par, seeded)) in enumerate(modes.items()):
        plain.set_params(par)
        plain.set_learned_start(net if seeded else None)
        # the tile shapes are tuned once, in the first mode's warm-up: every
        # mode then reduces the same rows in the same order, and the modes
        # differ only where they are meant to (coupling precision, sc0 start
        # and mixing)
        dt, conv = timed(plain, n, args.steps, args.warmup, label, tune=q == 0)
        info = [plain.start_info(i) for i in range(n)]
        runs[label] = {"dt": dt, "conv": conv, "stages": plain.timings(),
                       "sweeps": float(np.mean([plain.stats(i)["hf_its"] + plain.stats(i)["sc_its"]
                                                for i in range(n)])),
                       "sc_sweeps": np.array([plain.stats(i)["sc_its"] for i in range(n)]),
                       "fallbacks": int(sum(s["fell_back"] for s in info))}
        results[label] = [plain.result(i) for i in range(n)]
        print(f"# {label}: timed {args.steps} steps in {dt:.3f} s; last stages {runs[label]['stages']}; "
              f"{runs[label]['fallbacks']} fallbacks", file=sys.stderr, flush=True)
    plain.set_learned_start(None)
    modes = {k: v[0] for k, v in modes.items()}
    ref = results["fp64"]
    zs = [ZSIG_BY_NMAX[s.nmax] for s in specs]
    # the rounding floor of the FP64 solve itself: the same solve with other
    # Lanczos tile shapes (other reduction orders).  Where rounding alone moves
    # an instance's solutions by more than the bound (Krylov chains past
    # breakdown), only its Z-weighted moments are held to the bound.
    tuned, _ = plain.lanczos_shapes()
    plain.set_params(modes["fp64"])
    plain.set_lanczos_shapes(capi.LANC_ALT_SHAPES)
    plain.solve()
    floor = np.array([spectral_error(plain.result(i), ref[i], zs[i]) for i in range(n)])
    plain.set_lanczos_shapes(tuned)
    loose = floor > ERR_BOUND
    print(f"# rounding floor of fp64: {int(loose.sum())} of {n} instances above {ERR_BOUND} "
          f"(max {floor.max():.2e})", file=sys.stderr, flush=True)
    below = np.array([strength_below(ref[i], zs[i]) for i in range(n)])
    for label, r in runs.items():
        errs = np.zeros(n) if label == "fp64" else \
            np.array([spectral_error(results[label][i], ref[i], zs[i]) for i in range(n)])
        merr = np.zeros(n) if label == "fp64" else \
            np.array([moment_error(results[label][i], ref[i]) for i in range(n)])
        r["ok"] = (errs <= ERR_BOUND) | (loose & (merr <= ERR_BOUND))
        r["rate"] = float(np.sum(r["conv"] * r["ok"])) / r["dt"]
        r["max_err"] = float(errs.max())
        print(f"# {label}: {r['rate']:.1f} solves/s within {ERR_BOUND}; max error vs fp64 {r['max_err']:.2e}; "
              f"mean sweeps {r['sweeps']:.1f}", file=sys.stderr, flush=True)
    head = max(runs, key=lambda k: runs[k]["rate"])
    plain.set_params(modes[head])
    plain.set_learned_start(net if head == "fp64_learned" else None)
    rl, kernels, work = roofline(plain, "batched")
    print(f"# kernels {json.dumps(kernels)}", file=sys.stderr, flush=True)
    its = [plain.stats(i) for i in range(n)]
    sh, tms = plain.lanczos_shapes()
    shapes = {str(capi.LANC_BUCKET_MAX[k]): list(capi.LANC_SHAPES[sh[k]]) for k in range(capi.LANC_BUCKETS)
              if tms[k].max() > 0}
    plain.close()
    # plain FP64 on the GPU, one instance per solve: a sample of each Nmax
    sample = []
    for k in by_nmax:
        sample += [i for i in range(n) if specs[i].nmax == k and runs["fp64"]["conv"][i]][:SINGLE_PER_NMAX]
    if args.no_single:
        sample = []
    t_single, err_single = {}, 0.0
    for i in sample:
        b = capi.Batch([specs[i].make()], params(0, 0), lanc_n=[specs[i].lanc_n])
        b.solve()
        t1 = time.perf_counter()
        b.solve()
        t_single.setdefault(specs[i].nmax, []).append(time.perf_counter() - t1)
        err_single = max(err_single, spectral_error(b.result(0), ref[i], zs[i]))
        b.close()
    frac = {k: v / n for k, v in by_nmax.items()}
    single_rate = 1.0 / sum(frac[k] * float(np.mean(t_single[k])) for k in frac) if sample else float("nan")
    print(f"# single-instance fp64: {json.dumps({k: float(np.mean(v)) for k, v in t_single.items()})} s per solve "
          f"by Nmax; {single_rate:.2f} solves/s for the mix; max error vs batched {err_single:.2e}",
          file=sys.stderr, flush=True)
    # an Nmax 11 solve takes ~50 s on one core: the window holds at least one
    cpu = None if args.no_cpu else cpu_baseline_batched(specs, max(args.cpu_seconds, 70.0))
    value = runs[head]["rate"]
    line = {
        "metric": "converged SCGF solves/sec: mixed+batched vs plain FP64 GPU at <=1e-6 rel. err",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * runs[head]["dt"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if modes[head].precision == 0 else "f64 (f32 Sigma(2) coupling operands)",
        "data": "synthetic (seeded generator: HO j-scheme, hw U[10,30], sigma U[1,3], spe noise 0.5, "
                "Nmax U{5,7,9,11}, Lanczos U[30,80] per instance)",
        "config": {"workload": f"batched_{n}_heterogeneous_nmax5to11", "instances_per_step": n,
                   "instances_by_nmax": by_nmax, "lanczos_vectors_per_side": "per instance, U[30,80]",
                   "conv_max_de": CONV, "itr_max_hf": HF_CAP, "itr_max_sc0": SC_CAP,
                   "instances_reseeded_in_screening": reseeded, "headline_mode": head,
                   "headline_rule": "the fastest batched mode, counting only converged solves within 1e-6 of the "
                                    "FP64 solve of the same instance (plain iteration); accelerated and seeded "
                                    "modes check every result's residual with one plain sweep and redo the "
                                    "instance plainly above 1e-6",
                   "modes": {k: {"precision": "FP32 Sigma(2) coupling operands" if modes[k].precision else "FP64",
                                 "density_mixing": f"Anderson, {modes[k].accel} iterates" if modes[k].accel
                                 else "none",
                                 "learned_start": k == "fp64_learned",
                                 "fallbacks_to_plain": r["fallbacks"],
                                 "solves_per_s": r["rate"], "ms_per_step": 1e3 * r["dt"] / args.steps,
                                 "max_error_vs_fp64": r["max_err"], "instances_within_bound": int(r["ok"].sum()),
                                 "converged_instances": int(np.sum(r["conv"] == args.steps)),
                                 "mean_sweeps": r["sweeps"],
                                 "stages_ms": r["stages"]} for k, r in runs.items()},
                   "learned_start": None if net is None else dict(
                       learn_info, mean_sc0_sweeps_saved_vs_plain=float(
                           np.mean(runs["fp64"]["sc_sweeps"] - runs["fp64_learned"]["sc_sweeps"])),
                       note="sweeps include the residual-check sweep; saved < 0 means the seeded "
                            "iteration took more sweeps than the plain one"),
                   "plain_fp64_single_instance_solves_per_s": single_rate,
                   "plain_fp64_single_instance_sample": {str(k): len(v) for k, v in t_single.items()},
                   "single_vs_batched_max_error": err_single,
                   "batched_over_plain_single": value / single_rate,
                   "error_measure": "per subshell, every solution with Z above the Nmax's rounding floor "
                                    f"({ZSIG_BY_NMAX}) matched to the nearest in energy of the other solve: max "
                                    "rel. energy and abs. Z difference, and rel. error of the Z-weighted moments; "
                                    "against the FP64 plain-iteration GPU solve of the same instance.  Instances "
                                    "whose FP64 solutions move by more than the bound under another reduction "
                                    "order (Lanczos tile shapes) are held to the bound on the moments only",
                   "fp64_rounding_floor": {"instances_above_bound": int(loose.sum()),
                                           "max": float(floor.max()),
                                           "median": float(np.median(floor))},
                   "strength_below_z_floor": {"note": "share of sum Z of the fp64 solve in solutions below the "
                                                      "Nmax's Z floor (compared only through the moments)",
                                              "max": float(below.max()), "median": float(np.median(below))},
                   "lanczos_tile_shapes": shapes,
                   "mean_hf_sweeps": float(np.mean([s["hf_its"] for s in its])),
                   "mean_sc0_sweeps": float(np.mean([s["sc_its"] for s in its]))},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--threads", type=int, default=16, help="host threads generating the instances")
    ap.add_argument("--modes", default="fp64,fp64_anderson,fp64_learned,mixed_anderson",
                    help="batched: comma-separated modes to time (fp64 is always timed: it is the reference)")
    ap.add_argument("--no-screen", action="store_true",
                    help="batched: keep the generator's instances as drawn (short profiling runs)")
    ap.add_argument("--no-single", action="store_true",
                    help="batched: skip the one-instance-per-solve sample")
    ap.add_argument("--workload", choices=("batched", "nmax7", "twobody"), default="batched",
                    help="batched: the heterogeneous converged-solve benchmark (default); nmax7: one model "
                         "space; twobody: ladder + ring matrices")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; shard instances across GPUs "
                                   "(one process per GPU) to scale"}))
        sys.exit(2)
    if capi.lib().bcd_device_count() < 1:
        raise SystemExit("no GPU visible: the HIP path has no CPU fallback")
    if args.workload == "twobody":
        return main_twobody(args)
    if args.workload == "nmax7":
        return main_nmax7(args)
    return main_batched(args)


if __name__ == "__main__":
    main()
