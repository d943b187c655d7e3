"""The learned start and the fallback to the plain iteration (bcd_batch_
set_learned_start, bcd_batch_set_fallback_tol) on instances of the batched
workload's generator (Nmax 5 and 7 here, to keep the test short).

* A network trained on the device on the trajectories of one set of
  instances seeds the sc0 iteration of held-out instances: every instance
  that the plain iteration converges ends within 1e-6 (relative, solutions
  with Z >= 1e-5) of the plain result, seeded or redone by the fallback;
  the sweeps it saves (or costs) against the plain iteration are printed.
* The fallback itself: with a tolerance no residual meets, every instance
  seeded by a network (here an untrained one) is redone plainly from its HF
  propagator and then equals the plain solve bit for bit.
* Density mixing (accel > 0) under the same residual check."""
import numpy as np
import pytest

from bcdor_amd import capi, learn, synth

pytestmark = pytest.mark.gpu


def params(accel=0):
    p = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                          conv_kind=1, hf_conv=1e-8, sc_conv=1e-8)
    p.accel = accel
    return p


def specs(seed, n, nmax_max=7):
    return [s for s in synth.batched_specs(4 * n, seed) if s.nmax <= nmax_max][:n]


def batch(sp, accel=0):
    return capi.Batch.streamed(len(sp), lambda i: sp[i].make(), params(accel), lanc_n=[s.lanc_n for s in sp],
                               threads=8)


def worst_rel(a, b, zsig=1e-5):
    if a.ep.shape != b.ep.shape or a.eh.shape != b.eh.shape:
        return np.inf
    w = 0.0
    for x, y, z in ((a.ep, b.ep, b.zp), (a.eh, b.eh, b.zh)):
        m = z >= zsig
        if m.any():
            w = max(w, float(np.max(np.abs(x[m] - y[m]) / np.maximum(1.0, np.abs(y[m])))))
    return w


@pytest.fixture(scope="module")
def held_out():
    sp = specs(4, 32)
    b = batch(sp)
    b.solve()
    ref = [b.result(i) for i in range(b.n)]
    its = np.array([b.stats(i)["sc_its"] for i in range(b.n)])
    conv = np.array([b.converged(i) for i in range(b.n)])
    print(f"held-out: {conv.sum()} of {b.n} converge plainly, mean sc0 sweeps {its[conv].mean():.2f}", flush=True)
    return b, ref, its, conv


def test_learned_start(held_out):
    b, ref, its, conv = held_out
    tr = batch(specs(5, 96))
    x, y, _ = learn.collect(tr)
    tr.close()
    net, hist = learn.train(x, y, epochs=40)
    print(f"trained on {len(x)} channels: loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    b.set_learned_start(net)
    b.solve()
    info = [b.start_info(i) for i in range(b.n)]
    seeded_its = np.array([b.stats(i)["sc_its"] for i in range(b.n)])
    fb = np.array([s["fell_back"] for s in info])
    assert all(s["seeded"] for s in info)
    worst = 0.0
    for i in np.flatnonzero(conv):
        assert b.converged(i), i
        worst = max(worst, worst_rel(b.result(i), ref[i]))
    ok = conv & ~fb
    print(f"learned start: {fb[conv].sum()} fallbacks among the converging instances; sweeps (incl. the residual "
          f"sweep) {seeded_its[ok].mean():.2f} vs plain {its[ok].mean():.2f} on the {ok.sum()} seeded ones "
          f"(saved {its[ok].mean() - seeded_its[ok].mean():+.2f}); worst difference {worst:.1e}", flush=True)
    assert worst < 1e-6
    b.set_learned_start(None)


def test_fallback_redoes_plain_iteration(held_out):
    b, ref, its, conv = held_out
    b.set_learned_start(capi.Mlp(seed=11))
    b.set_fallback_tol(1e-300)
    b.solve()
    for i in range(b.n):
        s = b.start_info(i)
        assert s["fell_back"] and s["residual"] >= 0.0, (i, s)
        assert b.converged(i) == conv[i]
        g = b.result(i)
        assert np.array_equal(g.ep, ref[i].ep) and np.array_equal(g.eh, ref[i].eh), i
        assert np.array_equal(g.fqp, ref[i].fqp) and np.array_equal(g.fqh, ref[i].fqh), i
        assert b.stats(i)["sc_its"] == s["first_its"] + 1 + its[i], (i, b.stats(i), s, its[i])
    b.set_fallback_tol(1e-6)
    b.set_learned_start(None)


def test_density_mixing_with_residual_check(held_out):
    """Mixing runs in the sc0 stage only (HF, Sigma(2) and the Lanczos
    reduction stay those of the plain solve); every instance the plain
    iteration converges ends within 1e-6 of the plain result."""
    b, ref, its, conv = held_out
    b.set_params(params(accel=5))
    b.solve()
    mixed_its = np.array([b.stats(i)["sc_its"] for i in range(b.n)])
    fb = np.array([b.start_info(i)["fell_back"] for i in range(b.n)])
    worst = max(worst_rel(b.result(i), ref[i]) for i in np.flatnonzero(conv))
    ok = conv & ~fb
    print(f"density mixing: sweeps {mixed_its[ok].mean():.2f} vs plain {its[ok].mean():.2f}, "
          f"{fb[conv].sum()} fallbacks, worst difference {worst:.1e}", flush=True)
    assert all(b.converged(i) for i in np.flatnonzero(conv))
    assert worst < 1e-6
    b.set_params(params())
