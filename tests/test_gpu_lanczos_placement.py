"""The Lanczos launches of a batch run on three streams: placed by estimate
in the first solve after the tile shapes are tuned, then moved by the
launches' measured end times over the next solves (at most two moves, then
kept).  Every job computes the same numbers on whichever stream it runs, so
repeated solves of a batch whose launches span several pole-count buckets
give the same results bit for bit while the placement changes."""
import numpy as np
import pytest

from bcdor_amd import capi, synth

pytestmark = pytest.mark.gpu


def test_results_do_not_depend_on_the_placement():
    sp = synth.batched_specs(24, 7)
    print(f"placement: {len(sp)} instances, Nmax {sorted(s.nmax for s in sp)}", flush=True)
    par = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                            conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, accel=5)
    b = capi.Batch.streamed(len(sp), lambda i: sp[i].make(), par, lanc_n=[s.lanc_n for s in sp], threads=8)
    b.autotune()
    b.solve()           # tunes the tile shapes, one bucket at a time
    sh, _ = b.lanczos_shapes()
    b.solve()           # placement by estimate
    ref = [(b.result(i), b.stats(i)) for i in range(b.n)]
    for rep in range(4):
        b.solve()       # placement from the previous solve's measured ends
        for i in range(b.n):
            g, st = b.result(i), b.stats(i)
            assert st == ref[i][1], (rep, i, st, ref[i][1])
            assert np.array_equal(g.ep, ref[i][0].ep) and np.array_equal(g.eh, ref[i][0].eh), (rep, i)
            assert np.array_equal(g.fqp, ref[i][0].fqp) and np.array_equal(g.fqh, ref[i][0].fqh), (rep, i)
    assert np.array_equal(b.lanczos_shapes()[0], sh)
    print("  Lanczos stage ms", b.timings()["lanczos"], flush=True)
