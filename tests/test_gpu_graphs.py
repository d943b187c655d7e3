"""Sweep chunks replayed from captured HIP graphs (bcd_batch_set_graphs)
against the same sweeps launched directly: bit for bit, on a batch whose
channels fall into several arrowhead launch classes (run on concurrent
streams inside the capture), for the plain and the density-mixed iteration.
Each solve replays its chunk graphs many times with the host's flag reads in
between, and the graphed solve is repeated to check later replays too."""
import numpy as np
import pytest

from bcdor_amd import capi, synth

pytestmark = pytest.mark.gpu


def params(accel):
    return capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                             conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, accel=accel)


@pytest.mark.parametrize("accel", [0, 5])
def test_graph_replay_equals_direct_launches(accel):
    sp = [s for s in synth.batched_specs(64, 4) if s.nmax <= 9][:10]
    print(f"graphs accel={accel}: Nmax {[s.nmax for s in sp]}", flush=True)
    b = capi.Batch.streamed(len(sp), lambda i: sp[i].make(), params(accel), lanc_n=[s.lanc_n for s in sp],
                            threads=8)
    b.set_graphs(False)
    b.solve()
    ref = [(b.result(i), b.stats(i)) for i in range(b.n)]
    b.set_graphs(True)
    for rep in range(2):
        b.solve()
        for i in range(b.n):
            g, st = b.result(i), b.stats(i)
            assert st == ref[i][1], (rep, i, st, ref[i][1])
            assert np.array_equal(g.ep, ref[i][0].ep) and np.array_equal(g.eh, ref[i][0].eh), (rep, i)
            assert np.array_equal(g.fqp, ref[i][0].fqp) and np.array_equal(g.fqh, ref[i][0].fqh), (rep, i)
    print("  sc0 ms with graphs", b.timings()["sc0"], flush=True)


def test_batches_on_concurrent_host_threads():
    """Three batches solved at once from three host threads (each capturing
    its sweep chunks into graphs on its own streams while the others copy
    synchronously) give what each gives alone."""
    import threading
    sp = [s for s in synth.batched_specs(64, 4) if s.nmax <= 7][:12]
    parts = [sp[k::3] for k in range(3)]
    batches = [capi.Batch.streamed(len(p), lambda i, p=p: p[i].make(), params(5), lanc_n=[s.lanc_n for s in p],
                                   threads=4) for p in parts]
    alone = []
    for b in batches:
        b.solve()
        alone.append([b.result(i).ep.copy() for i in range(b.n)])
    errors = []

    def run(b):
        try:
            for _ in range(3):
                b.solve()
                for i in range(b.n):
                    b.result(i)
        except Exception as e:          # reported below, in the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=run, args=(b,)) for b in batches]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for b, ref in zip(batches, alone):
        for i in range(b.n):
            assert np.array_equal(b.result(i).ep, ref[i]), i
