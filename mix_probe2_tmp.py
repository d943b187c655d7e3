import sys, time
sys.path[:0] = ['boccadorata-public_amd', 'tests', '.']
import numpy as np
import bench
from bcdor_amd import capi
specs, b, _ = bench.screened_batch(256, 16)
def p(accel, ms=0.0):
    q = bench.params(0, accel); q.mix_start = ms; return q
b.solve(); b.solve()
its0 = np.array([b.stats(i)['sc_its'] for i in range(256)])
print(f"plain: sc0 {b.timings()['sc0']:.1f} ms; its mean {its0.mean():.2f} max {its0.max()} q90 {np.quantile(its0,0.9)}; unconv {sum(not b.converged(i) for i in range(256))}", flush=True)
for accel, ms in ((5, 1e-3), (5, 1e-4), (3, 1e-3), (8, 1e-2)):
    b.set_params(p(accel, ms)); b.solve(); b.solve()
    inf = [b.start_info(i) for i in range(256)]
    f = np.array([x['first_its'] for x in inf]); fb = [i for i in range(256) if inf[i]['fell_back']]
    print(f"accel {accel} start {ms:g}: sc0 {b.timings()['sc0']:.1f} ms; first its mean {f.mean():.2f} max {f.max()} q90 {np.quantile(f,0.9)}; fallbacks {fb} (plain its {[int(its0[i]) for i in fb]}, mixed {[int(f[i]) for i in fb]}, resid {[inf[i]['residual'] for i in fb]})", flush=True)
    b.set_fallback_tol(0.0); b.solve()
    print(f"   no check: sc0 {b.timings()['sc0']:.1f} ms", flush=True)
    b.set_fallback_tol(1e-6)
