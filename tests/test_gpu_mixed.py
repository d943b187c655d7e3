"""Mixed-precision Sigma(2) coupling contraction on the matrix cores
(bcd_solve_params.precision 1: FP32 operands, 2: BF16 operands, both with FP32
accumulation; 3: FP32 operands on the vector units with FP64 sums) against the
FP64 solve of the same instance, at Nmax 9, seed 1, hw 20, sigma 2 (the
mixed-precision workload's system), HF + Lanczos (50 vectors per side) + sc0.

These modes are not held to the FP64 bar: cancellation in the contraction
amplifies the operand rounding, so each mode's error is measured and printed
as a table (largest relative energy and absolute Z difference over the
solutions with Z >= 1e-3, the Z-weighted moments, and the Koltun energy).
The bounds asserted here are the orders of magnitude of that rounding:
FP32 operands within 1e-3, BF16 within 1e-1 (BF16 keeps 8 significant
bits), and the FP64 mode itself bit-identical between two solves."""
import numpy as np
import pytest

from bcdor_amd import capi, synth

pytestmark = pytest.mark.gpu


def params(precision):
    return capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                             conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, precision=precision)


def errors(g, ref, zsig=1e-3):
    e = z = m = 0.0
    for i in range(len(ref.np_)):
        for sa, sb in ((0, 1), (3, 4)):
            a = np.stack([g.subshell(i)[sa], g.subshell(i)[sb]], 1)
            b = np.stack([ref.subshell(i)[sa], ref.subshell(i)[sb]], 1)
            for x, y in ((a, b), (b, a)):
                for en, zz in x[x[:, 1] >= zsig]:
                    k = int(np.argmin(np.abs(y[:, 0] - en)))
                    e = max(e, abs(y[k, 0] - en) / max(1.0, abs(en)))
                    z = max(z, abs(y[k, 1] - zz))
            for p in range(4):
                ma, mb = np.sum(a[:, 1] * a[:, 0] ** p), np.sum(b[:, 1] * b[:, 0] ** p)
                m = max(m, abs(ma - mb) / (np.sum(np.abs(b[:, 1] * b[:, 0] ** p)) + 1.0))
    return e, z, m


def test_mixed_precision_error_table():
    print("mixed precision, Nmax 9 seed 1", flush=True)
    sysm = synth.make_system(9, 1, hw=20.0, sigma=2.0)
    b = capi.Batch([sysm], params(0))
    b.set_profile(True)
    b.solve()
    ref, st0 = b.result(0), b.stats(0)
    t0 = b.kernel_times()["sigma2"][0]
    b.solve()
    assert np.array_equal(b.result(0).ep, ref.ep) and np.array_equal(b.result(0).eh, ref.eh)
    print(f"  {'mode':34s} {'energy':>9s} {'Z':>9s} {'moments':>9s} {'E_Koltun':>9s} {'sigma2 ms':>9s}", flush=True)
    print(f"  {'FP64 (reference rows)':34s} {0:9.1e} {0:9.1e} {0:9.1e} {0:9.1e} {t0:9.3f}", flush=True)
    bounds = {1: 1e-3, 2: 1e-1, 3: 1e-3}
    names = {1: "matrix cores, FP32 operands", 2: "matrix cores, BF16 operands", 3: "vector units, FP32 operands"}
    for prec in (1, 2, 3):
        b.set_params(params(prec))
        b.solve()
        assert b.converged(0), prec
        e, z, m = errors(b.result(0), ref)
        ek = abs(b.stats(0)["e_kolt"] - st0["e_kolt"]) / abs(st0["e_kolt"])
        t = b.kernel_times()["sigma2"][0]
        print(f"  {names[prec]:34s} {e:9.1e} {z:9.1e} {m:9.1e} {ek:9.1e} {t:9.3f}", flush=True)
        assert max(e, z, m, ek) < bounds[prec], (prec, e, z, m, ek)
