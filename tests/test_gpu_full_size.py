"""Size-independent properties of the batched workload at its full size: 256
heterogeneous instances (Nmax 5-11, hw U[10,30], sigma U[1,3], a Lanczos
length U[30,80] per instance), solved in one batch as the benchmark solves
them (FP64, Anderson mixing in sc0 with the residual check).  The oracle
cannot run these sizes in a test; these properties hold for every channel of
every instance whatever its size:

- closure: the Dyson eigenvectors of a channel are orthonormal, so the
  quasiparticle amplitudes of all its solutions, quasiparticles and
  quasiholes together, satisfy sum_k x_k x_k^T = 1 (nI x nI); its trace is
  the sum rule sum_k Z_k = nI (DysonMatrix.cpp:436-484 stores every
  eigenpair; Sp_Propagator.cpp:728-734 recomputes Z as |x|^2);
- order: each side is stored in descending energy and every quasiparticle
  lies above every quasihole of its subshell (the split at the Fermi
  energy, DysonMatrix.cpp:442-471);
- fixed point: an instance whose mixed iteration converged reproduces its
  energies through one more plain sweep to within the residual tolerance,
  or was redone by the plain iteration."""
import numpy as np
import pytest

from bcdor_amd import capi, synth

pytestmark = pytest.mark.gpu

CLOSURE_TOL = 1e-10
RESID_TOL = 1e-6


def test_full_size_invariants():
    specs = synth.batched_specs(256, 4)
    print(f"full size: {len(specs)} instances, Nmax {sorted({s.nmax for s in specs})}", flush=True)
    par = capi.solve_params(mode=2, hf_itr_max=100, sc_itr_max=60, lanczos=True, lanc_kind=1, lanc_n=50,
                            conv_kind=1, hf_conv=1e-8, sc_conv=1e-8, accel=5)
    b = capi.Batch.streamed(len(specs), lambda i: specs[i].make(), par, lanc_n=[s.lanc_n for s in specs],
                            threads=16)
    b.set_fallback_tol(RESID_TOL)
    b.solve()
    print(" ", b.timings(), flush=True)
    worst_closure, worst_order, nchan, nconv = 0.0, 0, 0, 0
    for i, s in enumerate(specs):
        ms = b.systems[i][0]
        g = b.result(i)
        for k in range(ms.nsubsh):
            nI = int(ms.no[k])
            ep, zp, xp, eh, zh, xh = g.subshell(k)
            x = np.concatenate([xp[:, :nI], xh[:, :nI]])
            closure = x.T @ x - np.eye(nI)
            worst_closure = max(worst_closure, float(np.max(np.abs(closure))))
            z = np.concatenate([zp, zh])
            assert np.allclose(z, np.sum(x * x, axis=1), rtol=0.0, atol=1e-12), (i, k)
            if np.any(np.diff(ep) > 0.0) or np.any(np.diff(eh) > 0.0):
                worst_order += 1
            if ep.size and eh.size and not ep.min() > eh.max():
                worst_order += 1
            nchan += 1
        if b.converged(i):
            nconv += 1
            info = b.start_info(i)
            assert info["fell_back"] or 0.0 <= info["residual"] <= RESID_TOL, (i, info)
    print(f"  {nchan} channels, {nconv} converged instances; worst closure {worst_closure:.2e}", flush=True)
    assert worst_closure < CLOSURE_TOL
    assert worst_order == 0
    assert nconv > 0
