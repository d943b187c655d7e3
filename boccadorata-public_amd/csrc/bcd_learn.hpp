// The learned start of the sc0 iteration: per-channel inputs and targets of
// the network and the seeding of the sc0 iteration.
//
// The sc0 sweep reads its input propagator only through the quasihole
// density rho (Sigma^inf[rho], Sp_Propagator_a.cpp:39-130; Sigma(2) is fixed
// during the iteration), so the start the network supplies is a density: the
// first sc0 sweep runs plainly from the HF propagator, and the network then
// predicts, from the start of that trajectory, where it converges; the
// second sweep starts from the prediction.  Densities are expressed in the
// channel's HF orbital basis {u_k} (HF solutions sorted by energy,
// occupations n_k):
//   rho = sum_k n_k u_k u_k^T + sum_kk' D_kk' u_k u_k'^T,
// D as its upper triangle (at most kLearnNI orbits).  The network's inputs
// per channel: the HF energies, their distance to the Fermi energy, the
// occupations, the second-order estimate D2 from the Sigma(2) rows,
//   D2_kk' = - sum_{2p1h poles p} a_pk a_pk' / ((E_p - e_k)(E_p - e_k'))   (k, k' holes)
//   D2_kk' = + sum_{2h1p poles p} a_pk a_pk' / ((e_k - E_p)(e_k' - E_p))   (k, k' particles)
// with a_pk = m_p . u_k, the first sweep's D1, per instance hbar*omega, the
// interaction scale and the model-space size estimated from the inputs, and
// the channel's quantum numbers.  Its output is the remaining step D - D1 of
// the converged D.
#pragma once
#include "bcd_device.hpp"
#include "bcd_kernels.hpp"

namespace bcd {

constexpr int kLearnIn = 80;       // inputs per channel
constexpr int kLearnOut = 32;      // outputs per channel (the first kLearnTri used)
constexpr int kLearnNI = 6;        // orbits per channel covered by the network (Nmax <= 11)
constexpr int kLearnTri = kLearnNI * (kLearnNI + 1) / 2;
constexpr float kLearnScale = 10.f;   // D and D2 enter the network multiplied by this
constexpr int kLearnInstFeat = 4;  // per instance: hbar*omega estimate, rms of V, Nmax, spare

// Feature slots (offsets into the kLearnIn inputs of a channel).
constexpr int kFeatE = 0, kFeatEF = 6, kFeatOcc = 12, kFeatMask = 18, kFeatD2 = 24, kFeatGlob = 45, kFeatD1 = 53;

__host__ __device__ inline int learn_tri(int k, int kp) { return k * kLearnNI - k * (k - 1) / 2 + (kp - k); }

// inputs (after the Sigma(2) build, the propagator being the HF one) and a
// snapshot of the HF propagator: energies [n_gc x kMaxNI], amplitudes
// [n_gc x kMaxNI^2], counts [n_gc x 2].  features = 0: the snapshot only
// (what the residual check's fallback restarts from).
__global__ void k_learn_features(const InstDev *insts, const ChanDev *chans, int n_gc, PropView P,
                                 const float *inst_feat, const ChanState *cs, const long *row_fw_base,
                                 const long *row_bk_base, const double *rows_fw, const double *rows_bk, float *feat,
                                 double *hf_e, double *hf_x, int *hf_n, int features);
// D of the current propagator in the snapshot's basis, times kLearnScale,
// into out[gc * stride + off + t]: the first sc0 sweep's D1 (an input) and
// the converged D (the target)
__global__ void k_learn_density(const ChanDev *chans, int n_gc, PropView P, const double *hf_e, const double *hf_x,
                                const int *hf_n, float *out, int stride, int off);
// overwrite the density of the channels of seeded instances with the one
// built from D = D1 + the network's output
__global__ void k_learn_seed(const InstDev *insts, const ChanDev *chans, int n_gc, const int *seed_on,
                             const double *hf_e, const double *hf_x, const int *hf_n, const float *feat,
                             const float *pred, double *rho);
// put the HF snapshot back as the current propagator of the flagged instances
__global__ void k_learn_restore(const ChanDev *chans, int n_gc, PropView P, const int *flag, const double *hf_e,
                                const double *hf_x, const int *hf_n);

}  // namespace bcd
