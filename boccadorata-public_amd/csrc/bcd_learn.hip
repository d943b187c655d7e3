// Learned start of the sc0 iteration: the network's per-channel inputs and
// targets, the seeded density of the first sc0 sweep and the restore of the
// HF propagator for the fallback to the plain iteration.
#include <hip/hip_runtime.h>

#include "bcd_learn.hpp"

namespace bcd {

namespace {

__device__ __forceinline__ double lwave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ int lprop_cur(const PropView &P, int inst) { return P.is ? P.is[inst].cur : 0; }

// The channel's HF orbital basis from the snapshot: solutions sorted by
// energy (stable), their occupations and amplitude vectors.
struct LearnBasis {
  double e[kLearnNI];
  double u[kLearnNI][kLearnNI];
  int occ[kLearnNI];
};

__device__ bool learn_basis(const double *he, const double *hx, const int *hn, int nI, LearnBasis &B) {
  const int nqp = hn[0], nqh = hn[1];
  if (nI < 1 || nI > kLearnNI || nqp < 0 || nqh < 0 || nqp + nqh != nI) return false;
  int ord[kLearnNI];
  for (int k = 0; k < nI; ++k) ord[k] = k;
  for (int k = 1; k < nI; ++k)
    for (int j = k; j > 0 && he[ord[j - 1]] > he[ord[j]]; --j) {
      const int t = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = t;
    }
  for (int k = 0; k < kLearnNI; ++k) {
    const bool in = k < nI;
    B.e[k] = in ? he[ord[k]] : 0.0;
    B.occ[k] = in ? (ord[k] >= nqp) : 0;
    for (int d = 0; d < kLearnNI; ++d) B.u[k][d] = (in && d < nI) ? hx[ord[k] * nI + d] : 0.0;
  }
  return true;
}

__device__ __forceinline__ float clip(double v) {
  const double c = 50.0;
  return (float)(v > c ? c : (v < -c ? -c : v));
}

}  // namespace

__global__ __launch_bounds__(256) void k_learn_features(const InstDev *insts, const ChanDev *chans, int n_gc,
                                                        PropView P, const float *inst_feat, const ChanState *cs,
                                                        const long *row_fw_base, const long *row_bk_base,
                                                        const double *rows_fw, const double *rows_bk, float *feat,
                                                        double *hf_e, double *hf_x, int *hf_n, int features) {
  __shared__ LearnBasis B;
  __shared__ int ok;
  __shared__ double red[4][kLearnTri];
  const int gc = blockIdx.x;
  if (gc >= n_gc) return;
  const ChanDev C = chans[gc];
  const InstDev I = insts[C.inst];
  const int cur = lprop_cur(P, C.inst), nI = C.nI, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nqp = P.nqp[cur][gc], nqh = P.nqh[cur][gc];
  double *he = hf_e + (long)gc * kMaxNI, *hx = hf_x + (long)gc * kMaxNI * kMaxNI;
  const int ns = nqp + nqh <= kMaxNI ? nqp + nqh : kMaxNI;
  const double *e = P.e[cur] + P.e_off[gc], *x = P.x[cur] + P.x_off[gc];
  for (int i = tid; i < ns; i += 256) he[i] = e[i];
  for (int i = tid; i < ns * nI; i += 256) hx[i] = x[i];
  if (tid == 0) { hf_n[2 * gc] = nqp; hf_n[2 * gc + 1] = nqh; }
  if (!features) return;
  float *f = feat + (long)gc * kLearnIn;
  for (int i = tid; i < kLearnIn; i += 256) f[i] = 0.f;
  __syncthreads();
  if (tid == 0) ok = learn_basis(he, hx, hf_n + 2 * gc, nI, B);
  __syncthreads();
  if (!ok) return;
  // second-order estimate D2 over the unreduced rows of both sides
  double acc[kLearnTri];
#pragma unroll
  for (int t = 0; t < kLearnTri; ++t) acc[t] = 0.0;
  const int w = 2 + nI;
  for (int side = 0; side < 2; ++side) {
    const int M = side ? cs[gc].mbk : cs[gc].mfw;
    const double *rows = side ? rows_bk + row_bk_base[gc] : rows_fw + row_fw_base[gc];
    const int want = side ? 0 : 1;    // 2p1h poles: holes; 2h1p poles: particles
    const double sg = side ? 1.0 : -1.0;
    for (int p = tid; p < M; p += 256) {
      const double *r = rows + (long)p * w;
      const double E = r[0];
      double q[kLearnNI];
#pragma unroll
      for (int k = 0; k < kLearnNI; ++k) {
        q[k] = 0.0;
        if (k < nI && B.occ[k] == want) {
          double s = 0.0;
          for (int d = 0; d < nI; ++d) s += r[2 + d] * B.u[k][d];
          double den = side ? B.e[k] - E : E - B.e[k];
          den = den >= 0.0 ? (den < 1e-3 ? 1e-3 : den) : (den > -1e-3 ? -1e-3 : den);
          q[k] = s / den;
        }
      }
#pragma unroll
      for (int k = 0; k < kLearnNI; ++k)
#pragma unroll
        for (int kp = k; kp < kLearnNI; ++kp) acc[learn_tri(k, kp)] += sg * q[k] * q[kp];
    }
  }
#pragma unroll
  for (int t = 0; t < kLearnTri; ++t) {
    const double s = lwave_sum(acc[t]);
    if (lane == 0) red[wv][t] = s;
  }
  __syncthreads();
  if (tid != 0) return;
  const InstState *S = P.is ? P.is + C.inst : nullptr;
  const double ef = chan_efermi(C, S);
  const float *g = inst_feat + kLearnInstFeat * C.inst;
  const double hw = g[0] > 0.f ? (double)g[0] : 1.0;
  for (int k = 0; k < nI; ++k) {
    f[kFeatE + k] = clip(B.e[k] / hw);
    f[kFeatEF + k] = clip((B.e[k] - ef) / hw);
    f[kFeatOcc + k] = (float)B.occ[k];
    f[kFeatMask + k] = 1.f;
  }
  for (int t = 0; t < kLearnTri; ++t) {
    const double s = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    f[kFeatD2 + t] = clip(s * kLearnScale);
  }
  f[kFeatGlob + 0] = (float)(hw / 20.0);
  f[kFeatGlob + 1] = g[1] / 2.f;
  f[kFeatGlob + 2] = g[2] / 11.f;
  f[kFeatGlob + 3] = (float)I.l[C.ish] / 11.f;
  f[kFeatGlob + 4] = (float)(C.j2 + 1) / 24.f;
  f[kFeatGlob + 5] = (float)(C.ch & 1);
  f[kFeatGlob + 6] = (float)nI / (float)kLearnNI;
  f[kFeatGlob + 7] = 1.f;
}

__global__ __launch_bounds__(64) void k_learn_density(const ChanDev *chans, int n_gc, PropView P, const double *hf_e,
                                                      const double *hf_x, const int *hf_n, float *out, int stride,
                                                      int off) {
  __shared__ LearnBasis B;
  __shared__ int ok;
  __shared__ double rho[kLearnNI * kLearnNI];
  const int gc = blockIdx.x;
  if (gc >= n_gc) return;
  const ChanDev C = chans[gc];
  const int nI = C.nI, tid = threadIdx.x;
  float *t = out + (long)gc * stride + off;
  for (int i = tid; i < kLearnTri; i += 64) t[i] = 0.f;
  if (tid == 0) ok = learn_basis(hf_e + (long)gc * kMaxNI, hf_x + (long)gc * kMaxNI * kMaxNI, hf_n + 2 * gc, nI, B);
  __syncthreads();
  if (!ok) return;
  const int cur = lprop_cur(P, C.inst);
  const int nqp = P.nqp[cur][gc], nqh = P.nqh[cur][gc];
  const double *x = P.x[cur] + P.x_off[gc];
  // rho(d, d') = sum over the quasiholes, in stored order per thread, then over the wave
  for (int dd = 0; dd < nI * nI; ++dd) {
    const int d = dd / nI, dp = dd % nI;
    double s = 0.0;
    for (int h = tid; h < nqh; h += 64) s += x[(long)(nqp + h) * nI + d] * x[(long)(nqp + h) * nI + dp];
    s = lwave_sum(s);
    if (tid == 0) rho[dd] = s;
  }
  __syncthreads();
  if (tid != 0) return;
  for (int k = 0; k < nI; ++k)
    for (int kp = k; kp < nI; ++kp) {
      double s = 0.0;
      for (int d = 0; d < nI; ++d) {
        double r = 0.0;
        for (int dp = 0; dp < nI; ++dp) r += rho[d * nI + dp] * B.u[kp][dp];
        s += B.u[k][d] * r;
      }
      if (k == kp) s -= (double)B.occ[k];
      t[learn_tri(k, kp)] = (float)(s * kLearnScale);
    }
}

__global__ __launch_bounds__(64) void k_learn_seed(const InstDev *insts, const ChanDev *chans, int n_gc,
                                                   const int *seed_on, const double *hf_e, const double *hf_x,
                                                   const int *hf_n, const float *feat, const float *pred,
                                                   double *rho) {
  const int gc = blockIdx.x * 64 + threadIdx.x;
  if (gc >= n_gc) return;
  const ChanDev C = chans[gc];
  if (!seed_on[C.inst]) return;
  LearnBasis B;
  const int nI = C.nI;
  if (!learn_basis(hf_e + (long)gc * kMaxNI, hf_x + (long)gc * kMaxNI * kMaxNI, hf_n + 2 * gc, nI, B)) return;
  const InstDev I = insts[C.inst];
  const float *pr = pred + (long)gc * kLearnOut, *d1 = feat + (long)gc * kLearnIn + kFeatD1;
  double D[kLearnNI][kLearnNI];
  for (int k = 0; k < nI; ++k)
    for (int kp = k; kp < nI; ++kp) {
      const int t = learn_tri(k, kp);
      const double v = ((double)d1[t] + (double)pr[t]) / kLearnScale;
      D[k][kp] = D[kp][k] = v;
    }
  for (int k = 0; k < nI; ++k) D[k][k] += (double)B.occ[k];
  double *out = rho + I.rho_off + I.qoff[C.ish];
  for (int d = 0; d < nI; ++d)
    for (int dp = 0; dp < nI; ++dp) {
      double s = 0.0;
      for (int k = 0; k < nI; ++k) {
        double r = 0.0;
        for (int kp = 0; kp < nI; ++kp) r += D[k][kp] * B.u[kp][dp];
        s += B.u[k][d] * r;
      }
      out[d * nI + dp] = s;
    }
}

__global__ __launch_bounds__(64) void k_learn_restore(const ChanDev *chans, int n_gc, PropView P, const int *flag,
                                                      const double *hf_e, const double *hf_x, const int *hf_n) {
  const int gc = blockIdx.x;
  if (gc >= n_gc) return;
  const ChanDev C = chans[gc];
  if (!flag[C.inst]) return;
  const int cur = lprop_cur(P, C.inst), nI = C.nI;
  const int nqp = hf_n[2 * gc], nqh = hf_n[2 * gc + 1];
  const int ns = nqp + nqh <= kMaxNI ? nqp + nqh : kMaxNI;
  double *e = P.e[cur] + P.e_off[gc], *x = P.x[cur] + P.x_off[gc];
  for (int i = threadIdx.x; i < ns; i += 64) e[i] = hf_e[(long)gc * kMaxNI + i];
  for (int i = threadIdx.x; i < ns * nI; i += 64) x[i] = hf_x[(long)gc * kMaxNI * kMaxNI + i];
  if (threadIdx.x == 0) { P.nqp[cur][gc] = nqp; P.nqh[cur][gc] = nqh; }
}

}  // namespace bcd
