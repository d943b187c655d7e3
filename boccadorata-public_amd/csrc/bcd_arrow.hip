// bcd_arrow.hip -- Dyson sweeps by successive arrowhead updates on gfx950.
//
// k_poles_prep + k_tridiag/k_ql (poles mode) diagonalise every
// Lanczos-reduced side once per solve; each sweep then runs one
// k_arrow_sweep block per channel (DysonMatrix.cpp:123-532 for the matrix,
// :436-484 for the split and ordering, :542-594 for Calc_En).
#include "bcd_kernels.hpp"
#include "bcd_wave.hpp"

namespace bcd {

namespace {

// The block an arrowhead body runs on: reductions go through wave-level DPP
// sums (or ballots) and one LDS slot per wave.
struct DevBlock {
  int *status;
  double *work;                  // FLOP counter (may be null)
  unsigned long long *prof;      // this block's 8 phase-time slots in wall-clock ticks (may be null)
  // ends phase k (k < 0: starts timing) begun at t0; thread 0 adds the
  // elapsed ticks to its block's own slot
  __device__ unsigned long long lap(int k, unsigned long long t0) const {
    if (!prof) return 0;
    const unsigned long long t = wall_clock64();
    if (k >= 0 && threadIdx.x == 0) prof[k] += t - t0;
    return t;
  }
  __device__ int tid() const { return threadIdx.x; }
  __device__ int nt() const { return blockDim.x; }
  __device__ void sync() const { __syncthreads(); }
  __device__ void fail() const { atomicAdd(status, 1); }
  __device__ void add_work(double f) const { if (work) atomicAdd(work, f); }
  // 1/x: the hardware estimate refined by two Newton steps (each doubles the
  // correct bits), within an ulp for normal x
  __device__ double rcp(double x) const {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
  }
  // exclusive prefix sum of keep (0 or 1) over the block (thread order);
  // total in tot.  Within a wave the prefix is the count of set ballot bits
  // below the lane.
  __device__ int scan01(int *tmp, int keep, int &tot) const {
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned long long b = __ballot(keep != 0);
    const int below = __builtin_amdgcn_mbcnt_hi((unsigned int)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned int)b, 0));
    if ((threadIdx.x & 63) == 0) tmp[w] = __popcll(b);
    __syncthreads();
    int off = 0, t = 0;
    for (int i = 0; i < nw; ++i) { const int c = tmp[i]; off += i < w ? c : 0; t += c; }
    __syncthreads();
    tot = t;
    return off + below;
  }
  __device__ double sum(double *red, double v) const {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
  }
  // N sums at once, with one pair of barriers; red holds N slots per wave
  template <int N>
  __device__ void sums(double *red, double *v) const {
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int k = 0; k < N; ++k) red[w * N + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double s = 0.0;
      for (int i = 0; i < nw; ++i) s += red[i * N + k];
      v[k] = s;
    }
    __syncthreads();
  }
  __device__ double max(double *red, double v) const {
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
  }
};

}  // namespace

// One block per channel; the workspace sits in LDS (IN_LDS) or in the job's
// global slab.  Blocks of inactive channels (n <= 0) leave at once.
template <bool IN_LDS, int NIMAX>
__global__ __launch_bounds__(kArrowMaxThreads, 4) void k_arrow(const ArrowJob *jobs, int njobs, int *status) {
  extern __shared__ double sm[];
  if ((int)blockIdx.x >= njobs) return;
  const ArrowJob J = jobs[blockIdx.x];
  if (J.n <= 0 || J.nI > NIMAX) return;
  arrow_body<NIMAX>(DevBlock{status, nullptr, nullptr}, J, IN_LDS ? sm : J.ws);
}

#define BCD_ARROW_INST(L, N)                                                            \
  template __global__ void k_arrow<L, N>(const ArrowJob *, int, int *);                 \
  template __global__ void k_arrow_sweep<L, N>(ArrowSweep);


// Jobs that diagonalise both reduced sides of every channel: job 2*gc + side,
// for the njobs job indices in order (all 2*n_gc when order is null).
__global__ void k_poles_prep(const ChanDev *chans, int n_gc, const ChanState *cs, const double *lrows,
                             double *tri, double *mat, double *poles, DysonJob *jobs, const int *order, int njobs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= njobs) return;
  const int t = order ? order[k] : k;
  if (t < 0 || t >= 2 * n_gc) return;
  const int gc = t >> 1, side = t & 1;
  const ChanDev C = chans[gc];
  const ChanState &Q = cs[gc];
  DysonJob &J = jobs[t];
  const int L = side ? Q.lbk : Q.lfw, capS = side ? C.lbk_cap : C.lfw_cap;
  J.nI = C.nI; J.j2 = C.j2; J.ch = C.ch; J.inst = C.inst; J.path = pole_path(capS, C.nI);
  J.n = L > 0 ? L : 0;
  J.nfw = J.n; J.nbk = 0;
  J.den_fw = J.den_bk = Q.den;
  J.npiv_fw = J.n > 0 ? (side ? Q.ipiv_bk[0] : Q.ipiv_fw[0]) : 0; J.npiv_bk = 0;
  J.efermi = 0.0;
  J.h = nullptr; J.u1 = nullptr;
  J.rfw = lrows + (side ? C.lbk_off : C.lfw_off); J.rbk = nullptr;
  J.ipiv_fw = side ? Q.ipiv_bk : Q.ipiv_fw; J.ipiv_bk = nullptr;
  J.tri = tri + C.tri_off + (side ? (long)C.lfw_cap * (2 + C.nI) : 0);
  // the channel's dense slab (cap^2 >= lfw_cap^2 + lbk_cap^2) when the side
  // does not fit in LDS
  J.mat = mat ? mat + C.mat_off + (side ? (long)C.lfw_cap * C.lfw_cap : 0) : nullptr;
  J.out_e = nullptr; J.out_x = nullptr; J.out_nqp = nullptr; J.out_nqh = nullptr; J.kolt = nullptr;
  J.poles = 1;
  J.ct_ld = capS;
  J.out_lam = poles + C.pole_off + (side ? (long)(1 + C.nI) * C.lfw_cap : 0);
  J.out_ct = J.out_lam + capS;
}

// The arrowhead job of channel gc for this sweep: static part, pole source
// (none, unreduced rows, or the diagonalised reduced sides), Fermi energy of
// the sweep's input propagator, output slots in the other buffer.  n = 0 for
// channels of instances that no longer iterate.
__device__ ArrowJob arrow_job(const ArrowSweep &a, int gc) {
  const ChanDev &C = a.chans[gc];
  ArrowJob J;
  J.n = 0;
  const InstState *S = a.P.is ? a.P.is + C.inst : nullptr;
  if (S && !S->active) return J;
  const ChanState &Q = a.cs[gc];
  const int cur = S ? S->cur : 0, nxt = 1 - cur, nI = C.nI;
  J.nI = nI; J.j2 = C.j2; J.inst = C.inst;
  J.h = a.h + C.h_off; J.u1 = a.u1 + C.u1_off;
  for (int s = 0; s < 2; ++s) { J.np[s] = 0; J.es[s] = 1; J.ci[s] = 1; J.cr[s] = 1; J.E[s] = nullptr; J.C[s] = nullptr; }
  if (a.src == 1) {
    const int T = 2 + nI;
    J.np[0] = Q.mfw; J.np[1] = Q.mbk;
    J.E[0] = a.rows_fw + a.row_fw_base[gc]; J.E[1] = a.rows_bk + a.row_bk_base[gc];
    for (int s = 0; s < 2; ++s) { J.C[s] = J.E[s] + 2; J.es[s] = T; J.ci[s] = T; J.cr[s] = 1; }
  } else if (a.src == 2) {
    J.np[0] = Q.lfw > 0 ? Q.lfw : 0; J.np[1] = Q.lbk > 0 ? Q.lbk : 0;
    J.E[0] = a.poles + C.pole_off;
    J.C[0] = J.E[0] + C.lfw_cap; J.cr[0] = C.lfw_cap;
    J.E[1] = a.poles + C.pole_off + (long)(1 + nI) * C.lfw_cap;
    J.C[1] = J.E[1] + C.lbk_cap; J.cr[1] = C.lbk_cap;
  }
  J.n = nI + J.np[0] + J.np[1];
  if (J.n > C.cap) {
    J.n = 0;
    if (a.P.is && threadIdx.x == 0) a.P.is[C.inst].err = 2;
    return J;
  }
  J.efermi = chan_efermi(C, S);
  J.ws = a.aws ? a.aws + C.aws_off : nullptr;
  J.out_e = a.P.e[nxt] + a.P.e_off[gc];
  J.out_x = a.P.x[nxt] + a.P.x_off[gc];
  J.out_nqp = a.P.nqp[nxt] + gc;
  J.out_nqh = a.P.nqh[nxt] + gc;
  J.kolt = a.kolt + 3L * gc;
  J.warm = a.warm ? a.warm + C.warm_off : nullptr;
  J.warm_p = a.warm_p ? a.warm_p + (long)kArrowNI * gc : nullptr;
  J.wld = C.cap;
  return J;
}

// One Dyson sweep of every channel: one block per channel builds its job and
// solves it.
template <bool IN_LDS, int NIMAX>
__global__ __launch_bounds__(kArrowMaxThreads, 4) void k_arrow_sweep(ArrowSweep a) {
  extern __shared__ double sm[];
  if ((int)blockIdx.x >= a.n_gc) return;
  if (a.chans[blockIdx.x].nI <= a.skip_ni) return;
  if (a.nt_class && arrow_threads(a.chans[blockIdx.x].cap) != a.nt_class) return;
  if (a.nb_class && arrow_ni_bound(a.chans[blockIdx.x].nI) != a.nb_class) return;
  const ArrowJob J = arrow_job(a, blockIdx.x);
  if (J.n <= 0 || J.nI > NIMAX) return;
  arrow_body<NIMAX>(DevBlock{a.status, a.work, a.prof ? a.prof + 8L * blockIdx.x : nullptr}, J,
                    IN_LDS ? sm : J.ws);
}

// The HF sweep's Dyson problems: one thread per channel (static_dyson),
// writing what k_arrow_sweep writes for a channel without poles, including
// the pole count 0 of every head for the next sweep's warm start.
__global__ __launch_bounds__(64) void k_dyson_static(ArrowSweep a) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= a.n_gc) return;
  const ChanDev &C = a.chans[gc];
  if (C.nI > kStaticNI) return;
  const InstState *S = a.P.is ? a.P.is + C.inst : nullptr;
  if (S && !S->active) return;
  const int nxt = 1 - (S ? S->cur : 0);
  const double ef = chan_efermi(C, S);
  if (static_dyson<kStaticNI>(C.nI, C.j2, a.h + C.h_off, a.u1 + C.u1_off, ef, a.P.e[nxt] + a.P.e_off[gc],
                              a.P.x[nxt] + a.P.x_off[gc], a.P.nqp[nxt] + gc, a.P.nqh[nxt] + gc, a.kolt + 3L * gc))
    atomicAdd(a.status, 1);
  if (a.warm_p)
    for (int k = 0; k < C.nI && k < kArrowNI; ++k) a.warm_p[(long)kArrowNI * gc + k] = 0;
}

BCD_ARROW_INST(true, 4)
BCD_ARROW_INST(true, 8)
BCD_ARROW_INST(true, 16)
BCD_ARROW_INST(false, 4)
BCD_ARROW_INST(false, 8)
BCD_ARROW_INST(false, 16)

void launch_arrow_sweep(const ArrowSweep &a, int nimax, bool in_lds, int nt, size_t lds, hipStream_t st) {
  const int nb = arrow_ni_bound(nimax);
  const dim3 g(a.n_gc), b(nt);
  if (in_lds) {
    if (nb == 4) hipLaunchKernelGGL((k_arrow_sweep<true, 4>), g, b, lds, st, a);
    else if (nb == 8) hipLaunchKernelGGL((k_arrow_sweep<true, 8>), g, b, lds, st, a);
    else hipLaunchKernelGGL((k_arrow_sweep<true, 16>), g, b, lds, st, a);
  } else {
    if (nb == 4) hipLaunchKernelGGL((k_arrow_sweep<false, 4>), g, b, 0, st, a);
    else if (nb == 8) hipLaunchKernelGGL((k_arrow_sweep<false, 8>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_arrow_sweep<false, 16>), g, b, 0, st, a);
  }
}

void launch_arrow(const ArrowJob *jobs, int njobs, int *status, int nimax, bool in_lds, int nt, size_t lds,
                  hipStream_t st) {
  const int nb = arrow_ni_bound(nimax);
  const dim3 g(njobs), b(nt);
  if (in_lds) {
    if (nb == 4) hipLaunchKernelGGL((k_arrow<true, 4>), g, b, lds, st, jobs, njobs, status);
    else if (nb == 8) hipLaunchKernelGGL((k_arrow<true, 8>), g, b, lds, st, jobs, njobs, status);
    else hipLaunchKernelGGL((k_arrow<true, 16>), g, b, lds, st, jobs, njobs, status);
  } else {
    if (nb == 4) hipLaunchKernelGGL((k_arrow<false, 4>), g, b, 0, st, jobs, njobs, status);
    else if (nb == 8) hipLaunchKernelGGL((k_arrow<false, 8>), g, b, 0, st, jobs, njobs, status);
    else hipLaunchKernelGGL((k_arrow<false, 16>), g, b, 0, st, jobs, njobs, status);
  }
}

// Allow the large dynamic LDS on every in-LDS arrowhead variant.
int arrow_set_lds_limit(int bytes) {
  const void *f[6] = {(const void *)k_arrow_sweep<true, 4>, (const void *)k_arrow_sweep<true, 8>,
                      (const void *)k_arrow_sweep<true, 16>, (const void *)k_arrow<true, 4>,
                      (const void *)k_arrow<true, 8>, (const void *)k_arrow<true, 16>};
  for (const void *p : f)
    if (hipFuncSetAttribute(p, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) return -1;
  return 0;
}

}  // namespace bcd
