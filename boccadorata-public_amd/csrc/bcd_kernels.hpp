// bcd_kernels.hpp -- kernel limits and declarations shared by host and device code.
#pragma once

#include "bcd_arrow.hpp"
#include "bcd_device.hpp"

namespace bcd {

constexpr int kMaxVt = 2048;      // nI * (alpha,beta) pairs * gamma of one Sigma(2) block
constexpr int kMaxPairs = 64;     // (alpha,beta) pairs of one block (+1 slot for the count)
constexpr int kSigGrid = 16384;    // waves of the Sigma(2) build, striding over the work list
constexpr int kSigStage = 128;    // fragment amplitudes of one Sigma(2) block side staged in LDS
constexpr int kSigRowBuf = 128;   // (pole, orbit) couplings of one block side kept in LDS for pass 2
constexpr int kMaxPiv = 8;        // Lanczos pivots per channel
constexpr int kMaxNB = 16;        // coupling columns of a Lanczos input row
constexpr int kLancRed = 2 + kMaxNB + kMaxPiv;
constexpr int kMaxNI = 16;        // orbits per subshell
constexpr int kMaxSub = 64;       // subshells per instance
// Lanczos workspace of one (channel, side), in vectors of M doubles: the
// Krylov vectors, then the nb coupling columns and E * v_last of up to
// kMaxPiv pivot chains
__host__ __device__ inline long lanczos_ws_vectors(long nlanc, int nb) { return nlanc + nb + kMaxPiv; }
// Dynamic LDS of a Lanczos tile of NT threads x PER elements: each thread's
// elements of the latest Krylov vector, or 0 (no stash) where that exceeds
// 128 KiB.
constexpr int lanczos_stash_bytes(int nt, int per) { return 8 * nt * per <= 128 * 1024 ? 8 * nt * per : 0; }
// Stride of the workspace vectors of a job with M poles: even, so that
// element pairs are 16-byte aligned.
__host__ __device__ inline int lanczos_stride(int M) { return M + (M & 1); }
constexpr int kMixMax = 8;        // iterates of the Anderson density mixing
constexpr double kMixStart = 1e-3;  // default density residual norm below which the mixing starts

__global__ void k_wbuild(const InstDev *insts, const ChanDev *chans, int gc_begin, int gc_end, double *W);
__global__ void k_density(const InstDev *insts, const ChanDev *chans, int gc_begin, int gc_end,
                          PropView P, double *rho);
__global__ void k_sigma_inf(const InstDev *insts, const ChanDev *chans, int gc_begin, int gc_end,
                            const InstState *is, const double *W, const double *rho, double *h);
__global__ void k_mix(const InstDev *insts, InstState *is, double *rho, double *hist, long rho_tot, int m,
                      double mix_start);
__global__ void k_sigma2_plan(const ChanDev *chans, const int *bpack, int gc_begin, int gc_end,
                              PropView P, const InstDev *insts, ChanState *cs, int *blk_fw_off,
                              int *blk_bk_off, Sig2Work *work_list, int *work_count, int only_ish,
                              const int *order);
// V-table size classes of the Sigma(2) build: each class is launched with
// LDS for its largest table, so the many small blocks run at high occupancy
constexpr int kSigVtClass[] = {256, 768};
constexpr int kSigMf = 48;        // largest (alpha, beta) pairs, (na, nb) fragment pairs and nI * n_g of a
                                  // block side contracted on the matrix cores
template <class T, class A, int MF>
__global__ void k_sigma2(const InstDev *insts, const ChanDev *chans, const BlockDev *blocks, PropView P,
                         const int *blk_fw_off, const int *blk_bk_off, const Sig2Work *work_list,
                         const int *work_count, const long *row_fw_base, const long *row_bk_base,
                         double *rows_fw, double *rows_bk, int vt_lo, int vt_cap, int c_cap);
// dynamic LDS bytes of k_sigma2 for a V table of vt_cap doubles
// dynamic LDS bytes of k_sigma2: V table of vt_cap entries, the staged
// amplitudes and couplings, and the pair rows (c_cap = the largest nI * n_g)
inline size_t sigma2_lds(int vt_cap, int c_cap, bool mfma = false) {
  return sizeof(double) * ((size_t)vt_cap + kSigStage + kSigRowBuf) + sizeof(int) * (kMaxPairs + 2 * (size_t)c_cap) +
         (mfma ? 16 + sizeof(float) * 2 * kSigMf * kSigMf : 0);
}
constexpr int kMergeMaxNI = 4;    // J-degenerate pole groups are merged in channels up to this nI
constexpr int kMergeMaxJ = 16;    // J blocks of one (a, b, g) triple (J <= j_a + j_b)
__global__ void k_merge_poles(const ChanDev *chans, int n_gc, ChanState *cs, const BlockDev *blocks,
                              const int *trip_start, const int *trip_blk, PropView P, const InstDev *insts,
                              const int *blk_fw_off, const int *blk_bk_off, const long *row_fw_base,
                              const long *row_bk_base, const double *rows_fw, const double *rows_bk,
                              double *crows_fw, double *crows_bk, const int *order);
__global__ void k_lanczos_prep(const ChanDev *chans, int gc_begin, int gc_end, PropView P, ChanState *cs,
                               double *piv, int lanc_kind, int piv_round, const long *row_fw_base,
                               const long *row_bk_base, double *rows_fw, double *rows_bk, double *crows_fw,
                               double *crows_bk, double *lrows, double *lws, LancJob *jobs);
constexpr int kLancPer = 24;      // largest register tile (elements per thread) of k_lanczos
template <int NT, int PER>
__global__ void k_lanczos(const LancJob *jobs, const int *order, int njobs);
// Tile shapes of k_lanczos: NT threads per (channel, side) job, PER
// elements of each vector per thread, so a job of M poles needs
// M <= NT * PER.  Jobs are grouped by M into buckets (upper bounds
// kLancBucket); every bucket has a default shape, and the batch autotuner can
// pick another one by timing the bucket's own jobs.
struct LancShape { int nt, per; };
constexpr LancShape kLancShapes[] = {{64, 4},   {64, 8},   {128, 8},   {192, 8},   {256, 8},
                                     {128, 16}, {192, 16}, {256, 16},  {384, 8},   {192, 24},
                                     {256, 24}, {384, 16}, {384, 24},  {768, 24},  {1024, 24},
                                     {1024, 32}, {1024, 40}, {512, 16}, {768, 16}, {1024, 16}, {512, 24},
                                     {768, 32}};
constexpr int kNLancShapes = sizeof(kLancShapes) / sizeof(kLancShapes[0]);
// Whether a tile keeps the pole energies in registers: three PER-element
// vectors (current, projected-out, energies) plus working registers within
// the VGPRs a thread gets at NT threads per block on one CU; otherwise the
// energies are re-read from the workspace where used.
constexpr bool lanczos_ev_in_regs(int nt, int per) {
  return 6 * per + 64 <= (nt <= 256 ? 256 : 512 / ((nt + 255) / 256));
}
// The last two buckets take channels whose HF occupations moved (more
// quasiholes than the bare propagator gives), found at Nmax 11.
constexpr int kLancBucket[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 24576, 32768, 40960};
constexpr int kNLancBuckets = sizeof(kLancBucket) / sizeof(kLancBucket[0]);
inline int lanczos_bucket(int M) {
  for (int b = 0; b < kNLancBuckets; ++b)
    if (M <= kLancBucket[b]) return b;
  return -1;
}
// default shape of each bucket
constexpr int kLancDefault[kNLancBuckets] = {0, 1, 2, 4, 9, 12, 13, 14, 15, 16};
inline bool lanczos_shape_fits(int shape, int M) {
  return shape >= 0 && shape < kNLancShapes && M <= kLancShapes[shape].nt * kLancShapes[shape].per;
}
// Launch shape `shape` of k_lanczos over njobs jobs (order may be NULL).
void launch_lanczos(int shape, int njobs, const LancJob *jobs, const int *order, hipStream_t stream);
__global__ void k_dyson_prep(const ChanDev *chans, int gc_begin, int gc_end, const ChanState *cs,
                             PropView P, int src, const double *h, const double *u1,
                             const long *row_fw_base, const long *row_bk_base, const double *rows_fw,
                             const double *rows_bk, const double *lrows, double *tri, double *mat,
                             double *kolt, DysonJob *jobs);
template <bool IN_LDS>
__global__ void k_tridiag(const DysonJob *jobs, int njobs);
// The pole-block kernels take njobs jobs, job k being jobs[order[k]] (or
// jobs[k] when order is null).
__global__ void k_tridiag_poles(const DysonJob *jobs, const int *order, int njobs, int ld);
__global__ void k_ql(const DysonJob *jobs, int njobs, int *status);
constexpr int kQLGroupNI = 8;     // tracked rows of k_ql_poles (one lane of an 8-lane group each)
__global__ void k_ql_poles(const DysonJob *jobs, const int *order, int njobs, int *status, int ld, int qrows);
template <bool IN_LDS, int NIMAX>
__global__ void k_arrow(const ArrowJob *jobs, int njobs, int *status);
// mfw, mbk, lfw, lbk, err, lmfw, lmbk (and one spare) of every channel
constexpr int kCsPack = 8;
__global__ void k_cs_pack(const ChanState *cs, int n_gc, int *out);
__global__ void k_poles_prep(const ChanDev *chans, int n_gc, const ChanState *cs, const double *lrows,
                             double *tri, double *mat, double *poles, DysonJob *jobs, const int *order, int njobs);
// Everything one Dyson sweep over all channels reads (k_arrow_sweep).
struct ArrowSweep {
  const ChanDev *chans;
  int n_gc, src;                 // src: 0 no dynamic Sigma, 1 unreduced rows, 2 diagonalised reduced sides
  const ChanState *cs;
  PropView P;
  const double *h, *u1;
  const long *row_fw_base, *row_bk_base;
  const double *rows_fw, *rows_bk, *poles;
  double *aws, *kolt;
  int *status;
  double *work;                  // accumulates the arrowhead FLOPs (may be null)
  unsigned long long *prof;      // 8 phase-time counters per channel (may be null)
  double *warm;                  // previous sweep's roots per channel (ChanDev::warm_off; may be null)
  int *warm_p;                   // their pole counts, kArrowNI per channel
  int skip_ni;                   // channels with nI <= skip_ni are left to k_dyson_static (0: none)
  int nt_class;                  // > 0: only channels whose arrow_threads(cap) equals it (one launch per class)
  int nb_class;                  // > 0: only channels whose arrow_ni_bound(nI) equals it
};
// Threads of the arrowhead block of a channel with cap propagator slots.
// The block size sets the order of the kernel's reductions, so it depends on
// the channel alone: a channel gives the same bits in any batch.
__host__ __device__ inline int arrow_threads(int cap) { return cap <= 64 ? 64 : (cap <= 128 ? 128 : kArrowMaxThreads); }
// Largest reduced side of the one-wave pole-block kernels (two columns per
// lane in k_tridiag_poles).
constexpr int kPoleWaveN = 128;
// Pole block of a reduced side with room for cap rows: the one-wave kernels
// (k_tridiag_poles, k_ql_poles) when it has at most kPoleWaveN rows and the
// channel at most kQLGroupNI orbits, the general ones otherwise -- chosen per
// side, not per batch, for the same reason.
__host__ __device__ inline int pole_path(int cap, int nI) { return cap <= kPoleWaveN && nI <= kQLGroupNI ? 0 : 1; }
template <bool IN_LDS, int NIMAX>
__global__ void k_arrow_sweep(ArrowSweep a);
// A sweep without dynamic self-energy: one thread per channel with nI <=
// kStaticNI solves its static Dyson matrix (static_dyson); the arrowhead
// kernel takes the larger channels of the same sweep, so every channel's
// path depends on its own size only, not on the rest of the batch.
__global__ void k_dyson_static(ArrowSweep a);
// Launch one arrowhead kernel variant: the workspace in LDS or HBM, the
// accumulator bound from nimax (arrow_ni_bound).
void launch_arrow_sweep(const ArrowSweep &a, int nimax, bool in_lds, int nt, size_t lds, hipStream_t st);
void launch_arrow(const ArrowJob *jobs, int njobs, int *status, int nimax, bool in_lds, int nt, size_t lds,
                  hipStream_t st);
int arrow_set_lds_limit(int bytes);
__global__ void k_heap_rows(const double *rows, int M, int T, int col, int desc, double *keys, int *idx,
                            int lds_keys, double *out);
// rows whose heapsort keys and indices fit the dynamic LDS of k_heap_rows
constexpr int kHeapLdsRows = (64 * 1024) / 12;
__global__ void k_pivot_round(const double *x, double *out, long n);
__global__ void k_finalize(const InstDev *insts, const ChanDev *chans, PropView P, const ChanState *cs,
                           const double *kolt, int conv_kind, int init);

}  // namespace bcd
