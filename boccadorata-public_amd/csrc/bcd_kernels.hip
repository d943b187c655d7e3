// bcd_kernels.hip -- gfx950 kernels of the second-order SCGF hot path (FP64).
//
// Stage map (reference files under source/ of BoccaDorata v1.0):
//   k_wbuild, k_density, k_sigma_inf  Sigma^inf     Sp_Propagator_a.cpp:39-130
//   k_sigma2_plan, k_sigma2           Sigma(2)      Self_energy_ext_Op_c.cpp:35-113, 323-561
//   k_lanczos_prep, k_lanczos         Lanczos       Self_energy_ext_Op_e.cpp:55-232, Lanczos.cpp:44-286
//   k_dyson_prep, k_tridiag, k_ql     Dyson         DysonMatrix.cpp:123-532
//   k_finalize                        sweep control Main.cpp:403-451, Sp_Propagator.cpp:1189-1213,
//                                                   DysonMatrix.cpp:619-635
#include "bcd_kernels.hpp"
#include "bcd_order.hpp"
#include "bcd_wave.hpp"

namespace bcd {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ int wave_isum(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Sum nv values over the whole block; every thread gets the totals.  red
// must hold (blockDim/64)*nv doubles of LDS.  Deterministic order.
__device__ void block_sum(double *vals, int nv, double *red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < nv; ++k) vals[k] = wave_sum(vals[k]);
  if (lane == 0)
    for (int k = 0; k < nv; ++k) red[w * nv + k] = vals[k];
  __syncthreads();
  for (int k = 0; k < nv; ++k) {
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[i * nv + k];
    vals[k] = s;
  }
  __syncthreads();
}

__device__ __forceinline__ int prop_cur(const PropView &P, int inst) {
  return P.is ? P.is[inst].cur : 0;
}

__device__ __forceinline__ bool tri_ok(int a2, int b2, int c2) {
  const int s1 = a2 + b2 - c2, s2 = b2 + c2 - a2, s3 = c2 + a2 - b2;
  return s1 >= 0 && s2 >= 0 && s3 >= 0 && !(s1 & 1) && !(s2 & 1) && !(s3 & 1);
}

// ------------------------------------------------------------- Sigma^inf

// W[rc][q] = sum_J (2J+1)/(2j+1) f_rm f_cu V^J(r m; c u), rc = r(r+1)/2+c with
// r >= c, q = qoff[v] + u*no_v + m; f = sqrt2 on the channel's own orbit and
// that term dropped unless (j + j_v + J) is odd (Sp_Propagator_a.cpp:62-117).
__global__ void k_wbuild(const InstDev *insts, const ChanDev *chans, int gc_begin, int gc_end,
                         double *W) {
  const int gc = gc_begin + blockIdx.x;
  if (gc >= gc_end) return;
  const ChanDev C = chans[gc];
  const InstDev I = insts[C.inst];
  const int nI = C.nI, nrc = nI * (nI + 1) / 2, nq = I.nq;
  for (long idx = threadIdx.x; idx < (long)nrc * nq; idx += blockDim.x) {
    const int rc = (int)(idx / nq), q = (int)(idx % nq);
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= rc) ++r;
    const int c = rc - r * (r + 1) / 2;
    int v = 0;
    while (v + 1 < I.nsubsh && I.qoff[v + 1] <= q) ++v;
    const int nv = I.no[v], um = q - I.qoff[v], u = um / nv, m = um % nv;
    const bool self = (v == C.ish);
    const int ia = C.orb0 + r, ib = C.orb0 + c, im = I.orb0[v] + m, iu = I.orb0[v] + u;
    const int j2I = C.j2, j2v = I.j2[v];
    double acc = 0.0;
    for (int J2 = abs(j2I - j2v); J2 <= j2I + j2v; J2 += 2) {
      const int nJ = ((j2I + j2v + J2) / 2) % 2;
      if (self && m == r && nJ != 1) continue;
      if (self && u == c && nJ != 1) continue;
      const double x2 = (self && m == r) ? kSqrt2 : 1.0;
      const double x3 = (self && u == c) ? kSqrt2 : 1.0;
      const double x1 = (double)(J2 + 1) / (double)(j2I + 1);
      acc += x1 * (x2 * x3 * vget(I, ia, im, ib, iu, J2 / 2));
    }
    W[C.w_off + idx] = acc;
  }
}

// rho_v(u,m) = sum over the quasiholes k of subshell v of Y_k(u) Y_k(m), in
// stored order (the x6 sum of Sp_Propagator_a.cpp:103-110).
__global__ void k_density(const InstDev *insts, const ChanDev *chans, int gc_begin, int gc_end,
                          PropView P, double *rho) {
  const int gc = gc_begin + blockIdx.x;
  if (gc >= gc_end) return;
  const ChanDev C = chans[gc];
  if (P.is && !P.is[C.inst].active) return;
  const InstDev I = insts[C.inst];
  const int cur = prop_cur(P, C.inst), nI = C.nI;
  const int nqp = P.nqp[cur][gc], nqh = P.nqh[cur][gc];
  const double *x = P.x[cur] + P.x_off[gc];
  double *out = rho + I.rho_off + I.qoff[C.ish];
  for (int idx = threadIdx.x; idx < nI * nI; idx += blockDim.x) {
    const int u = idx / nI, m = idx % nI;
    double s = 0.0;
    for (int k = 0; k < nqh; ++k) s += x[(long)(nqp + k) * nI + u] * x[(long)(nqp + k) * nI + m];
    out[idx] = s;
  }
}

// h(r,c) = h(c,r) = sum_q W[rc][q] rho[q]: Sigma^inf of every channel.
__global__ void k_sigma_inf(const InstDev *insts, const ChanDev *chans, int gc_begin, int gc_end,
                            const InstState *is, const double *W, const double *rho, double *h) {
  const int gc = gc_begin + blockIdx.x;
  if (gc >= gc_end) return;
  const ChanDev C = chans[gc];
  if (is && !is[C.inst].active) return;
  const InstDev I = insts[C.inst];
  const int nI = C.nI, nq = I.nq, lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double *r0 = rho + I.rho_off;
  for (int rc = w; rc < nI * (nI + 1) / 2; rc += nw) {
    const double *wr = W + C.w_off + (long)rc * nq;
    double s = 0.0;
    for (int q = lane; q < nq; q += 64) s += wr[q] * r0[q];
    s = wave_sum(s);
    if (lane == 0) {
      int r = 0;
      while ((r + 1) * (r + 2) / 2 <= rc) ++r;
      const int c = rc - r * (r + 1) / 2;
      h[C.h_off + r * nI + c] = s;
      h[C.h_off + c * nI + r] = s;
    }
  }
}

// Anderson mixing of the hole density (type II: H. F. Walker and P. Ni, SIAM J.
// Numer. Anal. 49 (2011) 1715).  The sweep map is rho_in -> Sigma^inf -> Dyson
// -> rho_out; the reference iterates it plainly (rho_in of the next sweep =
// rho_out).  Here, with x = rho_in of the last sweep and f = rho_out - x,
//   x_new = x + f - sum_i gamma_i (dX_i + dF_i),
//   gamma = argmin | f - sum_i gamma_i dF_i |,
// over the differences dX_i, dF_i of the last m iterates; the fixed point
// rho_out(rho) = rho is the same.  One block per instance; hist holds the
// iterates X and residuals F in rings of m + 1 slots, each slot laid out like
// rho (rho_tot doubles).  Mixing starts once |f| <= mix_start: the random
// interactions here give several self-consistent solutions, and the plain
// iteration decides which one a solve converges to; until then the steps are
// plain.  A residual that grows tenfold restarts the history.
__global__ __launch_bounds__(256) void k_mix(const InstDev *insts, InstState *is, double *rho, double *hist,
                                             long rho_tot, int m, double mix_start) {
  InstState *S = is + blockIdx.x;
  if (!S->active) return;
  const InstDev I = insts[blockIdx.x];
  const int nq = I.nq, tid = threadIdx.x, mm = m + 1, k = S->mix_k;
  double *r = rho + I.rho_off;
  auto X = [&](int it) { return hist + (long)(it % mm) * rho_tot + I.rho_off; };
  auto F = [&](int it) { return hist + (long)(mm + it % mm) * rho_tot + I.rho_off; };
  if (k == 0) {
    // first sweep of the stage: rho_in is the starting propagator's density
    double *x0 = X(0);
    for (int q = tid; q < nq; q += blockDim.x) x0[q] = r[q];
    __syncthreads();
    if (tid == 0) { S->mix_k = 1; S->mix_s = 0; S->mix_fn = 0.0; }
    return;
  }
  constexpr int kNd = kMixMax * (kMixMax + 1) / 2 + kMixMax + 1;
  __shared__ double red[4][kNd];
  __shared__ double G[kNd];
  __shared__ double A[kMixMax][kMixMax + 1];
  __shared__ double gam[kMixMax];
  __shared__ int n_sh;
  const double *x = X(k - 1);
  double *f = F(k - 1);
  for (int q = tid; q < nq; q += blockDim.x) f[q] = r[q] - x[q];
  int s = S->mix_s;
  int n = min(m, k - 1 - s);
  // dot products of the differences dF_i = F(k-i) - F(k-i-1), i = 1..n
  double acc[kNd];
#pragma unroll
  for (int t = 0; t < kNd; ++t) acc[t] = 0.0;
  for (int q = tid; q < nq; q += blockDim.x) {
    double df[kMixMax];
    const double fq = f[q];
#pragma unroll
    for (int i = 0; i < kMixMax; ++i) df[i] = (i < n) ? F(k - 1 - i)[q] - F(k - 2 - i)[q] : 0.0;
    int t = 0;
#pragma unroll
    for (int i = 0; i < kMixMax; ++i) {
#pragma unroll
      for (int j = 0; j <= i; ++j) acc[t++] += df[i] * df[j];
    }
#pragma unroll
    for (int i = 0; i < kMixMax; ++i) acc[t++] += df[i] * fq;
    acc[t] += fq * fq;
  }
  // only the products of the n differences in use are reduced
  const int lane = tid & 63, wv = tid >> 6;
  constexpr int kTri = kMixMax * (kMixMax + 1) / 2;
#pragma unroll
  for (int t = 0; t < kNd; ++t) {
    const bool used = t < kTri ? t < n * (n + 1) / 2 : t - kTri < n || t == kNd - 1;
    if (!used) continue;
    const double v = wave_sum(acc[t]);
    if (lane == 0) red[wv][t] = v;
  }
  __syncthreads();
  if (tid == 0) {
    for (int t = 0; t < kNd; ++t) {
      const bool used = t < kTri ? t < n * (n + 1) / 2 : t - kTri < n || t == kNd - 1;
      G[t] = 0.0;
      for (int w = 0; used && w < (int)(blockDim.x >> 6); ++w) G[t] += red[w][t];
    }
    const double fn = sqrt(G[kNd - 1]);
    // plain steps until the residual is small (the plain iteration picks the
    // fixed point, mixing then converges onto it); restart on growth
    if (fn > mix_start || (n > 0 && fn > 10.0 * S->mix_fn)) { s = k - 1; n = 0; }
    // normal equations (dF^T dF) gamma = dF^T f, Gaussian elimination with
    // partial pivoting; a singular or non-finite system falls back to the
    // plain step
    int t = 0;
    for (int i = 0; i < kMixMax; ++i)
      for (int j = 0; j <= i; ++j, ++t)
        if (i < n) { A[i][j] = G[t]; A[j][i] = G[t]; }
    for (int i = 0; i < n; ++i) A[i][n] = G[kMixMax * (kMixMax + 1) / 2 + i];
    double tr = 0.0;
    for (int i = 0; i < n; ++i) tr += A[i][i];
    for (int i = 0; i < n; ++i) A[i][i] += 1e-13 * tr;
    bool ok = n > 0 && tr > 0.0;
    for (int c = 0; c < n && ok; ++c) {
      int piv = c;
      for (int i = c + 1; i < n; ++i)
        if (fabs(A[i][c]) > fabs(A[piv][c])) piv = i;
      if (!(fabs(A[piv][c]) > 1e-300)) { ok = false; break; }
      if (piv != c)
        for (int j = 0; j <= n; ++j) { const double tmp = A[c][j]; A[c][j] = A[piv][j]; A[piv][j] = tmp; }
      for (int i = c + 1; i < n; ++i) {
        const double l = A[i][c] / A[c][c];
        for (int j = c; j <= n; ++j) A[i][j] -= l * A[c][j];
      }
    }
    for (int i = n - 1; i >= 0 && ok; --i) {
      double v = A[i][n];
      for (int j = i + 1; j < n; ++j) v -= A[i][j] * gam[j];
      gam[i] = v / A[i][i];
      if (!isfinite(gam[i])) ok = false;
    }
    if (!ok) n = 0;
    n_sh = n;
    S->mix_k = k + 1;
    S->mix_s = s;
    S->mix_fn = fn;
  }
  __syncthreads();
  n = n_sh;
  double *xn = X(k);
  for (int q = tid; q < nq; q += blockDim.x) {
    double v = x[q] + f[q];
    for (int i = 0; i < n; ++i)
      v -= gam[i] * (X(k - 1 - i)[q] - X(k - 2 - i)[q] + F(k - 1 - i)[q] - F(k - 2 - i)[q]);
    xn[q] = v;
    r[q] = v;
  }
}

// ------------------------------------------------------------- Sigma(2)

__global__ void k_cs_pack(const ChanState *cs, int n_gc, int *out) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= n_gc) return;
  const ChanState &S = cs[gc];
  int *q = out + (long)kCsPack * gc;
  q[0] = S.mfw; q[1] = S.mbk; q[2] = S.lfw; q[3] = S.lbk; q[4] = S.err; q[5] = S.lmfw; q[6] = S.lmbk; q[7] = 0;
}

// Pole counts of each (J,a,b,g) block from the fragment counts of the current
// propagator (Count_2p1h_2h1p_poles, Self_energy_ext_Op_c.cpp:76-90), row
// offsets in build order, and a compact list of the non-empty blocks.
__global__ __launch_bounds__(256) void k_sigma2_plan(const ChanDev *chans, const int *bpack, int gc_begin,
                                                    int gc_end, PropView P, const InstDev *insts, ChanState *cs,
                                                    int *blk_fw_off, int *blk_bk_off, Sig2Work *work_list,
                                                    int *work_count, int only_ish, const int *order) {
  // channels gc_begin + order[k] (or in index order), k < gc_end - gc_begin
  if ((int)blockIdx.x >= gc_end - gc_begin) return;
  const int gc = gc_begin + (order ? order[blockIdx.x] : (int)blockIdx.x);
  const ChanDev C = chans[gc];
  if (only_ish >= 0 && C.ish != only_ish) {
    if (threadIdx.x == 0) { cs[gc].mfw = 0; cs[gc].mbk = 0; }
    return;
  }
  __shared__ int wsum[3][4];
  __shared__ int list_base;
  if (threadIdx.x == 0) cs[gc].err = 0;
  __syncthreads();
  const InstDev I = insts[C.inst];
  const int cur = prop_cur(P, C.inst), tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = blockDim.x;
  // pole counts of block b per side, and the six fragment counts
  auto counts = [&](int b, int &cfw, int &cbk, int *f6) {
    cfw = 0; cbk = 0;
    if (b >= C.blk_cnt) return;
    const int pk = bpack[C.blk_off + b];
    const int ba = pk & 0xff, bb = (pk >> 8) & 0xff, bg = (pk >> 16) & 0xff, bJab = (pk >> 24) & 1;
    const int ga = I.gc0 + ba, gb = I.gc0 + bb, gg = I.gc0 + bg;
    int na = P.nqp[cur][ga], nb = P.nqp[cur][gb], ng = P.nqh[cur][gg];
    if (f6) { f6[0] = na; f6[1] = nb; f6[2] = ng; }
    cfw = na * nb * ng;
    if (ba == bb) { nb = na + 1; if (bJab == 1) nb -= 2; cfw = (na * nb / 2) * ng; }
    na = P.nqh[cur][ga]; nb = P.nqh[cur][gb]; ng = P.nqp[cur][gg];
    if (f6) { f6[3] = na; f6[4] = nb; f6[5] = ng; }
    cbk = na * nb * ng;
    if (ba == bb) { nb = na + 1; if (bJab == 1) nb -= 2; cbk = (na * nb / 2) * ng; }
  };
  // first pass: the channel's non-empty blocks, then one atomic for its
  // slice of the work list
  int nz = 0;
  for (int b0 = 0; b0 < C.blk_cnt; b0 += nt) {
    int cfw, cbk;
    counts(b0 + tid, cfw, cbk, nullptr);
    nz += cfw + cbk > 0;
  }
  for (int o = 32; o > 0; o >>= 1) nz += __shfl_xor(nz, o, 64);
  if (lane == 0) wsum[2][wv] = nz;
  __syncthreads();
  if (tid == 0) {
    int tot = 0;
    for (int w = 0; w < (nt >> 6); ++w) tot += wsum[2][w];
    list_base = tot ? atomicAdd(work_count, tot) : 0;
  }
  __syncthreads();
  int base_fw = 0, base_bk = 0, base_nz = list_base;
  for (int b0 = 0; b0 < C.blk_cnt; b0 += nt) {
    const int b = b0 + tid;
    int cfw, cbk, f6[6] = {0, 0, 0, 0, 0, 0};
    counts(b, cfw, cbk, f6);
    const bool take = cfw + cbk > 0;
    const unsigned long long mask = __ballot(take);
    // inclusive scans: within each wave, then the waves before it
    int sfw = cfw, sbk = cbk;
    for (int o = 1; o < 64; o <<= 1) {
      const int tf = __shfl_up(sfw, o, 64), tb = __shfl_up(sbk, o, 64);
      if (lane >= o) { sfw += tf; sbk += tb; }
    }
    if (lane == 63) { wsum[0][wv] = sfw; wsum[1][wv] = sbk; }
    if (lane == 0) wsum[2][wv] = __popcll(mask);
    __syncthreads();
    int pf = 0, pb = 0, pz = 0, tf = 0, tb = 0, tz = 0;
    for (int w = 0; w < (nt >> 6); ++w) {
      if (w < wv) { pf += wsum[0][w]; pb += wsum[1][w]; pz += wsum[2][w]; }
      tf += wsum[0][w]; tb += wsum[1][w]; tz += wsum[2][w];
    }
    __syncthreads();
    if (b < C.blk_cnt) {
      const int ofw = base_fw + pf + sfw - cfw, obk = base_bk + pb + sbk - cbk;
      blk_fw_off[C.blk_off + b] = ofw;
      blk_bk_off[C.blk_off + b] = obk;
      if (take) {
        // the work item packs fragment counts into 16 bits
        if ((f6[0] | f6[1] | f6[2] | f6[3] | f6[4] | f6[5]) >> 16) cs[gc].err = 3;
        Sig2Work W;
        W.bidx = (int)(C.blk_off + b); W.gc = gc; W.inst = C.inst;
        W.n01 = f6[0] | (f6[1] << 16); W.n23 = f6[2] | (f6[3] << 16); W.n45 = f6[4] | (f6[5] << 16);
        W.off_fw = ofw; W.off_bk = obk;
        {
          const int pk = bpack[C.blk_off + b];
          const int ba = pk & 0xff, bb = (pk >> 8) & 0xff, bg = (pk >> 16) & 0xff, bJab = (pk >> 24) & 1;
          const int noa = I.no[ba], nob = I.no[bb];
          const int npair = ba == bb ? noa * (noa + 1) / 2 - bJab * noa : noa * nob;
          W.nvt = C.nI * npair * I.no[bg];
          W._pad = 0;
        }
        work_list[base_nz + pz + __popcll(mask & ((1ull << lane) - 1ull))] = W;
      }
    }
    base_fw += tf;
    base_bk += tb;
    base_nz += tz;
  }
  if (tid == 0) {
    cs[gc].mfw = base_fw;
    cs[gc].mbk = base_bk;
  }
}

// Pass 1 of the coupling build for one block side: lane q owns the coupling
// m(pole p, orbit d), q = p * nI + d, summed over (alpha beta, gamma) in the
// reference's loop order (Self_energy_ext_Op_c.cpp:454-468) with contraction
// off, so the value is bit-identical to the reference's.  The amplitudes
// come from S-typed arrays: the block's fragments staged in LDS, or the
// propagator in HBM when they do not fit the staging area.
// The (alpha beta, gamma) sum of one (pole, d) for a compile-time gamma
// count NOG: the pole's gamma amplitudes stay in registers and the NOG table
// entries of each (alpha, beta) are loaded together; the terms are added in
// the reference's order.  The entries of gamma == d when nJgd = 1 are exact
// zeros in the table, and adding them leaves the sum's bits unchanged (it
// starts at +0 and never becomes -0), so they need no test.
template <class T, class A, class S, int NOG>
__device__ __forceinline__ A sigma2_term_sum(const S *ya, const S *yb, const S *yg, const T *v, int noa, int nob,
                                             int a_eq_b, int nJab, T xJab, T kS2) {
#pragma clang fp contract(off)
  T yq[NOG];
#pragma unroll
  for (int gm = 0; gm < NOG; ++gm) yq[gm] = (T)yg[gm];
  A x1 = (A)0.0;
#pragma unroll 1
  for (int al = 0; al < noa; ++al) {
    const T yaa = (T)ya[al], yba = (T)yb[al];
#pragma unroll 1
    for (int be = a_eq_b ? al + nJab : 0; be < nob; ++be, v += NOG) {
      T vv[NOG];
#pragma unroll
      for (int gm = 0; gm < NOG; ++gm) vv[gm] = v[gm];
      T x2 = yaa * (T)yb[be];
      if (a_eq_b) {
        x2 += xJab * (T)ya[be] * yba;
        if (al == be) x2 /= kS2;
      }
#pragma unroll
      for (int gm = 0; gm < NOG; ++gm) x1 += (A)(vv[gm] * x2 * yq[gm]);
    }
  }
  return x1;
}

template <class T, class A, class S, bool STAGED>
__device__ __forceinline__ void sigma2_pass1(const S *xa, const S *xb, const S *xg, const T *vt, int noa, int nob,
                                             int nog, int nI, int npole, int a_eq_b, int nJab, int nJgd, int cA,
                                             int cB, int cG, int orb_g, int orb_d0, double *rows, int w,
                                             double *rb) {
#pragma clang fp contract(off)
  const T xJab = (T)(1 - 2 * nJab);
  const T kS2 = (T)kSqrt2;
  for (int q = threadIdx.x; q < npole * nI; q += 64) {
    const int p = q / nI, d = q % nI;
    // fragments of pole p, in the reference's enumeration order
    int na, nb, ng, r = p / cG;
    ng = p % cG;
    if (a_eq_b) {
      na = 0;
      int len = cA - nJab;
      while (r >= len) { r -= len; ++na; --len; }
      nb = na + nJab + r;
    } else {
      na = r / cB; nb = r % cB;
    }
    const S *ya = xa + na * noa, *yb = xb + nb * nob, *yg = xg + ng * nog;
    const int od = orb_d0 + d;
    T x4 = (T)1.0;
    if (a_eq_b && na == nb) x4 /= kS2;
    const T *v = vt + d * (a_eq_b ? (noa * (noa + 1) / 2 - nJab * noa) : noa * nob) * nog;
    A x1 = (A)0.0;
    // staged amplitudes (S == T): gamma count specialised; the rare path
    // that reads the propagator in HBM keeps the general loop
    switch (STAGED ? nog : 0) {
      case 1: x1 = sigma2_term_sum<T, A, S, 1>(ya, yb, yg, v, noa, nob, a_eq_b, nJab, xJab, kS2); break;
      case 2: x1 = sigma2_term_sum<T, A, S, 2>(ya, yb, yg, v, noa, nob, a_eq_b, nJab, xJab, kS2); break;
      case 3: x1 = sigma2_term_sum<T, A, S, 3>(ya, yb, yg, v, noa, nob, a_eq_b, nJab, xJab, kS2); break;
      case 4: x1 = sigma2_term_sum<T, A, S, 4>(ya, yb, yg, v, noa, nob, a_eq_b, nJab, xJab, kS2); break;
      default:
    for (int al = 0; al < noa; ++al) {
      const T yaa = (T)ya[al], yba = (T)yb[al];
      for (int be = a_eq_b ? al + nJab : 0; be < nob; ++be, v += nog) {
        T x2 = yaa * (T)yb[be];
        if (a_eq_b) {
          x2 += xJab * (T)ya[be] * yba;
          if (al == be) x2 /= kS2;
        }
        for (int gm = 0; gm < nog; ++gm) {
          if (nJgd && orb_g + gm == od) continue;
          x1 += (A)(v[gm] * x2 * (T)yg[gm]);
        }
      }
    }
    }
    x1 *= (A)x4;
    if (rb) rb[q] = (double)x1;
    else rows[(long)p * w + 2 + d] = (double)x1;
  }
}

typedef short sig_bf16x4 __attribute__((ext_vector_type(4)));
typedef float sig_f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ short sig_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// Pass 1 of one block side on the matrix cores (mixed precision): the
// coupling contraction m(p, d) = x4 sum_{(alpha beta), gamma} Vt(d; alpha beta,
// gamma) X2(alpha beta; na nb) y_g(ng; gamma) of Self_energy_ext_Op_c.cpp:
// 454-468 as the GEMM T[(d, gamma), (na, nb)] = Vt x X2 over the (alpha, beta)
// pairs -- operands FP32 (MF 1, v_mfma_f32_16x16x4f32) or BF16 (MF 2,
// v_mfma_f32_16x16x16bf16_1k), FP32 accumulation -- followed by the gamma sum
// with y_g in FP32.  X2 is the pair amplitude x2 of the reference (with the
// exchange term and the 1/sqrt2 of alpha == beta when a == b).  Operands come
// from LDS: the V table, the staged amplitudes, and two scratch tiles for X2
// and T (pitch kSigMf).  One wave.
template <int MF>
__device__ void sigma2_pass1_mfma(const float *xa, const float *xb, const float *xg, const float *vt, int noa,
                                  int nob, int nog, int nI, int npole, int a_eq_b, int nJab, int cA, int cB, int cG,
                                  float *x2s, float *ts, double *rows, int w, double *rb) {
  const int lane = threadIdx.x;
  const float xJab = (float)(1 - 2 * nJab), kS2 = (float)kSqrt2;
  const int npair = a_eq_b ? noa * (noa + 1) / 2 - nJab * noa : noa * nob;
  const int R = npole / cG, Mr = nI * nog;
  auto frag = [&](int r, int &na, int &nb) {
    if (a_eq_b) {
      na = 0;
      int len = cA - nJab;
      while (r >= len) { r -= len; ++na; --len; }
      nb = na + nJab + r;
    } else {
      na = r / cB; nb = r % cB;
    }
  };
  for (int e = lane; e < npair * R; e += 64) {
    const int pi = e / R, r = e % R;
    int al, be, na, nb;
    if (a_eq_b) {
      int q = pi, len = noa - nJab;
      al = 0;
      while (q >= len) { q -= len; ++al; --len; }
      be = al + nJab + q;
    } else {
      al = pi / nob; be = pi % nob;
    }
    frag(r, na, nb);
    const float *ya = xa + na * noa, *yb = xb + nb * nob;
    float x2 = ya[al] * yb[be];
    if (a_eq_b) {
      x2 += xJab * ya[be] * yb[al];
      if (al == be) x2 /= kS2;
    }
    x2s[pi * kSigMf + r] = x2;
  }
  __syncthreads();
  const int l16 = lane & 15, lq = lane >> 4;
  for (int m0 = 0; m0 < Mr; m0 += 16)
    for (int n0 = 0; n0 < R; n0 += 16) {
      sig_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const int m = m0 + l16, n = n0 + l16;
      const int d = m / nog, g = m % nog;
      if constexpr (MF == 2) {
        for (int k0 = 0; k0 < npair; k0 += 16) {
          sig_bf16x4 a, b;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int k = k0 + 4 * lq + i;
            a[i] = sig_bf16((k < npair && m < Mr) ? vt[(d * npair + k) * nog + g] : 0.f);
            b[i] = sig_bf16((k < npair && n < R) ? x2s[k * kSigMf + n] : 0.f);
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
      } else {
        for (int k0 = 0; k0 < npair; k0 += 4) {
          const int k = k0 + lq;
          const float a = (k < npair && m < Mr) ? vt[(d * npair + k) * nog + g] : 0.f;
          const float b = (k < npair && n < R) ? x2s[k * kSigMf + n] : 0.f;
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int mr = m0 + 4 * lq + i;
        if (mr < Mr && n < R) ts[mr * kSigMf + n] = acc[i];
      }
    }
  __syncthreads();
  for (int q = lane; q < npole * nI; q += 64) {
    const int p = q / nI, d = q % nI, r = p / cG, ng = p % cG;
    int na, nb;
    frag(r, na, nb);
    float s = 0.f;
    for (int g = 0; g < nog; ++g) s += ts[(d * nog + g) * kSigMf + r] * xg[ng * nog + g];
    if (a_eq_b && na == nb) s /= kS2;
    if (rb) rb[q] = (double)s;
    else rows[(long)p * w + 2 + d] = (double)s;
  }
}

// The coupling build of Build_Dynamic_SlfEn_2ndOrder (Self_energy_ext_Op_c.cpp:
// 380-539), one wave per block, grid-stride over the compact list of
// non-empty blocks.  Per block the interaction part Vt(d; alpha beta, gamma)
// is gathered into LDS once; per side the fragment amplitudes the poles use
// are staged next to it, then pass 1 builds the couplings (one lane per
// (pole, d)) and pass 2 the energies and sums of squares (one lane per pole,
// over d in order).  Rows match the reference to the last bit.
// T: arithmetic of the contraction -- double (the reference's) or float
// (FP32 operands and products); A: type of the sums; MF: 0 on the vector
// units, 1 / 2 on the matrix cores with FP32 / BF16 operands (T = float;
// sides whose sizes exceed kSigMf, or whose amplitudes are not staged, take
// the vector path).  Rows are stored in FP64.
template <class T, class A, int MF>
__global__ __launch_bounds__(64) void k_sigma2(const InstDev *insts, const ChanDev *chans,
                                               const BlockDev *blocks, PropView P,
                                               const int *blk_fw_off, const int *blk_bk_off,
                                               const Sig2Work *work_list, const int *work_count,
                                               const long *row_fw_base, const long *row_bk_base,
                                               double *rows_fw, double *rows_bk, int vt_lo, int vt_cap,
                                               int c_cap) {
#pragma clang fp contract(off)
  // dynamic LDS: the block's V table (vt_cap entries, the batch's largest),
  // kSigStage staged amplitudes, kSigRowBuf pass-1 couplings, then the pair
  // rows of (alpha, beta) and (d, gamma) with the exchange signs of the latter
  extern __shared__ double sig_lds[];
  T *vt = (T *)sig_lds;
  T *stg = (T *)(sig_lds + vt_cap);
  double *rbuf = sig_lds + vt_cap + kSigStage;
  int *rowA = (int *)(rbuf + kSigRowBuf), *rowC = rowA + kMaxPairs, *sgnC = rowC + c_cap;
  // matrix-core scratch (MF > 0): X2 and T tiles, 16-byte aligned
  float *mfx = (float *)(((unsigned long)(sgnC + c_cap) + 15) & ~15ul), *mft = mfx + kSigMf * kSigMf;
  const int lane = threadIdx.x;
  const int total = *work_count, G = gridDim.x;
  // The wave's work items are t = blockIdx.x + k * G in order.  This launch
  // builds the blocks of one V-table size class (vt_lo, vt_cap] (vt_lo = -1
  // for the first class: blocks without pairs have nvt = 0): each lane reads
  // the class of one of the next 64 items, and the wave runs the ones of its
  // class, so items of other classes cost a 64th of a load each.
  for (int t0 = blockIdx.x; t0 < total; t0 += 64 * G) {
    const int tl = t0 + lane * G;
    const int nvl = tl < total ? work_list[tl].nvt : -2;
    unsigned long long pend = __ballot(tl < total && nvl > vt_lo && nvl <= vt_cap);
    while (pend) {
      const int t = t0 + (__ffsll((long long)pend) - 1) * G;
      pend &= pend - 1;
      const Sig2Work Wk = work_list[t];
      const BlockDev B = blocks[Wk.bidx];
      const ChanDev C = chans[Wk.gc];
      const InstDev I = insts[Wk.inst];
      const int cur = prop_cur(P, Wk.inst);
      // fragment counts: qp(a), qp(b), qh(g), qh(a), qh(b), qp(g)
      const int f0 = Wk.n01 & 0xffff, f1 = Wk.n01 >> 16, f2 = Wk.n23 & 0xffff, f3 = Wk.n23 >> 16;
      const int f4 = Wk.n45 & 0xffff, f5 = Wk.n45 >> 16;
      const int nI = C.nI, a = B.a, b = B.b, g = B.g, J = B.J, nJab = B.nJab, nJgd = B.nJgd;
      const int noa = I.no[a], nob = I.no[b], nog = I.no[g];
      const int ga = I.gc0 + a, gb = I.gc0 + b, gg = I.gc0 + g;
      const int oa0 = I.orb0[a], ob0 = I.orb0[b], og0 = I.orb0[g];
      // (alpha, beta) pairs in the reference order (beta >= alpha + nJab when a == b)
      const int npair = (a == b) ? noa * (noa + 1) / 2 - nJab * noa : noa * nob;
      const int nvt = nI * npair * nog, nC = nI * nog;
      const double x3 = sqrt((double)(J + J + 1) / (double)(1 + C.j2));
      // V table Vt(d; alpha beta, gamma) = <alpha beta; J|V|d gamma; J> as VppInt_t::get
      // returns it (Vpp_potential.cpp:557-607): the pair rows of (alpha, beta) and of
      // (d, gamma) first, then one load per entry.  (alpha, beta) never needs the
      // exchange swap; (d, gamma) does when d > gamma, with the phase
      // (-1)^((j_d + j_gamma)/2 - J + 1).
      const long pbase = (long)J * I.norb;
      const bool jok = J <= I.jmax;
      for (int pi = lane; pi < npair; pi += 64) {
        int al, be;
        if (a == b) {
          int r = pi, len = noa - nJab;
          al = 0;
          while (r >= len) { r -= len; ++al; --len; }
          be = al + nJab + r;
        } else {
          al = pi / nob; be = pi % nob;
        }
        rowA[pi] = jok ? I.pidx[(pbase + oa0 + al) * I.norb + ob0 + be] : -1;
      }
      const int phC = abs((C.j2 + I.j2[g]) / 2 - J + 1) & 1;
      for (int q = lane; q < nC; q += 64) {
        const int od = C.orb0 + q / nog, og = og0 + q % nog;
        const int lo = od < og ? od : og, hi = od < og ? og : od;
        rowC[q] = jok ? I.pidx[(pbase + lo) * I.norb + hi] : -1;
        sgnC[q] = od > og ? phC : 0;
      }
      const long voff_b = B.c2b >= 0 ? I.voff[B.c2b] : 0;
      const int np_b = B.c2b >= 0 ? I.npair[B.c2b] : 0;
      __syncthreads();
      for (int e = lane; e < nvt; e += 64) {
        const int ga_ = e % nog, pi = (e / nog) % npair, d = e / (nog * npair);
        const int od = C.orb0 + d, og = og0 + ga_;
        double val = 0.0;
        if (!(og == od && nJgd)) {
          const int pa = rowA[pi], q = d * nog + ga_, pc = rowC[q];
          double v = 0.0;
          if (pa >= 0 && pc >= 0 && (pa >> 16) == (pc >> 16)) {
            const int ca = pa >> 16;
            const long vo = ca == B.c2b ? voff_b : I.voff[ca];
            const int np = ca == B.c2b ? np_b : I.npair[ca];
            v = I.V[vo + (long)(pa & 0xffff) * np + (pc & 0xffff)];
            if (sgnC[q]) v = -v;
          }
          val = x3 * v;
          if (og == od) val *= kSqrt2;
        }
        vt[e] = (T)val;
      }
      for (int side = 0; side < 2; ++side) {
        // side 0: 2p1h from qp(a), qp(b), qh(g); side 1: 2h1p from qh(a), qh(b), qp(g)
        const int cA = side ? f3 : f0, cB = side ? f4 : f1, cG = side ? f5 : f2;
        // the quasiholes of a subshell are stored after its quasiparticles
        const int oA = side ? f0 : 0, oB = side ? f1 : 0, oG = side ? 0 : f5;
        int npole;
        if (a == b) { int nb2 = cA + 1; if (nJab == 1) nb2 -= 2; npole = (cA * nb2 / 2) * cG; }
        else npole = cA * cB * cG;
        if (npole <= 0) continue;
        const double *eA = P.e[cur] + P.e_off[ga] + oA, *eB = P.e[cur] + P.e_off[gb] + oB;
        const double *eG = P.e[cur] + P.e_off[gg] + oG;
        const double *xA = P.x[cur] + P.x_off[ga] + (long)oA * noa;
        const double *xB = P.x[cur] + P.x_off[gb] + (long)oB * nob;
        const double *xG = P.x[cur] + P.x_off[gg] + (long)oG * nog;
        const int row_off = side ? Wk.off_bk : Wk.off_fw;
        double *rows = (side ? rows_bk + row_bk_base[B.gc] : rows_fw + row_fw_base[B.gc]) + (long)row_off * (2 + nI);
        const int w = 2 + nI;
        // stage the amplitudes of the fragments this side uses (for a == b the
        // two orbits share one fragment list)
        const int nsa = cA * noa, nsb = (a == b) ? 0 : cB * nob, nsg = cG * nog;
        const bool staged = nsa + nsb + nsg <= kSigStage;
        if (staged) {
          for (int k = lane; k < nsa; k += 64) stg[k] = (T)xA[k];
          for (int k = lane; k < nsb; k += 64) stg[nsa + k] = (T)xB[k];
          for (int k = lane; k < nsg; k += 64) stg[nsa + nsb + k] = (T)xG[k];
        }
        __syncthreads();
        // pass-1 results stay in LDS for pass 2 when they fit
        double *rb = npole * nI <= kSigRowBuf ? rbuf : nullptr;
        const int npair_s = (a == b) ? noa * (noa + 1) / 2 - nJab * noa : noa * nob;
        if (MF > 0 && staged && npair_s <= kSigMf && npole / cG <= kSigMf && nI * nog <= kSigMf)
          sigma2_pass1_mfma<MF>((const float *)stg, (const float *)((a == b) ? stg : stg + nsa),
                                (const float *)(stg + nsa + nsb), (const float *)vt, noa, nob, nog, nI, npole, a == b,
                                nJab, cA, cB, cG, mfx, mft, rows, w, rb);
        else if (staged)
          sigma2_pass1<T, A, T, true>(stg, (a == b) ? stg : stg + nsa, stg + nsa + nsb, vt, noa, nob, nog, nI, npole,
                                a == b, nJab, nJgd, cA, cB, cG, og0, C.orb0, rows, w, rb);
        else
          sigma2_pass1<T, A, double, false>(xA, xB, xG, vt, noa, nob, nog, nI, npole, a == b, nJab, nJgd, cA, cB, cG,
                                     og0, C.orb0, rows, w, rb);
        __syncthreads();
        // pass 2: energy and sum of squares (over d in order) per pole
        for (int p = lane; p < npole; p += 64) {
          int na, nb, ng, r = p / cG;
          ng = p % cG;
          if (a == b) {
            na = 0;
            int len = cA - nJab;
            while (r >= len) { r -= len; ++na; --len; }
            nb = na + nJab + r;
          } else {
            na = r / cB; nb = r % cB;
          }
          double *row = rows + (long)p * w;
          row[0] = eA[na] + eB[nb] - eG[ng];
          double sum2 = 0.0;
          if (rb) {
            for (int d = 0; d < nI; ++d) {
              const double m = rb[p * nI + d];
              row[2 + d] = m;
              sum2 += m * m;
            }
          } else {
            for (int d = 0; d < nI; ++d) sum2 += row[2 + d] * row[2 + d];
          }
          row[1] = sum2;
        }
        __syncthreads();
      }
    }
  }
}

#define BCD_SIGMA2_INST(T, A, MF)                                                                           \
  template __global__ void k_sigma2<T, A, MF>(const InstDev *, const ChanDev *, const BlockDev *, PropView, const int *, \
                                          const int *, const Sig2Work *, const int *, const long *, const long *,  \
                                          double *, double *, int, int, int);
BCD_SIGMA2_INST(double, double, 0)
BCD_SIGMA2_INST(float, double, 0)
BCD_SIGMA2_INST(float, double, 1)
BCD_SIGMA2_INST(float, double, 2)
#undef BCD_SIGMA2_INST

// ------------------------------------------------------------- Lanczos

// Pivots and iteration counts of every channel from the current (HF)
// propagator: quasiholes then quasiparticles by ascending energy with Z >= 60%
// (SpProp_t::DysPivots, Sp_Propagator.cpp:1225-1289), rounded as the pivots
// file carries them when piv_round is set, the instance's lanc_n (ChanDev)
// iterations per pivot (kind 0) or lanc_n vectors per side split over the
// pivots (kind 1),
// rescaled as Reduce_SelfEn_Lncz does when they exceed the pole count
// (Self_energy_ext_Op_e.cpp:147-220).  Writes one LancJob per side.
// Poles of the same fragments (na, nb, ng) of one (a, b, g) triple in the
// different J blocks share one energy eA + eB - eG: diag(E) is degenerate
// on each such group.  Every Krylov vector restricted to a group stays in the
// span of the group's coupling rows, so only the Gram matrix S = sum_p m_p m_p^T
// of the group enters the Lanczos recursion (Lanczos.cpp:101-241: its inner
// products, E-weighted products and coupling contractions).  A group of k >
// nI rows is replaced by the nI rows of R = L^T, S = L L^T (Cholesky), which
// gives the same reduced rows in exact arithmetic; groups of k <= nI rows are
// copied.  One block per (channel, side), one thread per triple: count the
// triple's output rows, scan, write them.  Rows come out triple by triple,
// fragments in order, each group's rows in J order.  Blocks take the (channel,
// side) jobs 2 * gc + side in the given order (the host passes them largest
// first, so that the longest blocks do not start last), or in index order.
__global__ __launch_bounds__(256) void k_merge_poles(const ChanDev *chans, int n_gc, ChanState *cs,
                                                    const BlockDev *blocks, const int *trip_start,
                                                    const int *trip_blk, PropView P, const InstDev *insts,
                                                    const int *blk_fw_off, const int *blk_bk_off,
                                                    const long *row_fw_base, const long *row_bk_base,
                                                    const double *rows_fw, const double *rows_bk, double *crows_fw,
                                                    double *crows_bk, const int *order) {
  // the chunk's non-empty triples: J blocks, a == b, fragment counts, tuples
  // (inclusive prefix), diagonal-group size and output base, and per J block
  // its first row and nJab as (row << 1) | nJab
  __shared__ int q_nb[256], q_eq[256], q_cA[256], q_cB[256], q_cG[256], q_tp[256], q_n0[256], q_base[256];
  __shared__ int q_jb[kMergeMaxJ][256];
  __shared__ int wtot[3][4];
  if ((int)blockIdx.x >= 2 * n_gc || !crows_fw) return;
  const int job = order ? order[blockIdx.x] : blockIdx.x;
  const int gc = job >> 1, side = job & 1, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const ChanDev C = chans[gc];
  if (!C.merge) {
    // a triple with more J blocks than the LDS tables hold: the rows stay
    // as built (no merged count, so the Lanczos job reads the built rows)
    if (tid == 0) (side ? cs[gc].cbk : cs[gc].cfw) = 0;
    return;
  }
  const InstDev I = insts[C.inst];
  const int cur = prop_cur(P, C.inst), nI = C.nI, w = 2 + nI;
  const double *rows = side ? rows_bk + row_bk_base[gc] : rows_fw + row_fw_base[gc];
  double *out = side ? crows_bk + row_bk_base[gc] : crows_fw + row_fw_base[gc];
  const int *boff = side ? blk_bk_off : blk_fw_off;
  auto rows_of = [&](int k) { return (k > nI && nI <= kMergeMaxNI) ? nI : k; };
  int base = 0;
  for (int t0 = 0; t0 < C.ntrip; t0 += 256) {
    // pass 1, one thread per triple (na <= nb when a == b; each J block keeps
    // the tuples with nb >= na + nJab)
    int s0 = 0, nblk = 0, eq = 0, cA = 0, cB = 0, cG = 0, T = 0, n0 = 0, cnt = 0;
    if (t0 + tid < C.ntrip) {
      s0 = trip_start[C.trip_off + t0 + tid];
      nblk = trip_start[C.trip_off + t0 + tid + 1] - s0;
      const BlockDev B0 = blocks[trip_blk[s0]];
      eq = B0.a == B0.b;
      const int ga = I.gc0 + B0.a, gb = I.gc0 + B0.b, gg = I.gc0 + B0.g;
      cA = side ? P.nqh[cur][ga] : P.nqp[cur][ga];
      cB = side ? P.nqh[cur][gb] : P.nqp[cur][gb];
      cG = side ? P.nqp[cur][gg] : P.nqh[cur][gg];
      T = (eq ? cA * (cA + 1) / 2 : cA * cB) * cG;
      if (T > 0) {
        for (int x = 0; x < nblk; ++x) n0 += blocks[trip_blk[s0 + x]].nJab == 0;
        cnt = eq ? cA * cG * rows_of(n0) + (T - cA * cG) * rows_of(nblk) : T * rows_of(nblk);
      }
    }
    // compact the non-empty triples; scan their tuples and output rows
    const bool live = T > 0;
    const unsigned long long m = __ballot(live);
    int iT = T, iC = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const int y1 = __shfl_up(iT, o, 64), y2 = __shfl_up(iC, o, 64);
      if (lane >= o) { iT += y1; iC += y2; }
    }
    if (lane == 63) { wtot[0][wv] = iT; wtot[1][wv] = iC; wtot[2][wv] = __popcll(m); }
    __syncthreads();
    int pT = 0, pC = 0, pL = 0, nT = 0, nC = 0, nL = 0;
    for (int v = 0; v < 4; ++v) {
      if (v < wv) { pT += wtot[0][v]; pC += wtot[1][v]; pL += wtot[2][v]; }
      nT += wtot[0][v]; nC += wtot[1][v]; nL += wtot[2][v];
    }
    if (live) {
      const int k = pL + __popcll(m & ((1ull << lane) - 1ull));
      q_nb[k] = min(nblk, kMergeMaxJ); q_eq[k] = eq; q_cA[k] = cA; q_cB[k] = cB; q_cG[k] = cG;
      q_tp[k] = pT + iT; q_n0[k] = n0; q_base[k] = base + pC + iC - cnt;
      for (int x = 0; x < nblk && x < kMergeMaxJ; ++x) {
        const int bx = trip_blk[s0 + x];
        q_jb[x][k] = (boff[bx] << 1) | (blocks[bx].nJab & 1);
      }
    }
    __syncthreads();
    // pass 2, one thread per fragment tuple of the chunk's live triples
    for (int f = tid; f < nT; f += 256) {
      int lo = 0, hi = nL - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (q_tp[mid] > f) hi = mid; else lo = mid + 1;
      }
      const int k_ = lo, u = f - (q_tp[k_] - (q_eq[k_] ? q_cA[k_] * (q_cA[k_] + 1) / 2 : q_cA[k_] * q_cB[k_]) * q_cG[k_]);
      const int a_eq_b = q_eq[k_], tA = q_cA[k_], tB = q_cB[k_], tG = q_cG[k_], nb_ = q_nb[k_], tn0 = q_n0[k_];
      int q = u / tG, na = 0, nb;
      const int ng = u % tG;
      if (a_eq_b) {
        int len = tA;
        while (q >= len) { q -= len; ++na; --len; }
        nb = na + q;
      } else {
        na = q / tB; nb = q % tB;
      }
      const bool diag = a_eq_b && na == nb;
      const int kk = diag ? tn0 : nb_;
      if (kk == 0) continue;
      const int rn0 = rows_of(tn0), rnb = rows_of(nb_);
      // rows of the tuples before this one: the pairs before (na, nb) times
      // tG tuples each, plus the earlier ng of this pair
      const int pq = u / tG, D = a_eq_b ? na + (nb > na ? 1 : 0) : 0;
      int pos = q_base[k_] + (D * rn0 + (pq - D) * rnb) * tG + ng * (diag ? rn0 : rnb);
      auto row_of = [&](int x, int nJ) -> const double * {
        const int pr = a_eq_b ? na * (tA - nJ) - na * (na - 1) / 2 + (nb - na - nJ) : na * tB + nb;
        return rows + (long)((q_jb[x][k_] >> 1) + pr * tG + ng) * w;
      };
      if (!(kk > nI && nI <= kMergeMaxNI)) {
        for (int x = 0; x < nb_; ++x) {
          const int nJ = q_jb[x][k_] & 1;
          if (a_eq_b && nb < na + nJ) continue;
          const double *src = row_of(x, nJ);
          for (int c = 0; c < w; ++c) out[(long)pos * w + c] = src[c];
          ++pos;
        }
        continue;
      }
      double S_[kMergeMaxNI][kMergeMaxNI] = {}, L[kMergeMaxNI][kMergeMaxNI] = {};
      double e = 0.0;
      for (int x = 0; x < nb_; ++x) {
        const int nJ = q_jb[x][k_] & 1;
        if (a_eq_b && nb < na + nJ) continue;
        const double *src = row_of(x, nJ);
        e = src[0];
#pragma unroll
        for (int ii = 0; ii < kMergeMaxNI; ++ii)
#pragma unroll
          for (int jj = 0; jj <= ii; ++jj)
            if (ii < nI) S_[ii][jj] += src[2 + ii] * src[2 + jj];
      }
      double tr = 0.0;
#pragma unroll
      for (int ii = 0; ii < kMergeMaxNI; ++ii) if (ii < nI) tr += S_[ii][ii];
#pragma unroll
      for (int jj = 0; jj < kMergeMaxNI; ++jj) {
        if (jj >= nI) continue;
        double d = S_[jj][jj];
#pragma unroll
        for (int mm = 0; mm < jj; ++mm) d -= L[jj][mm] * L[jj][mm];
        L[jj][jj] = d > 1e-28 * tr ? sqrt(d) : 0.0;
#pragma unroll
        for (int ii = 0; ii < kMergeMaxNI; ++ii) {
          if (ii <= jj || ii >= nI) continue;
          double v = S_[ii][jj];
#pragma unroll
          for (int mm = 0; mm < jj; ++mm) v -= L[ii][mm] * L[jj][mm];
          L[ii][jj] = L[jj][jj] > 0.0 ? v / L[jj][jj] : 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < kMergeMaxNI; ++r) {
        if (r >= nI) continue;
        double *dst = out + (long)(pos + r) * w;
        double s2 = 0.0;
        dst[0] = e;
#pragma unroll
        for (int d = 0; d < kMergeMaxNI; ++d) {
          if (d >= nI) continue;
          const double mv = d >= r ? L[d][r] : 0.0;
          dst[2 + d] = mv;
          s2 += mv * mv;
        }
        dst[1] = s2;
      }
    }
    base += nC;
    __syncthreads();
  }
  if (tid == 0) (side ? cs[gc].cbk : cs[gc].cfw) = base;
}

__global__ void k_lanczos_prep(const ChanDev *chans, int gc_begin, int gc_end, PropView P,
                               ChanState *cs, double *piv, int lanc_kind, int piv_round,
                               const long *row_fw_base, const long *row_bk_base, double *rows_fw,
                               double *rows_bk, double *crows_fw, double *crows_bk, double *lrows, double *lws,
                               LancJob *jobs) {
  const int gc = gc_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= gc_end) return;
  const ChanDev C = chans[gc];
  ChanState &S = cs[gc];
  const int cur = prop_cur(P, C.inst), nI = C.nI, lanc_n = C.lanc_n;
  const int nqp = P.nqp[cur][gc], nqh = P.nqh[cur][gc];
  const double *e = P.e[cur] + P.e_off[gc];
  const double *x = P.x[cur] + P.x_off[gc];
  double *pv = piv + C.piv_off;
  int npv = 0;
  for (int side = 0; side < 2 && npv < kMaxPiv; ++side) {
    const int cnt = side ? nqp : nqh, off = side ? 0 : nqp;
    double x2 = -1.0e10;
    int pick = cnt > 0 ? 1 : -100;
    while (pick >= 0 && npv < kMaxPiv) {
      double x1 = 1.0e10;
      pick = -100;
      for (int j = 0; j < cnt; ++j) {
        const double ej = e[off + j];
        double z = 0.0;
        for (int k = 0; k < nI; ++k) z += x[(long)(off + j) * nI + k] * x[(long)(off + j) * nI + k];
        if (ej < x1 && ej > x2 && 100. * z >= 60.0) { x1 = ej; pick = j; }
      }
      if (pick >= 0) {
        x2 = x1;
        for (int k = 0; k < nI; ++k) {
          const double xk = x[(long)(off + pick) * nI + k];
          pv[npv * nI + k] = piv_round ? dys_round(xk) : xk;
        }
        ++npv;
      }
    }
  }
  S.npiv_tot = npv;
  S.den = npv + 1 > 2 ? npv + 1 : 2;
  S.err = 0;
  for (int side = 0; side < 2; ++side) {
    int *itr = side ? S.itr_bk : S.itr_fw;
    int *ipv = side ? S.ipiv_bk : S.ipiv_fw;
    for (int k = 0; k < kMaxPiv + 2; ++k) ipv[k] = -100;
    for (int k = 0; k < npv; ++k) itr[k] = lanc_kind == 0 ? lanc_n : lanc_n / npv + (k < lanc_n % npv ? 1 : 0);
    int tot = 0;
    for (int k = 0; k < npv; ++k) if (itr[k] > 0) tot += itr[k];
    // the J-merged rows span the same Krylov spaces; they are used when
    // they still hold every requested iteration, so the reduction keeps the
    // reference's shape (no saturation that the full rows would not reach)
    const int Mfull = side ? S.mbk : S.mfw, Mc = crows_fw ? (side ? S.cbk : S.cfw) : 0;
    const bool merged = Mc > 0 && Mc >= tot;
    const int M = merged ? Mc : Mfull;
    (side ? S.lmbk : S.lmfw) = 0;
    int nl = tot < M ? tot : M;
    LancJob &Jb = jobs[2 * gc + side];
    Jb.nLanc = 0;
    Jb.inst = -1;
    int *lout = side ? &S.lbk : &S.lfw;
    *lout = 0;
    if (M < 1 || nl < 1) { ipv[0] = 0; continue; }
    if (tot > M) {
      const double xs = (double)M / (double)tot;
      for (int k = 0; k < npv; ++k) itr[k] = (int)(xs * itr[k] + 1.0);
    }
    int nLanc = 0;
    for (int k = 0; k < npv; ++k) if (itr[k] > 0) nLanc += itr[k];
    if (nLanc > M) nLanc = M;
    const int cap = side ? C.lbk_cap : C.lfw_cap;
    if (nLanc > (nl > 1 ? nl : 1) || nLanc > cap) { S.err = 1; continue; }
    Jb.Min = merged ? (side ? crows_bk + row_bk_base[gc] : crows_fw + row_fw_base[gc])
                    : (side ? rows_bk + row_bk_base[gc] : rows_fw + row_fw_base[gc]);
    (side ? S.lmbk : S.lmfw) = M;
    Jb.Tin = 2 + nI; Jb.NA_in = 2; Jb.NB = nI; Jb.M = M;
    Jb.piv = pv; Jb.nt_pvts = npv; Jb.nLanc = nLanc; Jb.itr = itr;
    Jb.Mout = lrows + (side ? C.lbk_off : C.lfw_off);
    Jb.Tout = S.den + nI; Jb.NA_out = S.den;
    Jb.i_piv_out = ipv; Jb.n_out = lout;
    Jb.V = lws + (side ? C.lws_bk_off : C.lws_fw_off);
  }
}

// Sum NV values over an NT-thread block with one barrier: every thread gets
// the totals.  Two LDS slots alternate (ph), so a slot is only rewritten after
// every thread has passed the barrier of the reduction in between.  NV is a
// compile-time count so that vals stays in registers.
template <int NT, int NV>
__device__ __forceinline__ void lanc_reduce(double *vals, double *red, int &ph) {
  if constexpr (NT == 64) {
    // one wave: the shuffle sums are the totals
#pragma unroll
    for (int k = 0; k < NV; ++k) vals[k] = wave_sum(vals[k]);
    return;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double *slot = red + ph * (NT / 64) * kLancRed;
#pragma unroll
  for (int k = 0; k < NV; ++k) vals[k] = wave_sum(vals[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) slot[w * kLancRed + k] = vals[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += slot[i * kLancRed + k];
    vals[k] = s;
  }
  ph ^= 1;
}

// Raw buffer access: one per-lane 32-bit offset plus a wave-uniform scalar
// offset, so a register tile of many elements needs no per-element address
// registers.  Offsets are in bytes; callers keep them inside the buffer.
constexpr int kBufWord3 = 0x00020000;   // gfx9-family buffer descriptor word 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, kBufWord3);
}

__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void buf_st(double x, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, voff, soff, 0);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Two neighbouring doubles in one 16-byte access.
__device__ __forceinline__ double2 buf_ld2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

__device__ __forceinline__ void buf_st2(double a, double b, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  double2 x;
  x.x = a;
  x.y = b;
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, x), r, voff, soff, 0);
}

// Multi-pivot Lanczos on diag(E) with full reorthogonalisation in the
// reference's order (Reduce_Mtx_Lncz, Lanczos.cpp:101-241): the current vector
// lives in registers, the earlier vectors in HBM, each thread touching only
// its own elements of them; each Gram-Schmidt projection, latest vector first,
// is one block-wide reduction of <v_i, v> and <v_i, v_i> followed by
// v -= (x1/x2) v_i.  A thread holds PER/2 pairs of neighbouring elements,
// pair q = tid + NT*t being elements 2q and 2q+1, so that every access moves
// 16 bytes per lane; vectors are stored with an even stride Ms, and the
// padding element of an odd M is kept at zero.
// E stays in registers for the whole job (read once from the input rows)
// where the tile leaves room for it (lanczos_ev_in_regs); otherwise it is
// kept as one more contiguous workspace vector, in the last E * v_last slot
// (the last pivot chain never fills it), and read where it is used.
// Behind the nLanc vectors the workspace holds the NB coupling columns of the
// input rows as contiguous vectors, and E * v_last of every finished pivot
// chain, so the per-step sums read contiguous vectors only.  With a stash
// (NT * PER doubles of LDS), each thread also keeps its own elements of the
// latest vector there, and the first projection of every step reads them
// from LDS instead of HBM; only the thread that wrote an element reads it, so
// no barrier is involved.
template <int NT, int PER>
__device__ __forceinline__ void lanczos_job(const LancJob &Jb, double *red, double *u, int *last, double *part,
                                            double2 *stash) {
  static_assert(PER % 2 == 0, "elements come in pairs");
  constexpr int PP = PER / 2;
  const int tid = threadIdx.x, M = Jb.M, NB = Jb.NB, Tin = Jb.Tin, NA = Jb.NA_in;
  const int Tout = Jb.Tout, NAo = Jb.NA_out, nLanc = Jb.nLanc, Ms = lanczos_stride(M);
  const double *Min = Jb.Min;
  for (long i = tid; i < (long)Tout * nLanc; i += NT) Jb.Mout[i] = 0.0;
  // pair t of a vector: byte offset 16*tid (lane) + 16*NT*t (scalar)
  const long nvec = lanczos_ws_vectors(nLanc, NB);
  const __amdgpu_buffer_rsrc_t rV = buf_rsrc(Jb.V, 8L * nvec * Ms), rM = buf_rsrc(Min, 8L * M * Tin);
  const int vo = 16 * tid, mo = 16 * tid * Tin;
  const int co = 8 * nLanc * Ms, wo = co + 8 * NB * Ms;   // couplings, E * v_last
  constexpr bool kEvReg = lanczos_ev_in_regs(NT, PER);
  const int eo = wo + 8 * (kMaxPiv - 1) * Ms;               // E when not in registers
  double ev[kEvReg ? PER : 2];
#pragma unroll
  for (int t = 0; t < PP; ++t) {
    const int p = 2 * (tid + NT * t), ms = 16 * NT * t * Tin;
    double e0 = 0.0, e1 = 0.0;
    if (p < M) {
      const bool h1 = p + 1 < M;
      if (h1) e1 = buf_ld(rM, mo, ms + 8 * Tin);
      e0 = buf_ld(rM, mo, ms);
      if constexpr (!kEvReg) buf_st2(e0, e1, rV, vo, eo + 16 * NT * t);
      for (int l = 0; l < NB; ++l) {
        double c1 = 0.0;
        if (h1) c1 = buf_ld(rM, mo, ms + 8 * (Tin + NA + l));
        buf_st2(buf_ld(rM, mo, ms + 8 * (NA + l)), c1, rV, vo, co + 8 * l * Ms + 16 * NT * t);
      }
    }
    if constexpr (kEvReg) {
      ev[2 * t] = e0;
      ev[2 * t + 1] = e1;
    }
  }
  // E at pair t (0 past the end)
  auto evp = [&](int t) -> double2 {
    double2 r;
    if constexpr (kEvReg) {
      r.x = ev[2 * t];
      r.y = ev[2 * t + 1];
    } else {
      r.x = r.y = 0.0;
      if (2 * (tid + NT * t) < M) r = buf_ld2(rV, vo, eo + 16 * NT * t);
    }
    return r;
  };
  double v[PER], vi[PER];
  int i_pvt = -1, n_pvt = -1, j_next = 0, ph = 0;
  int status = 0;
  for (int j = 0; j < nLanc; ++j) {
    const bool fresh = (j == j_next);
    if (fresh) {
      ++i_pvt;
      while (i_pvt < Jb.nt_pvts && Jb.itr[i_pvt] < 1) ++i_pvt;
      if (i_pvt >= Jb.nt_pvts) { status = -100; break; }
      ++n_pvt;
      j_next += Jb.itr[i_pvt];
      if (tid == 0) {
        last[n_pvt] = j_next - 1;
        double x1 = 0.0;
        for (int l = 0; l < NB; ++l) x1 += Jb.piv[i_pvt * NB + l] * Jb.piv[i_pvt * NB + l];
        x1 = sqrt(x1);
        for (int l = 0; l < NB; ++l) u[l] = Jb.piv[i_pvt * NB + l] / x1;
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < PP; ++t) {
        const int p = 2 * (tid + NT * t);
        double a0 = 0.0, a1 = 0.0;
        if (p < M)
          for (int l = 0; l < NB; ++l) {
            const double2 c = buf_ld2(rV, vo, co + 8 * l * Ms + 16 * NT * t);
            a0 += c.x * u[l];
            a1 += c.y * u[l];
          }
        v[2 * t] = a0;
        v[2 * t + 1] = a1;
      }
    } else {
#pragma unroll
      for (int t = 0; t < PP; ++t) {
        const double2 e = evp(t);
        v[2 * t] *= e.x;
        v[2 * t + 1] *= e.y;
      }
    }
  // Gram-Schmidt against every earlier vector, latest first (Lanczos.cpp:164-173)
    for (int i = j - 1; i >= 0; --i) {
      const int io = 8 * i * Ms;
      const bool from_lds = stash && i == j - 1;
      double s[2] = {0.0, 0.0};
#pragma unroll
      for (int t = 0; t < PP; ++t) {
        const int p = 2 * (tid + NT * t);
        double2 x;
        x.x = x.y = 0.0;
        if (from_lds) x = stash[tid + NT * t];
        else if (p < M) x = buf_ld2(rV, vo, io + 16 * NT * t);
        vi[2 * t] = x.x;
        vi[2 * t + 1] = x.y;
        s[0] += x.x * v[2 * t];
        s[1] += x.x * x.x;
        s[0] += x.y * v[2 * t + 1];
        s[1] += x.y * x.y;
      }
      lanc_reduce<NT, 2>(s, red, ph);
      const double c = s[0] / s[1];
#pragma unroll
      for (int t = 0; t < PER; ++t) v[t] -= c * vi[t];
    }
    // a, b, g_i and the couplings m^L (Lanczos.cpp:181-216).  Values k:
    // 0 sum E v^2, 1 sum v^2, 2..2+NB the couplings to the orbits, then the
    // g_i of the pivots; one value at a time, each reduced over its wave
    // into part[k][wave], then summed over the waves after one barrier.
    const int nred = 2 + NB + n_pvt, lane = tid & 63, wv = tid >> 6;
    for (int k = 0; k < nred; ++k) {
      double acc = 0.0;
      if (k == 0) {
#pragma unroll
        for (int t = 0; t < PP; ++t) {
          const double2 e = evp(t);
          acc += (e.x * v[2 * t]) * v[2 * t];
          acc += (e.y * v[2 * t + 1]) * v[2 * t + 1];
        }
      } else if (k == 1) {
#pragma unroll
        for (int t = 0; t < PER; ++t) acc += v[t] * v[t];
      } else {
        const int lo = k < 2 + NB ? co + 8 * (k - 2) * Ms : wo + 8 * (k - 2 - NB) * Ms;
#pragma unroll
        for (int t = 0; t < PP; ++t) {
          const int p = 2 * (tid + NT * t);
          if (p < M) {
            const double2 x = buf_ld2(rV, vo, lo + 16 * NT * t);
            acc += x.x * v[2 * t];
            acc += x.y * v[2 * t + 1];
          }
        }
      }
      acc = wave_sum(acc);
      if (lane == 0) part[k * (NT / 64) + wv] = acc;
    }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) { s0 += part[w]; s1 += part[NT / 64 + w]; }
    const double xa = s0 / s1, xb = sqrt(s1);
    double *row = Jb.Mout + (long)j * Tout;
    if (tid == 0) {
      row[0] = xa;
      row[1] = xb;
      for (int k = 2; k < nred; ++k) {
        double sk = 0.0;
        for (int w = 0; w < NT / 64; ++w) sk += part[k * (NT / 64) + w];
        if (k < 2 + NB) row[NAo + k - 2] = sk / xb;
        else row[k - NB] = sk / xb;
      }
      if (fresh) row[1] = row[1 + n_pvt];
    }
    if (n_pvt + 2 + NB > Tout) { status = -200; break; }
    // a chain that ends here leaves E * v_last for the g_i of later rows
    const bool chain_end = j == last[n_pvt] && j + 1 < nLanc;
#pragma unroll
    for (int t = 0; t < PP; ++t) {
      const int p = 2 * (tid + NT * t);
      v[2 * t] /= xb;
      v[2 * t + 1] /= xb;
      if (stash) {
        double2 x;
        x.x = v[2 * t];
        x.y = v[2 * t + 1];
        stash[tid + NT * t] = x;
      }
      if (p < M) {
        buf_st2(v[2 * t], v[2 * t + 1], rV, vo, 8 * j * Ms + 16 * NT * t);
        if (chain_end) {
          const double2 e = evp(t);
          buf_st2(v[2 * t] * e.x, v[2 * t + 1] * e.y, rV, vo, wo + 8 * n_pvt * Ms + 16 * NT * t);
        }
      }
    }
  }
  if (tid == 0) {
    if (status == 0) {
      ++n_pvt;
      Jb.i_piv_out[0] = n_pvt;
      Jb.i_piv_out[1] = 1;
      for (int i = 0; i < n_pvt; ++i) Jb.i_piv_out[2 + i] = last[i] + 1;
      if (last[n_pvt - 1] + 1 != nLanc && !(last[n_pvt - 1] + 1 > nLanc && M == nLanc)) status = -201;
    }
    *Jb.n_out = status == 0 ? nLanc : status;
  }
}

// One block of NT threads per job, jobs taken in the given order (longest
// first), PER elements of each vector per thread (the host keeps M <= NT*PER).
// Forcing more waves per SIMD (4 at 16 elements, 5 at 8) fits the tiles
// without spills but measured 2-5% slower: the loads already saturate HBM.
template <int NT, int PER>
__global__ __launch_bounds__(NT) void k_lanczos(const LancJob *jobs, const int *order, int njobs) {
  __shared__ double red[2 * (NT / 64) * kLancRed];
  __shared__ double u[kMaxNB];
  __shared__ double part[kLancRed * (NT / 64)];
  __shared__ int last[kMaxPiv + 1];
  extern __shared__ double2 lanc_stash[];
  if ((int)blockIdx.x >= njobs) return;
  const LancJob Jb = jobs[order ? order[blockIdx.x] : blockIdx.x];
  if (Jb.nLanc <= 0 || Jb.M > PER * NT) return;
  lanczos_job<NT, PER>(Jb, red, u, last, part, lanczos_stash_bytes(NT, PER) ? lanc_stash : nullptr);
}

template <int I>
static void launch_shape(int njobs, const LancJob *jobs, const int *order, hipStream_t stream) {
  constexpr int NT = kLancShapes[I].nt, PER = kLancShapes[I].per;
  constexpr int lds = lanczos_stash_bytes(NT, PER);
  // above 64 KiB the kernel needs the larger dynamic-LDS limit on the current
  // device (a failure shows up as a launch error)
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void *)k_lanczos<NT, PER>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  hipLaunchKernelGGL((k_lanczos<NT, PER>), dim3(njobs), dim3(NT), lds, stream, jobs, order, njobs);
}

void launch_lanczos(int shape, int njobs, const LancJob *jobs, const int *order, hipStream_t stream) {
  if (njobs <= 0) return;
  switch (shape) {
    case 0: launch_shape<0>(njobs, jobs, order, stream); break;
    case 1: launch_shape<1>(njobs, jobs, order, stream); break;
    case 2: launch_shape<2>(njobs, jobs, order, stream); break;
    case 3: launch_shape<3>(njobs, jobs, order, stream); break;
    case 4: launch_shape<4>(njobs, jobs, order, stream); break;
    case 5: launch_shape<5>(njobs, jobs, order, stream); break;
    case 6: launch_shape<6>(njobs, jobs, order, stream); break;
    case 7: launch_shape<7>(njobs, jobs, order, stream); break;
    case 8: launch_shape<8>(njobs, jobs, order, stream); break;
    case 9: launch_shape<9>(njobs, jobs, order, stream); break;
    case 10: launch_shape<10>(njobs, jobs, order, stream); break;
    case 11: launch_shape<11>(njobs, jobs, order, stream); break;
    case 12: launch_shape<12>(njobs, jobs, order, stream); break;
    case 13: launch_shape<13>(njobs, jobs, order, stream); break;
    case 14: launch_shape<14>(njobs, jobs, order, stream); break;
    case 15: launch_shape<15>(njobs, jobs, order, stream); break;
    case 16: launch_shape<16>(njobs, jobs, order, stream); break;
    case 17: launch_shape<17>(njobs, jobs, order, stream); break;
    case 18: launch_shape<18>(njobs, jobs, order, stream); break;
    case 19: launch_shape<19>(njobs, jobs, order, stream); break;
    case 20: launch_shape<20>(njobs, jobs, order, stream); break;
    default: launch_shape<21>(njobs, jobs, order, stream); break;
  }
}
static_assert(kNLancShapes == 22, "launch_lanczos covers every shape");

// ------------------------------------------------------------- Dyson

// Fill the Dyson job of every channel for this sweep: static part, pole rows
// (none, unreduced or Lanczos-reduced), Fermi energy of the sweep's input
// propagator and the output slots of the other propagator buffer.
__global__ void k_dyson_prep(const ChanDev *chans, int gc_begin, int gc_end, const ChanState *cs,
                             PropView P, int src, const double *h, const double *u1,
                             const long *row_fw_base, const long *row_bk_base, const double *rows_fw,
                             const double *rows_bk, const double *lrows, double *tri, double *mat,
                             double *kolt, DysonJob *jobs) {
  const int gc = gc_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= gc_end) return;
  const ChanDev C = chans[gc];
  DysonJob &J = jobs[gc - gc_begin];
  const InstState *S = P.is ? P.is + C.inst : nullptr;
  if (S && !S->active) { J.n = 0; return; }
  const ChanState &Q = cs[gc];
  const int cur = S ? S->cur : 0, nxt = 1 - cur, nI = C.nI;
  J.nI = nI; J.j2 = C.j2; J.ch = C.ch; J.inst = C.inst; J.poles = 0;
  J.h = h + C.h_off; J.u1 = u1 + C.u1_off;
  if (src == 0) {
    J.nfw = J.nbk = 0; J.den_fw = J.den_bk = 2; J.npiv_fw = J.npiv_bk = 0;
    J.rfw = J.rbk = nullptr; J.ipiv_fw = J.ipiv_bk = nullptr;
  } else if (src == 1) {
    J.nfw = Q.mfw; J.nbk = Q.mbk; J.den_fw = J.den_bk = 2; J.npiv_fw = J.npiv_bk = 0;
    J.rfw = rows_fw + row_fw_base[gc]; J.rbk = rows_bk + row_bk_base[gc];
    J.ipiv_fw = J.ipiv_bk = nullptr;
  } else {
    J.nfw = Q.lfw > 0 ? Q.lfw : 0; J.nbk = Q.lbk > 0 ? Q.lbk : 0;
    J.den_fw = J.den_bk = Q.den;
    J.npiv_fw = J.nfw > 0 ? Q.ipiv_fw[0] : 0; J.npiv_bk = J.nbk > 0 ? Q.ipiv_bk[0] : 0;
    J.rfw = lrows + C.lfw_off; J.rbk = lrows + C.lbk_off;
    J.ipiv_fw = Q.ipiv_fw; J.ipiv_bk = Q.ipiv_bk;
  }
  J.n = nI + J.nfw + J.nbk;
  if (J.n > C.cap) { J.n = 0; if (P.is) P.is[C.inst].err = 2; return; }
  J.efermi = chan_efermi(C, S);
  J.tri = tri + C.tri_off;
  J.mat = mat ? mat + C.mat_off : nullptr;
  J.out_e = P.e[nxt] + P.e_off[gc];
  J.out_x = P.x[nxt] + P.x_off[gc];
  J.out_nqp = P.nqp[nxt] + gc;
  J.out_nqh = P.nqh[nxt] + gc;
  J.kolt = kolt + 3L * gc;
}

// Assemble the Dyson matrix (DysonMatrix.cpp:269-390) and reduce it to
// tridiagonal form with Householder reflections (Golub & Van Loan, Alg.
// 8.3.1), tracking only the first nI rows of the orthogonal factor, which are
// all the quasiparticle amplitudes need.  Four waves per matrix; the matrix
// sits in LDS (IN_LDS) or in a global scratch slab.  In poles mode the same
// reduction diagonalises one Lanczos-reduced side for the arrowhead solver.
template <bool IN_LDS>
__global__ __launch_bounds__(256) void k_tridiag(const DysonJob *jobs, int njobs) {
  extern __shared__ double sm[];
  if ((int)blockIdx.x >= njobs) return;
  const DysonJob J = jobs[blockIdx.x];
  const int n = J.n;
  if (n <= 0 || (J.poles && J.path == 0)) return;
  const int nI = J.nI, tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
  double *A = IN_LDS ? sm : J.mat;
  double *aux = IN_LDS ? sm + (long)n * n : sm;
  double *v = aux, *p = v + n, *Qh = p + n, *qv = Qh + (long)nI * n, *red = qv + nI;
  double *td = J.tri, *te = td + n, *tq = te + n;
  const long ld = n;
  // poles mode: the matrix is one side's pole block alone and the tracked
  // rows start as its couplings, so they end as U^T M
  const bool poles = J.poles != 0;
  const int hd = poles ? 0 : nI;
  for (long i = tid; i < (long)n * n; i += blockDim.x) A[i] = 0.0;
  for (long i = tid; i < (long)nI * n; i += blockDim.x) Qh[i] = (!poles && (i / n) == (i % n)) ? 1.0 : 0.0;
  __syncthreads();
  if (!poles)
    for (int i = tid; i < nI * nI; i += blockDim.x) A[(i / nI) * ld + (i % nI)] = J.u1[i] + J.h[i];
  for (int side = 0; side < 2; ++side) {
    const int nd = side ? J.nbk : J.nfw, den = side ? J.den_bk : J.den_fw, T = den + nI;
    const int base = hd + (side ? J.nfw : 0);
    const double *rows = side ? J.rbk : J.rfw;
    for (int i = tid; i < nd; i += blockDim.x) {
      const double *row = rows + (long)i * T;
      A[(base + i) * ld + base + i] = row[0];
      for (int r = 0; r < nI; ++r) {
        if (poles) {
          Qh[(long)r * n + base + i] = row[den + r];
        } else {
          A[(base + i) * ld + r] = row[den + r];
          A[(long)r * ld + base + i] = row[den + r];
        }
      }
    }
  }
  __syncthreads();
  for (int side = 0; side < 2; ++side) {
    const int npiv = side ? J.npiv_bk : J.npiv_fw;
    if (npiv <= 0) continue;
    const int nd = side ? J.nbk : J.nfw, den = side ? J.den_bk : J.den_fw, T = den + nI;
    const int base = hd + (side ? J.nfw : 0);
    const double *rows = side ? J.rbk : J.rfw;
    const int *ipv = side ? J.ipiv_bk : J.ipiv_fw;
    for (int i = 1 + tid; i < nd; i += blockDim.x) {
      const double b = rows[(long)i * T + 1];
      A[(base + i) * ld + base + i - 1] = b;
      A[(base + i - 1) * ld + base + i] = b;
    }
    __syncthreads();
    // fishbone branches (DysonMatrix.cpp:303-333): branch nc starts at row
    // i_piv[nc+1] and couples its predecessor to every later row
    for (int nc = 1; nc <= npiv; ++nc) {
      const int st = ipv[nc + 1];
      if (st < 1 || st >= nd) break;
      for (int r = st + tid; r < nd; r += blockDim.x) {
        const double gv = rows[(long)r * T + 1 + nc];
        A[(base + r) * ld + base + st - 1] = gv;
        A[(base + st - 1) * ld + base + r] = gv;
      }
    }
    __syncthreads();
  }
  // Householder reduction
  for (int k = 0; k + 2 < n; ++k) {
    const int m0 = k + 1;
    double s[1] = {0.0};
    for (int j = m0 + tid; j < n; j += blockDim.x) { const double a = A[k * ld + j]; s[0] += a * a; }
    block_sum(s, 1, red);
    const double s2 = s[0], x0 = A[k * ld + m0], tail = s2 - x0 * x0;
    if (tail <= 0.0) { if (tid == 0) te[k] = x0; continue; }
    const double nrm = sqrt(s2), alpha = x0 > 0.0 ? -nrm : nrm, v0 = x0 - alpha;
    const double beta = 2.0 / (tail + v0 * v0);
    for (int j = m0 + tid; j < n; j += blockDim.x) v[j] = (j == m0) ? v0 : A[k * ld + j];
    if (tid == 0) te[k] = alpha;
    __syncthreads();
    for (int i = m0 + w; i < n; i += nw) {
      double acc = 0.0;
      for (int j = m0 + lane; j < n; j += 64) acc += A[i * ld + j] * v[j];
      acc = wave_sum(acc);
      if (lane == 0) p[i] = beta * acc;
    }
    for (int r = w; r < nI; r += nw) {
      double acc = 0.0;
      for (int j = m0 + lane; j < n; j += 64) acc += Qh[(long)r * n + j] * v[j];
      acc = wave_sum(acc);
      if (lane == 0) qv[r] = beta * acc;
    }
    __syncthreads();
    s[0] = 0.0;
    for (int i = m0 + tid; i < n; i += blockDim.x) s[0] += p[i] * v[i];
    block_sum(s, 1, red);
    const double K = 0.5 * beta * s[0];
    for (int i = m0 + tid; i < n; i += blockDim.x) p[i] -= K * v[i];
    __syncthreads();
    for (int i = m0 + w; i < n; i += nw) {
      const double vi = v[i], wi = p[i];
      for (int j = m0 + lane; j < n; j += 64) A[i * ld + j] -= vi * p[j] + wi * v[j];
    }
    for (int r = w; r < nI; r += nw) {
      const double q = qv[r];
      for (int j = m0 + lane; j < n; j += 64) Qh[(long)r * n + j] -= q * v[j];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += blockDim.x) {
    td[i] = A[i * ld + i];
    if (i == n - 2) te[i] = A[i * ld + i + 1];
    if (i == n - 1) te[i] = 0.0;
  }
  for (long i = tid; i < (long)nI * n; i += blockDim.x) tq[i] = Qh[i];
}

// k_tridiag for one Lanczos-reduced side in poles mode with n <= 128: one
// wave per matrix, lane j owning columns j and j + 64.  The symmetric block
// sits in LDS with an odd row stride and is only ever read along rows, so
// lanes touch consecutive words; the tracked rows (the couplings) stay in
// registers; every reduction is a wave shuffle over the lane's two columns,
// so no step waits on a block barrier.  A column past n holds zeros, and
// adding them leaves every sum of an n <= 64 side as with one column per
// lane.  Same reflections as k_tridiag (Golub & Van Loan, Alg. 8.3.1).  ld:
// LDS row stride (odd, >= n); dynamic LDS: ld * ld + 2 * kPoleWaveN doubles.
// NIMAX bounds the tracked rows held per lane (nI <= NIMAX).
template <int NIMAX>
__device__ __forceinline__ void tridiag_poles_body(const DysonJob &J, int ld, double *sm) {
  const int n = J.n, nI = J.nI, j = threadIdx.x;
  double *A = sm, *v = A + (long)ld * ld, *wv = v + kPoleWaveN;
  const int col[2] = {j, j + 64};
  double q[2][NIMAX];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int r = 0; r < NIMAX; ++r) q[c][r] = 0.0;
#pragma unroll
  for (int c = 0; c < 2; ++c)
    if (col[c] < n)
      for (int i = 0; i < n; ++i) A[(long)i * ld + col[c]] = 0.0;
  __syncthreads();
  // the fishbone block and its couplings (DysonMatrix.cpp:303-333)
  const int nd = J.nfw, den = J.den_fw, T = den + nI, npiv = J.npiv_fw;
  const double *rows = J.rfw;
#pragma unroll
  for (int c = 0; c < 2; ++c)
    if (col[c] < nd) {
      const double *row = rows + (long)col[c] * T;
      A[(long)col[c] * ld + col[c]] = row[0];
#pragma unroll
      for (int r = 0; r < NIMAX; ++r)
        if (r < nI) q[c][r] = row[den + r];
    }
  __syncthreads();
  if (npiv > 0) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (col[c] >= 1 && col[c] < nd) {
        const double b = rows[(long)col[c] * T + 1];
        A[(long)col[c] * ld + col[c] - 1] = b;
        A[(long)(col[c] - 1) * ld + col[c]] = b;
      }
    __syncthreads();
    for (int nc = 1; nc <= npiv; ++nc) {
      const int st = J.ipiv_fw[nc + 1];
      if (st < 1 || st >= nd) break;
#pragma unroll
      for (int c = 0; c < 2; ++c)
        if (col[c] >= st && col[c] < nd) {
          const double gv = rows[(long)col[c] * T + 1 + nc];
          A[(long)col[c] * ld + st - 1] = gv;
          A[(long)(st - 1) * ld + col[c]] = gv;
        }
      __syncthreads();
    }
  }
  double *td = J.tri, *te = td + n, *tq = te + n;
  for (int k = 0; k + 2 < n; ++k) {
    const int m0 = k + 1;
    double akj[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) akj[c] = col[c] > k && col[c] < n ? A[(long)k * ld + col[c]] : 0.0;
    const double s2 = wave_sum(akj[0] * akj[0] + akj[1] * akj[1]);
    const double x0 = A[(long)k * ld + m0], tail = s2 - x0 * x0;
    if (tail <= 0.0) {
      if (j == 0) te[k] = x0;
      continue;
    }
    const double nrm = sqrt(s2), alpha = x0 > 0.0 ? -nrm : nrm, v0 = x0 - alpha;
    const double beta = 2.0 / (tail + v0 * v0);
    double vj[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      vj[c] = col[c] == m0 ? v0 : akj[c];
      if (col[c] < kPoleWaveN) v[col[c]] = vj[c];
    }
    if (j == 0) te[k] = alpha;
    __syncthreads();
    // p = beta A v over rows > k (A symmetric: row i read along row j of A^T)
    double pj[2] = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (col[c] > k && col[c] < n)
        for (int i = m0; i < n; ++i) pj[c] += A[(long)i * ld + col[c]] * v[i];
      pj[c] *= beta;
    }
    double qv[NIMAX];
#pragma unroll
    for (int r = 0; r < NIMAX; ++r) qv[r] = r < nI ? beta * wave_sum(q[0][r] * vj[0] + q[1][r] * vj[1]) : 0.0;
    const double K = 0.5 * beta * wave_sum(pj[0] * vj[0] + pj[1] * vj[1]);
    double wj[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      wj[c] = pj[c] - K * vj[c];
      if (col[c] < kPoleWaveN) wv[col[c]] = wj[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (col[c] > k && col[c] < n)
        for (int i = m0; i < n; ++i) A[(long)i * ld + col[c]] -= v[i] * wj[c] + wv[i] * vj[c];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < NIMAX; ++r) q[c][r] -= qv[r] * vj[c];
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int jc = col[c];
    if (jc < n) {
      td[jc] = A[(long)jc * ld + jc];
      if (jc == n - 2) te[jc] = A[(long)jc * ld + jc + 1];
      if (jc == n - 1) te[jc] = 0.0;
#pragma unroll
      for (int r = 0; r < NIMAX; ++r)
        if (r < nI) tq[(long)r * n + jc] = q[c][r];
    }
  }
}

// One launch for every one-wave job: each runs the body of its own
// accumulator bound (4 or 8), so that the unrolled code that runs a job does
// not depend on the batch.
__global__ __launch_bounds__(64) void k_tridiag_poles(const DysonJob *jobs, const int *order, int njobs, int ld) {
  extern __shared__ double sm[];
  if ((int)blockIdx.x >= njobs) return;
  const DysonJob J = jobs[order ? order[blockIdx.x] : blockIdx.x];
  if (J.n <= 0 || J.n > kPoleWaveN || J.n > ld || J.nI > kQLGroupNI || !J.poles || J.path != 0) return;
  if (J.nI <= 4) tridiag_poles_body<4>(J, ld, sm);
  else tridiag_poles_body<8>(J, ld, sm);
}

template __global__ void k_tridiag<true>(const DysonJob *, int);
template __global__ void k_tridiag<false>(const DysonJob *, int);

// sqrt(a^2 + b^2) without the library hypot's scaling where no over- or
// underflow can occur (the arguments are uniform across the wave).
__device__ __forceinline__ double ql_hyp(double a, double b) {
  const double m = fmax(fabs(a), fabs(b));
  if (m > 1e150 || (m < 1e-150 && m > 0.0)) return hypot(a, b);
  return sqrt(a * a + b * b);
}

// The implicit QL iteration below for n <= 64: lane i holds d_i and e_i in
// registers, the rotation chain reads them with readlane and each lane keeps
// its own entry up to date; the first negligible e_m at or after lo comes
// from one ballot.  The tracked rows stay in LDS (lane r owns row r).  On
// return d holds the eigenvalues.  Returns 1 when an eigenvalue needs more
// than 60 iterations.
__device__ int ql_reg(double *d, double *e, double *Q, int n, int nI, int lane) {
  const double eps = 2.2204460492503131e-16;
  double dr = lane < n ? d[lane] : 0.0, er = lane < n ? e[lane] : 0.0;
  int fail = 0;
  for (int lo = 0; lo < n && !fail; ++lo) {
    for (int it = 0;; ++it) {
      const double dn = __shfl_down(dr, 1, 64);
      const bool small = lane >= lo && lane < n && (lane == n - 1 || fabs(er) <= eps * (fabs(dr) + fabs(dn)));
      const int m = __ffsll((long long)__ballot(small)) - 1;
      if (m == lo) break;
      if (it >= 60) { fail = 1; break; }
      const double dlo = lane_get(dr, lo), elo = lane_get(er, lo);
      double g = (lane_get(dr, lo + 1) - dlo) / (2.0 * elo);
      double r = ql_hyp(g, 1.0);
      double di1 = lane_get(dr, m);
      g = di1 - dlo + elo / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      bool defl = false;
      for (int i = m - 1; i >= lo; --i) {
        const double ei = lane_get(er, i), di = lane_get(dr, i);
        const double f = s * ei, b = c * ei;
        r = ql_hyp(f, g);
        if (lane == i + 1) er = r;
        if (r == 0.0) {
          if (lane == i + 1) dr = di1 - pp;
          if (lane == m) er = 0.0;
          defl = true;
          break;
        }
        s = f / r; c = g / r;
        g = di1 - pp;
        r = (di - g) * s + 2.0 * c * b;
        pp = s * r;
        if (lane == i + 1) dr = g + pp;
        g = c * r - b;
        if (lane < nI) {
          double *row = Q + (long)lane * n;
          const double t = row[i + 1];
          row[i + 1] = s * row[i] + c * t;
          row[i] = c * row[i] - s * t;
        }
        di1 = di;
      }
      if (defl) continue;
      if (lane == lo) { dr -= pp; er = g; }
      if (lane == m) er = 0.0;
    }
  }
  if (lane < n) d[lane] = dr;
  __syncthreads();
  return fail;
}

// Implicit QL with shifts on the tridiagonal (Bowdler, Martin, Reinsch and
// Wilkinson 1968), rotations applied to the nI tracked rows; then the split
// at the Fermi energy, the descending sort of quasiparticles and quasiholes
// (DysonMatrix.cpp:436-484) and the Calc_En partial sums (:542-594).  One wave
// per matrix: the rotation chain is uniform across lanes, lanes < nI own one
// tracked row each.
__global__ __launch_bounds__(64) void k_ql(const DysonJob *jobs, int njobs, int *status) {
  extern __shared__ double sm[];
  if ((int)blockIdx.x >= njobs) return;
  const DysonJob J = jobs[blockIdx.x];
  const int n = J.n;
  if (n <= 0 || (J.poles && J.path == 0)) return;
  const int nI = J.nI, lane = threadIdx.x;
  double *d = sm, *e = d + n, *Q = e + n;
  for (int i = lane; i < n; i += 64) { d[i] = J.tri[i]; e[i] = J.tri[n + i]; }
  for (long i = lane; i < (long)nI * n; i += 64) Q[i] = J.tri[2L * n + i];
  __syncthreads();
  const double eps = 2.2204460492503131e-16;
  int fail = 0;
  if (n <= 64) {
    fail = ql_reg(d, e, Q, n, nI, lane);
  } else
  for (int lo = 0; lo < n && !fail; ++lo) {
    for (int it = 0;; ++it) {
      int m = lo;
      for (; m + 1 < n; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= eps * dd) break;
      }
      if (m == lo) break;
      if (it >= 60) { fail = 1; break; }
      double g = (d[lo + 1] - d[lo]) / (2.0 * e[lo]);
      double r = hypot(g, 1.0);
      g = d[m] - d[lo] + e[lo] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      bool defl = false;
      // every lane runs the same scalar chain and stores the same values, so
      // each lane's own program order keeps d and e consistent
      for (int i = m - 1; i >= lo; --i) {
        const double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          defl = true;
          break;
        }
        s = f / r; c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int rr = lane; rr < nI; rr += 64) {
          double *row = Q + (long)rr * n;
          const double t = row[i + 1];
          row[i + 1] = s * row[i] + c * t;
          row[i] = c * row[i] - s * t;
        }
      }
      if (defl) continue;
      d[lo] -= pp; e[lo] = g; e[m] = 0.0;
    }
  }
  if (fail) { if (lane == 0) atomicAdd(status, 1); return; }
  __syncthreads();
  if (J.poles) {
    for (int k = lane; k < n; k += 64) {
      J.out_lam[k] = d[k];
      for (int r = 0; r < nI; ++r) J.out_ct[(long)r * J.ct_ld + k] = Q[(long)r * n + k];
    }
    return;
  }
  const double ef = J.efermi;
  int cnt = 0;
  for (int k = lane; k < n; k += 64) cnt += (d[k] <= ef);
  const int nqh = wave_isum(cnt), nqp = n - nqh;
  double kb[9];
  for (int i = 0; i < 9; ++i) kb[i] = 0.0;
  for (int k = lane; k < n; k += 64) {
    const double dk = d[k];
    int rank = 0;
    for (int j = 0; j < n; ++j) { const double dj = d[j]; rank += (dj < dk) || (dj == dk && j < k); }
    const int slot = dk > ef ? (n - 1 - rank) : nqp + (nqh - 1 - rank);
    J.out_e[slot] = dk;
    double x1 = 0.0, x2 = 0.0;
    for (int r = 0; r < nI; ++r) {
      const double xr = Q[(long)r * n + k];
      J.out_x[(long)slot * nI + r] = xr;
      x1 += xr * xr;
    }
    if (dk <= ef) {
      for (int r = 0; r < nI; ++r)
        for (int c = 0; c < nI; ++c) x2 += J.u1[r * nI + c] * Q[(long)r * n + k] * Q[(long)c * n + k];
      const int bk = x1 > 0.01 ? 0 : (x1 > 1.e-5 ? 1 : 2);
      kb[3 * bk] += x2; kb[3 * bk + 1] += x1 * dk; kb[3 * bk + 2] += x1;
    }
  }
  for (int i = 0; i < 9; ++i) kb[i] = wave_sum(kb[i]);
  if (lane == 0) {
    *J.out_nqp = nqp;
    *J.out_nqh = nqh;
    const double f = (double)J.j2 + 1.0;
    J.kolt[0] = (kb[0] + kb[3] + kb[6]) * f;
    J.kolt[1] = (kb[1] + kb[4] + kb[7]) * f;
    J.kolt[2] = (kb[2] + kb[5] + kb[8]) * f;
  }
}

// A group's negligibility mask: bit p (p < 128) in word p / 64, kept in two
// scalars so that no step indexes a local array.
struct NegMask {
  unsigned long long w0 = 0ull, w1 = 0ull;
  __device__ bool get(int p) const { return ((p < 64 ? w0 >> p : w1 >> (p - 64)) & 1ull) != 0ull; }
  __device__ void put(int p, bool v) {
    if (p < 64) w0 = v ? (w0 | (1ull << p)) : (w0 & ~(1ull << p));
    else w1 = v ? (w1 | (1ull << (p - 64))) : (w1 & ~(1ull << (p - 64)));
  }
  // first set (want) or clear bit at or after p < 128; 128 when there is none
  __device__ int find(int p, bool want) const {
    unsigned long long x0 = want ? w0 : ~w0, x1 = want ? w1 : ~w1;
    if (p < 64) {
      x0 &= ~0ull << p;
      if (x0) return __ffsll((long long)x0) - 1;
    } else {
      x1 &= ~0ull << (p - 64);
    }
    return x1 ? 64 + __ffsll((long long)x1) - 1 : 128;
  }
};

// The pole-block diagonalisation (poles mode) of k_ql for n <= kPoleWaveN and
// nI <= 8, eight matrices per wave: each group of eight lanes runs one
// matrix's implicit QL (the same iteration as k_ql) with d and e in its own
// LDS slice; lane r < nI of a group rotates tracked row r where the
// tridiagonalisation left it, in the job's buffer (tri + 2n), and finally
// copies it out.  Every lane of a group computes the same scalar chain, so a
// wave does the work of eight matrices in the instructions k_ql spends on
// one; with only d and e in LDS, sixteen waves fit a CU, which hides the
// chain's latency (the tracked rows in LDS allowed four).
// The block end m of k_ql's search (the first p >= lo with |e_p| <=
// eps (|d_p| + |d_p+1|), else n - 1) comes from a mask of that test per p,
// kept current as the rotations change d and e (each changed position is
// tested again once its three values are final), so no step scans e.  The
// next step's d, e and tracked-row entries are loaded while the current one
// computes (the row entry a step ahead hides its load), and the tracked
// row's intermediate entry stays in a register.  ld: LDS doubles per d or e
// array (the batch's largest side); qrows: the batch's largest nI.
__global__ __launch_bounds__(64) void k_ql_poles(const DysonJob *jobs, const int *order, int njobs, int *status, int ld,
                                                  int qrows) {
  extern __shared__ double sm[];
  const int lane = threadIdx.x, grp = lane >> 3, gl = lane & 7;
  const int job = blockIdx.x * 8 + grp;
  double *d = sm + (long)grp * 2 * ld, *e = d + ld;
  const double eps = 2.2204460492503131e-16;
  int n = 0, nI = 0;
  const DysonJob *J = job < njobs ? jobs + (order ? order[job] : job) : nullptr;
  if (J) { n = J->path == 0 ? J->n : 0; nI = J->nI; }
  if (n > ld || n > kPoleWaveN || nI > qrows || nI > kQLGroupNI) {
    // does not fit this kernel's slices (the host never launches such jobs here)
    if (gl == 0 && n > 0) atomicAdd(status, 1);
    n = 0;
  }
  for (int i = gl; i < n; i += 8) { d[i] = J->tri[i]; e[i] = J->tri[n + i]; }
  __syncthreads();
  auto negl = [&](double ep, double dp, double dq) { return fabs(ep) <= eps * (fabs(dp) + fabs(dq)); };
  // the mask; position n - 1 stands for k_ql's default block end
  NegMask nb;
  for (int b0 = 0; b0 < ld && b0 < kPoleWaveN; b0 += 8) {
    const int p = b0 + gl;
    const bool neg = p < n - 1 && negl(e[p], d[p], d[p + 1]);
    const unsigned long long bits = (__ballot(neg) >> (8 * grp)) & 0xffull;
    if (b0 < 64) nb.w0 |= bits << b0;
    else nb.w1 |= bits << (b0 - 64);
  }
  if (n > 0) nb.put(n - 1, true);
  const bool track = n > 0 && gl < nI;
  double *row = track ? J->tri + 2L * n + (long)gl * n : nullptr;
  // phase 0: find the block end m for lo; 1: one rotation of the chain; 2: done
  int lo = 0, it = 0, m = 0, i = 0, phase = n > 0 ? 0 : 2, fail = 0;
  double s = 1.0, c = 1.0, g = 0.0, pp = 0.0, di1 = 0.0;
  // this step's e_i, d_i and tracked entry, the tracked row's carried entry
  // (position i + 1) and d_{i+2} as the previous step left it
  double ce = 0.0, cd = 0.0, cq = 0.0, qc = 0.0, dab = 0.0;
  while (__ballot(phase != 2)) {
    if (phase == 0) {
      // positions whose e is negligible end a block of one: lo moves past
      // them (k_ql's m == lo case, one lo at a time)
      const int nlo = nb.get(lo) ? nb.find(lo, false) : lo;
      if (nlo != lo) { lo = nlo; it = 0; }
      if (lo >= n) {
        phase = 2;
      } else if (it >= 60) {
        fail = 1;
        phase = 2;
      } else if ((m = nb.find(lo, true)) >= n) {
        // cannot happen (bit n - 1 stays set); kept so that no index leaves the slice
        fail = 1;
        phase = 2;
      } else {
        const double dlo = d[lo], elo = e[lo];
        g = (d[lo + 1] - dlo) / (2.0 * elo);
        const double r = ql_hyp(g, 1.0);
        di1 = d[m];
        g = di1 - dlo + elo / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
        s = 1.0; c = 1.0; pp = 0.0;
        i = m - 1;
        ce = e[i]; cd = d[i];
        if (track) { cq = row[i]; qc = row[m]; }
        phase = 1;
      }
    } else if (phase == 1) {
      const double ei = ce, di = cd, qi = cq;
      if (i - 1 >= lo) {
        ce = e[i - 1]; cd = d[i - 1];
        if (track) cq = row[i - 1];
      }
      const double f = s * ei, b = c * ei;
      const double eh = ql_hyp(f, g);
      double r = eh;
      e[i + 1] = eh;
      if (r == 0.0) {
        const double dn = di1 - pp;
        d[i + 1] = dn;
        e[m] = 0.0;
        if (track) row[i + 1] = qc;
        if (i + 1 < m) nb.put(i + 1, negl(eh, dn, dab));
        nb.put(i, negl(ei, di, dn));
        ++it;
        phase = 0;
      } else {
        s = f / r; c = g / r;
        g = di1 - pp;
        r = (di - g) * s + 2.0 * c * b;
        pp = s * r;
        const double dn = g + pp;
        d[i + 1] = dn;
        g = c * r - b;
        if (track) {
          row[i + 1] = s * qi + c * qc;
          qc = c * qi - s * qc;
        }
        if (i + 1 < m) nb.put(i + 1, negl(eh, dn, dab));
        dab = dn;
        di1 = di;
        if (--i < lo) {
          const double dl = d[lo] - pp;
          d[lo] = dl;
          e[lo] = g;
          e[m] = 0.0;
          if (track) row[lo] = qc;
          nb.put(lo, negl(g, dl, dab));
          ++it;
          phase = 0;
        }
      }
    }
  }
  __syncthreads();
  if (!J || n == 0) return;
  if (fail) { if (gl == 0) atomicAdd(status, 1); return; }
  for (int k = gl; k < n; k += 8) J->out_lam[k] = d[k];
  if (track)
    for (int k = 0; k < n; ++k) J->out_ct[(long)gl * J->ct_ld + k] = row[k];
}

// The reference's heapsort of pole rows by column col (Sort_u_double2dim /
// Sort_d_double2dim, Utilities.cpp:25-94), one block: thread 0 sorts (key,
// row index) pairs with the reference's sift order (heap_perm) -- in LDS when
// lds_keys is set, else in the global scratch keys/idx -- then every thread
// gathers rows into out.  Equal keys end up where the reference puts them.
__global__ __launch_bounds__(256) void k_heap_rows(const double *rows, int M, int T, int col, int desc,
                                                   double *keys, int *idx, int lds_keys, double *out) {
  extern __shared__ double hs[];
  double *K = lds_keys ? hs : keys;
  int *X = lds_keys ? (int *)(hs + M) : idx;
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    K[i] = rows[(long)i * T + col];
    X[i] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0) heap_perm(K, X, M, desc);
  __syncthreads();
  for (long e = threadIdx.x; e < (long)M * T; e += blockDim.x) {
    const long r = e / T, c = e % T;
    out[e] = rows[(long)X[r] * T + c];
  }
}

// x as the pivots file carries it (dys_round), element-wise.
__global__ void k_pivot_round(const double *x, double *out, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = dys_round(x[i]);
}

// ------------------------------------------------------------- sweep control

// One block per instance.  init = 1: Fermi energies of the current propagator
// only.  Otherwise: compare the new propagator (buffer 1-cur) with the current
// one by storage index (quick_n_dirty_diff, Sp_Propagator.cpp:1189-1213; sum or
// max), add up the Koltun parts (Main.cpp:387-390), flip the buffers, stop the
// instance when converged or at ItrMax (Main.cpp:216, 451), and set the Fermi
// energies of the next sweep from the new propagator (Seek_EFermi,
// DysonMatrix.cpp:619-635).
__global__ __launch_bounds__(256) void k_finalize(const InstDev *insts, const ChanDev *chans,
                                                  PropView P, const ChanState *cs, const double *kolt,
                                                  int conv_kind, int init) {
  // per subshell: fragment counts of both buffers (prefix sums), the start of
  // its records and its charge, all in LDS so that the record loops below
  // are flat over the instance's records
  __shared__ int pq[2][kMaxSub + 1], ph[2][kMaxSub + 1], chg[kMaxSub];
  __shared__ long eo[kMaxSub];
  __shared__ double red[4 * 4];
  const int inst = blockIdx.x, tid = threadIdx.x;
  const InstDev I = insts[inst];
  InstState &S = P.is[inst];
  if (!init && !S.active) return;
  const int ns = I.nsubsh, cur = S.cur, nxt = 1 - cur;
  const int useb = init ? cur : nxt, bu = init ? 0 : 1;
  for (int t = tid; t < 2 * ns; t += blockDim.x) {
    const int b = t / ns, s = t % ns, buf = b == 0 ? cur : nxt;
    pq[b][s + 1] = P.nqp[buf][I.gc0 + s];
    ph[b][s + 1] = P.nqh[buf][I.gc0 + s];
  }
  for (int s = tid; s < ns; s += blockDim.x) {
    eo[s] = P.e_off[I.gc0 + s];
    chg[s] = I.ch[s] & 1;
  }
  __syncthreads();
  if (tid < 2) {
    int *q = pq[tid], *h = ph[tid];
    q[0] = 0; h[0] = 0;
    for (int s = 0; s < ns; ++s) { q[s + 1] += q[s]; h[s + 1] += h[s]; }
  }
  __syncthreads();
  // subshell of the j-th record of a prefix-summed count list
  auto owner = [&](const int *pre, int j) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (pre[mid + 1] > j) hi = mid; else lo = mid + 1;
    }
    return lo;
  };
  // Fermi energies from buffer useb (counts in LDS slot bu)
  {
    double mn[2] = {1e300, 1e300}, mx[2] = {-1e300, -1e300};
    const double *E = P.e[useb];
    const int totp = pq[bu][ns], toth = ph[bu][ns];
    for (int j = tid; j < totp + toth; j += blockDim.x) {
      const bool hole = j >= totp;
      const int jj = hole ? j - totp : j;
      const int s = owner(hole ? ph[bu] : pq[bu], jj);
      const double e = E[eo[s] + (hole ? pq[bu][s + 1] - pq[bu][s] + jj - ph[bu][s] : jj - pq[bu][s])];
      if (hole) {
        if (chg[s]) mx[1] = fmax(mx[1], e); else mx[0] = fmax(mx[0], e);
      } else {
        if (chg[s]) mn[1] = fmin(mn[1], e); else mn[0] = fmin(mn[0], e);
      }
    }
    double v4[4] = {-mn[0], -mn[1], mx[0], mx[1]};
    for (int k = 0; k < 4; ++k) v4[k] = wave_max(v4[k]);
    if ((tid & 63) == 0) for (int k = 0; k < 4; ++k) red[(tid >> 6) * 4 + k] = v4[k];
    __syncthreads();
    if (tid == 0) {
      for (int k = 0; k < 4; ++k) {
        double m = red[k];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, red[i * 4 + k]);
        red[k] = m;
      }
      // first stored quasiparticle / quasihole of the propagator as the start
      // values, as Seek_EFermi does
      double e0p = 0.0, e0h = 0.0;
      for (int s = 0; s < ns; ++s)
        if (pq[bu][s + 1] > pq[bu][s]) { e0p = E[eo[s]]; break; }
      for (int s = 0; s < ns; ++s)
        if (ph[bu][s + 1] > ph[bu][s]) { e0h = E[eo[s] + (pq[bu][s + 1] - pq[bu][s])]; break; }
      for (int c = 0; c < 2; ++c) {
        const double x1 = fmin(e0p + 100., -red[c]), x2 = fmax(e0h - 100., red[2 + c]);
        S.ef[c] = 0.5 * (x1 + x2);
      }
    }
    __syncthreads();
  }
  if (init) {
    // the host has set which instances the stage iterates
    if (tid == 0) {
      if (S.failed) S.active = 0;
      if (S.active) { S.its = 0; S.diff = 0.0; }
    }
    return;
  }
  // storage-index difference between buffers cur (old) and nxt (new)
  double acc[2] = {0.0, 0.0};
  for (int side = 0; side < 2; ++side) {
    const int tot_old = side ? ph[0][ns] : pq[0][ns], tot_new = side ? ph[1][ns] : pq[1][ns];
    const int tmax = tot_old > tot_new ? tot_old : tot_new;
    for (int j = tid; j < tmax; j += blockDim.x) {
      double ev[2] = {0.0, 0.0};
      for (int b = 0; b < 2; ++b) {
        const int *pre = side ? ph[b] : pq[b];
        if (j >= pre[ns]) continue;
        const int s = owner(pre, j);
        ev[b] = P.e[b == 0 ? cur : nxt][eo[s] + (side ? pq[b][s + 1] - pq[b][s] : 0) + (j - pre[s])];
      }
      const double dlt = fabs(ev[0] - ev[1]);
      acc[0] += dlt;
      acc[1] = fmax(acc[1], dlt);
    }
  }
  const double dsum = wave_sum(acc[0]), dmax = wave_max(acc[1]);
  if ((tid & 63) == 0) { red[(tid >> 6) * 4] = dsum; red[(tid >> 6) * 4 + 1] = dmax; }
  __syncthreads();
  // Koltun sums over the subshells: first wave, one subshell per lane
  double e1 = 0.0, es = 0.0, as = 0.0;
  if (tid < 64) {
    for (int s = tid; s < ns; s += 64) {
      const double *k3 = kolt + 3L * (I.gc0 + s);
      e1 += k3[0]; es += (k3[0] + k3[1]) / 2.0; as += k3[2];
    }
    e1 = wave_sum(e1); es = wave_sum(es); as = wave_sum(as);
  }
  if (tid == 0) {
    double sum = 0.0, mxd = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) { sum += red[i * 4]; mxd = fmax(mxd, red[i * 4 + 1]); }
    S.u1_tot = e1; S.e_kolt = es; S.a_tot = as;
    const double diff = conv_kind == 1 ? mxd : sum;
    S.diff = diff;
    S.its += 1;
    S.cur = nxt;
    if (diff < S.conv || (S.itr_max >= 0 && S.its > S.itr_max)) S.active = 0;
  }
}

}  // namespace bcd
