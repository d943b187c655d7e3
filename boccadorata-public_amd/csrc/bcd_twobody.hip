// bcd_twobody.hip -- ladder (gII_prop_t) and ring (Pol_prop_t) matrices on gfx950.
//
//   k_t2_zero   clears the extent Fill_*_mtx clears (gII_b.cpp:350-352, PolProp_b.cpp:310-312)
//   k_t2_sixj   (2J'+1) {ja jb J; jg jd J'} tables (Ang_momenta.cpp:150-182) per J
//   k_t2_amp    folded amplitude matrices T (one row per basis entry)
//   k_t2_pandya Pandya-transformed interaction per ring channel (PolProp_b.cpp:569-614)
//   k_t2_gemm   FP64 MFMA GEMMs G = T K and G T^T with the block epilogue
//   k_t2_diag   free-pole energies on the diagonal (gII_b.cpp:552-556, PolProp_b.cpp:542-556)
#include "bcd_hipsync.hpp"
#include "bcd_twobody.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "bcdor_amd.h"

namespace bcd {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kTK = 16, kTPad = 1;   // K step of the GEMMs
constexpr int kNBuf = 1;             // LDS stages of the GEMM K loop
constexpr int kBigTile = 128;   // block tile of tasks whose both dimensions reach it (else 64)
constexpr bool kUseBigTiles = false;
constexpr int kXcd = 8;          // XCDs of an MI355X
constexpr int kPatch = 8;        // tile patch edge kept on one XCD
constexpr int kBin = 60;          // Pascal table rows (covers 2j <= 2*27)
constexpr long kZeroChunk = 1 << 16;

// ------------------------------------------------------------------ kernels

__global__ void k_t2_zero(const ZeroTask *tasks, const int2 *chunks) {
  const int2 c = chunks[blockIdx.x];
  const ZeroTask t = tasks[c.x];
  const long lo = (long)c.y * kZeroChunk;
  const long hi = min(t.n, lo + kZeroChunk);
  for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) t.out[i] = 0.0;
}

// {a/2 b/2 c/2; d/2 e/2 f/2} in the binomial form of Ang_momenta.cpp:150-182.
__device__ double s6j_dev(const double *bn, int ja, int jb, int jc, int la, int lb, int lc) {
  auto tri = [](int a, int b, int c) {
    const int s1 = a + b - c, s2 = b + c - a, s3 = c + a - b;
    return s1 >= 0 && s2 >= 0 && s3 >= 0 && !((s1 | s2 | s3) & 1);
  };
  if (!tri(ja, jb, jc) || !tri(ja, lb, lc) || !tri(la, jb, lc) || !tri(la, lb, jc)) return 0.0;
  const int mt = (ja + jb + jc) / 2 + 1, ma = (ja + lb + lc) / 2 + 1;
  const int mb = (la + jb + lc) / 2 + 1, mc = (la + lb + jc) / 2 + 1;
  const int na = mt - ja, nb = mt - jb, nc = mt - jc, ka = ma - lc, kb = mb - lc, kc = mc - jc;
  auto B = [bn](int n, int k) { return bn[n * kBin + k]; };
  double norm = B(mt, ja + 1) * B(ja, nc - 1) /
                (B(ma, ja + 1) * B(ja, ka - 1) * B(mb, la + 1) * B(la, kb - 1) * B(mc, la + 1) * B(la, kc - 1));
  norm = sqrt(norm) / (la + 1.0);
  const int lo = max(max(mt, ma), max(mb, mc));
  const int hi = min(min(ma + na, mb + nb), mc + nc);
  double x = 0.0;
  for (int l = lo; l < hi; ++l) x = -x + B(l, mt) * B(na - 1, l - ma) * B(nb - 1, l - mb) * B(nc - 1, l - mc);
  return (double)(1 - 2 * (hi & 1)) * norm * x;
}

// tab[((slot * nj^4 + quad) * nJp + Jp)] = (2Jp+1) {ja jb J; jg jd Jp}, quad = ((a*nj+b)*nj+g)*nj+d
__global__ void k_t2_sixj(double *tab, const int *Js, int nslot, int nj, const int *jv, int nJp, const double *bn) {
  const long nq = (long)nj * nj * nj * nj;
  const long total = (long)nslot * nq * nJp;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int Jp = (int)(i % nJp);
    long q = (i / nJp) % nq;
    const int slot = (int)(i / (nq * nJp));
    const int d = (int)(q % nj); q /= nj;
    const int g = (int)(q % nj); q /= nj;
    const int b = (int)(q % nj); q /= nj;
    const int a = (int)q;
    tab[i] = (double)(2 * Jp + 1) * s6j_dev(bn, jv[a], jv[b], 2 * Js[slot], jv[g], jv[d], 2 * Jp);
  }
}

// One block per (task, row): T[r][p] for all pairs p of the task.
__global__ void k_t2_amp(const AmpTask *tasks, const int2 *rows, InstDev I, FragDev F) {
  const int2 tr = rows[blockIdx.x];
  const AmpTask t = tasks[tr.x];
  const int r = tr.y;
  const int u = t.u[r], w = t.w[r];
  const int su = F.sub[t.ukind][u], sw = F.sub[t.wkind][w];
  const double *xu = F.fq[t.ukind] + (long)u * F.nspb;
  const double *xw = F.fq[t.wkind] + (long)w * F.nspb;
  const int u0 = I.orb0[su], w0 = I.orb0[sw];
  double *out = t.T + (long)r * t.ld;
  for (int p = threadIdx.x; p < t.npair; p += blockDim.x) {
    const int pr = t.pair[p];
    const int a = pr >> 16, b = pr & 0xffff;
    const int sa = I.orb_sh[a], sb = I.orb_sh[b];
    double v = 0.0;
    if (t.kind == 0) {
      if (sa == su && sb == sw) {
        v = xu[a - u0] * xw[b - w0];
        if (a == b) v *= kSqrt2;
      }
      if (a != b && sb == su && sa == sw) {
        const double x = xu[b - u0] * xw[a - w0];
        const int ph = (I.j2[sa] + I.j2[sb]) / 2 - t.J + 1;
        v += (abs(ph) & 1) ? -x : x;
      }
    } else if (t.kind == 1) {
      if (sa == su && sb == sw) v = xu[a - u0] * xw[b - w0];
    } else {
      if (sb == su && sa == sw) v = xu[b - u0] * xw[a - w0];
    }
    out[p] = v;
  }
}

// One block per (task, row p): K[p][q] for all q.
__global__ void k_t2_pandya(const PandyaTask *tasks, const int2 *rows, InstDev I, const double *sixj,
                            const int *jidx, int nj, int nJp) {
  const int2 tr = rows[blockIdx.x];
  const PandyaTask t = tasks[tr.x];
  const int p = tr.y;
  const int i = t.pair[p] >> 16, j = t.pair[p] & 0xffff;
  const int ji = I.j2[I.orb_sh[i]], jj = I.j2[I.orb_sh[j]];
  const long nq = (long)nj * nj * nj * nj;
  const double *W = sixj + (long)t.slot * nq * nJp;
  const int xi = jidx[ji], xj = jidx[jj];
  double *out = t.K + (long)p * t.ld;
  for (int q = threadIdx.x; q < t.npair; q += blockDim.x) {
    const int k = t.pair[q] >> 16, l = t.pair[q] & 0xffff;
    const int jk = I.j2[I.orb_sh[k]], jl = I.j2[I.orb_sh[l]];
    const double *w = W + ((((long)xi * nj + xj) * nj + jidx[jk]) * nj + jidx[jl]) * nJp;
    const int lo = max(abs(ji - jl), abs(jj - jk)), hi = min(ji + jl, jj + jk);
    double v = 0.0;
    for (int J2 = lo; J2 <= hi; J2 += 2) v -= w[J2 / 2] * vget(I, i, l, k, j, J2 / 2);
    if (i == l) v *= kSqrt2;
    if (k == j) v *= kSqrt2;
    out[q] = v;
  }
}

// C = A B^T on FP64 MFMA (v_mfma_f64_16x16x4f64).  Block of four waves in a
// 2x2 arrangement; each wave owns TI x TI MFMA tiles of 16x16, so the block
// tile is (32 TI) x (32 TI): 64 for TI = 2, 128 for TI = 4.  K advances in
// steps of TK through a double-buffered LDS stage.  Operand layout of the
// MFMA (measured on gfx950): lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and receives C[l/16 + 4v][l%16] in item v = 0..3.
template <bool EPI, int TI, int TK, int NBUF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_t2_gemm(const GemmTask *tasks,
                                                                                           const int4 *tiles) {
  constexpr int BT = 32 * TI;                 // block tile edge
  constexpr int LPT = BT * TK / 256;          // doubles per thread per operand per K step
  constexpr int TPR = TK / LPT;               // threads per tile row
  constexpr int RP = TK + 2;                  // LDS row pitch (doubles): 16 lanes of a read hit distinct banks
  static_assert(TK % 8 == 0 && LPT % 2 == 0, "K step in whole 16-byte runs per lane");
  const int4 tl = tiles[blockIdx.x];
  if (tl.x < 0) return;            // idle entry of the XCD-interleaved tile list
  const GemmTask t = tasks[tl.x];
  const int m0 = tl.y * BT, n0 = tl.z * BT;
  // Within a row, k = 4 kk + q is stored at q * (TK/4) + kk: the TK/4
  // operands a lane feeds to the MFMAs of a K step (same q, every kk) are
  // adjacent.
  __shared__ double smem[2 * NBUF * BT * RP];
  double(*As)[BT][RP] = reinterpret_cast<double(*)[BT][RP]>(smem);
  double(*Bs)[BT][RP] = reinterpret_cast<double(*)[BT][RP]>(smem + NBUF * BT * RP);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 16 * TI, wn = (wave & 1) * 16 * TI;
  const int lr = tid / TPR, lk = (tid % TPR) * LPT;
  const double *Ar = t.A + (long)(m0 + lr) * t.lda;
  const double *Br = t.B + (long)(n0 + lr) * t.ldb;
  double ra[LPT], rb[LPT];
  auto gload = [&](int k0) {
    const double2 *pa = (const double2 *)(Ar + k0 + lk), *pb = (const double2 *)(Br + k0 + lk);
#pragma unroll
    for (int q = 0; q < LPT / 2; ++q) {
      const double2 x = pa[q], y = pb[q];
      ra[2 * q] = x.x; ra[2 * q + 1] = x.y;
      rb[2 * q] = y.x; rb[2 * q + 1] = y.y;
    }
  };
  // this thread holds k = lk + q, q < LPT
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int k = lk + q, pos = (k & 3) * (TK / 4) + (k >> 2);
      As[buf][lr][pos] = ra[q];
      Bs[buf][lr][pos] = rb[q];
    }
  };
  d4 acc[TI][TI];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TI; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  const int nk = (t.K + TK - 1) / TK;
  gload(0);
  if (NBUF == 2) {
    sstore(0);
    __syncthreads();
  }
  for (int kt = 0; kt < nk; ++kt) {
    // two LDS stages: compute on one while the other is refilled (one barrier
    // per K step); one stage: refill, barrier, compute, barrier
    const int cur = NBUF == 2 ? (kt & 1) : 0;
    if (NBUF == 1) {
      sstore(0);
      __syncthreads();
    }
    if (kt + 1 < nk) gload((kt + 1) * TK);
    constexpr int KK = TK / 4;
    const int q4 = (lane >> 4) * KK;
    double a[TI][KK], b[TI][KK];
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const double2 *p = (const double2 *)&As[cur][wm + i * 16 + (lane & 15)][q4];
#pragma unroll
      for (int h = 0; h < KK / 2; ++h) {
        const double2 x = p[h];
        a[i][2 * h] = x.x; a[i][2 * h + 1] = x.y;
      }
    }
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const double2 *p = (const double2 *)&Bs[cur][wn + j * 16 + (lane & 15)][q4];
#pragma unroll
      for (int h = 0; h < KK / 2; ++h) {
        const double2 x = p[h];
        b[j][2 * h] = x.x; b[j][2 * h + 1] = x.y;
      }
    }
#pragma unroll
    for (int kk = 0; kk < TK / 4; ++kk)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TI; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][kk], b[j][kk], acc[i][j], 0, 0, 0);
    if (NBUF == 2 && kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  if constexpr (EPI && TI == 2) {
    // Staged epilogue: half a tile (32 rows) at a time through LDS, then each
    // placement written with consecutive lanes on consecutive addresses
    // (along the tile rows or along the tile columns, whichever is
    // contiguous in the output for that placement).
    constexpr int HR = 32, YP = BT + 1;
    static_assert(2 * NBUF * BT * RP >= HR * YP, "LDS stage too small for half a tile");
    double *Y = smem;
    for (int h = 0; h < 2; ++h) {
      if (wm == h * HR) {
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TI; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              Y[(i * 16 + (lane >> 4) + 4 * v) * YP + wn + j * 16 + (lane & 15)] = acc[i][j][v];
      }
      __syncthreads();
      const int rbase = m0 + h * HR;
      for (int k = 0; k < t.nw; ++k) {
        const bool tr = t.tr[k] != 0;
        const long c0 = t.c0[k], r0 = t.r0[k];
        const bool neg = t.sgn[k] < 0;
        for (int mirror = 0; mirror < (t.sym ? 2 : 1); ++mirror) {
          const bool rowfast = tr == (mirror != 0);
          for (int it = 0; it < HR * BT / 256; ++it) {
            const int e = it * 256 + tid;
            const int rl = rowfast ? (e % HR) : (e / BT);
            const int cl = rowfast ? (e / HR) : (e % BT);
            const int row = rbase + rl, col = n0 + cl;
            if (row >= t.Mv || col >= t.Nv) continue;
            if (t.sym && (row > col || (mirror && row == col))) continue;
            double y = Y[rl * YP + cl];
            if (t.x3) y = t.x3[col] * y;
            if (t.rs && t.rs[row]) y /= kSqrt2;
            if (t.cs && t.cs[col]) y /= kSqrt2;
            const bool swap = tr != (mirror != 0);      // entry (C, R) = (c0 + row, r0 + col)
            const long C = swap ? c0 + row : c0 + col;
            const long R = swap ? r0 + col : r0 + row;
            t.out[C * t.ld + R] = neg ? -y : y;
          }
        }
      }
      __syncthreads();
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TI; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
      if (col >= t.Nv) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm + i * 16 + (lane >> 4) + 4 * v;
        if (row >= t.Mv) continue;
        double y = acc[i][j][v];
        if (!EPI) {
          t.C[(long)row * t.ldc + col] = y;
          continue;
        }
        if (t.sym && row > col) continue;
        if (t.x3) y = t.x3[col] * y;
        if (t.rs && t.rs[row]) y /= kSqrt2;
        if (t.cs && t.cs[col]) y /= kSqrt2;
        for (int k = 0; k < t.nw; ++k) {
          const long C = t.tr[k] ? t.c0[k] + row : t.c0[k] + col;
          const long R = t.tr[k] ? t.r0[k] + col : t.r0[k] + row;
          t.out[C * t.ld + R] = t.sgn[k] < 0 ? -y : y;
          if (t.sym && row != col) {
            const long Cm = t.tr[k] ? t.c0[k] + col : t.c0[k] + row;
            const long Rm = t.tr[k] ? t.r0[k] + row : t.r0[k] + col;
            t.out[Cm * t.ld + Rm] = t.sgn[k] < 0 ? -y : y;
          }
        }
      }
    }
}

__global__ void k_t2_diag(const DiagTask *tasks, int ntask) {
  for (int k = blockIdx.x; k < ntask; k += gridDim.x) {
    const DiagTask t = tasks[k];
    for (int i = threadIdx.x; i < t.n; i += blockDim.x) {
      double &d = t.out[(long)(t.k0 + i) * (t.ld + 1)];
      d = t.sgn < 0 ? d - t.e[i] : d + t.e[i];
    }
  }
}

// ------------------------------------------------------------------- system

namespace {

bool tri2(int a, int b, int c) {
  const int s1 = a + b - c, s2 = b + c - a, s3 = c + a - b;
  return s1 >= 0 && s2 >= 0 && s3 >= 0 && !((s1 | s2 | s3) & 1);
}

template <class T>
int dev_copy(std::vector<void *> &bufs, const T *src, size_t n, T **dst) {
  void *p = nullptr;
  if (bcd::sync_malloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return BCD_ERR_NOMEM;
  bufs.push_back(p);
  if (n && bcd::sync_memcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return BCD_ERR_DEVICE;
  *dst = (T *)p;
  return BCD_OK;
}

}  // namespace

void T2System::release() {
  for (void *p : bufs) (void)bcd::sync_free(p);
  bufs.clear();
  d_pairs = nullptr;
  d_jidx = nullptr;
}

int T2System::upload(const HostSystem &S) {
  release();
  std::memset(&I, 0, sizeof I);
  I.nsubsh = S.nsubsh; I.norb = S.norb; I.nspb = S.nspb; I.jmax = S.jmax; I.nchan2b = S.nchan2b;
  int *p;
  long *lp;
  double *dp;
  int rc;
#define UP(field, vec)                                              \
  if ((rc = dev_copy(bufs, vec.data(), vec.size(), &p)) != BCD_OK) \
    return rc;                                                      \
  I.field = p;
  UP(l, S.l) UP(j2, S.j2) UP(ip, S.ip) UP(ch, S.ch) UP(no, S.no) UP(orb0, S.orb0) UP(orb_sh, S.orb_sh)
  UP(pidx, S.pidx) UP(npair, S.npair)
#undef UP
  if ((rc = dev_copy(bufs, S.voff.data(), S.voff.size(), &lp)) != BCD_OK) return rc;
  I.voff = lp;
  if ((rc = dev_copy(bufs, S.V.data(), S.V.size(), &dp)) != BCD_OK) return rc;
  I.V = dp;
  // pair list of every dense V block, and the (J, parity, charge) -> block map
  pairs.assign(S.nchan2b, {});
  chan_of.assign((size_t)(S.jmax + 1) * 2 * 8, -1);
  for (int J = 0; J <= S.jmax; ++J)
    for (int a = 0; a < S.norb; ++a)
      for (int b = a; b < S.norb; ++b) {
        const int pe = S.pidx[((size_t)J * S.norb + a) * S.norb + b];
        if (pe < 0) continue;
        const int c = pe >> 16, row = pe & 0xffff;
        std::vector<int> &pl = pairs[c];
        if ((int)pl.size() <= row) pl.resize(row + 1, -1);
        pl[row] = (a << 16) | b;
        const int sa = S.orb_sh[a], sb = S.orb_sh[b];
        chan_of[((size_t)J * 2 + (S.l[sa] + S.l[sb]) % 2) * 8 + (S.ch[sa] + S.ch[sb])] = c;
      }
  std::vector<int> flat;
  pair_off.assign(S.nchan2b + 1, 0);
  for (int c = 0; c < S.nchan2b; ++c) {
    pair_off[c] = (long)flat.size();
    flat.insert(flat.end(), pairs[c].begin(), pairs[c].end());
  }
  pair_off[S.nchan2b] = (long)flat.size();
  if ((rc = dev_copy(bufs, flat.data(), flat.size(), &d_pairs)) != BCD_OK) return rc;
  // distinct 2j values
  int jmax2 = 0;
  for (int s = 0; s < S.nsubsh; ++s) jmax2 = std::max(jmax2, S.j2[s]);
  jidx.assign(jmax2 + 1, -1);
  nj = 0;
  for (int v = 0; v <= jmax2; ++v)
    for (int s = 0; s < S.nsubsh; ++s)
      if (S.j2[s] == v) { jidx[v] = nj++; break; }
  if ((rc = dev_copy(bufs, jidx.data(), jidx.size(), &d_jidx)) != BCD_OK) return rc;
  return BCD_OK;
}

// -------------------------------------------------------------------- bases

namespace {

// icut == 0 keeps every pair; otherwise (n1(u)+n1(w))*icut >= ntcut, with
// ntcut -> 1-ntcut for icut < 0 (gII_b.cpp:67-70, PolProp_b.cpp:48-50).
bool cut_keep(int icut, int ntcut, int n1u, int n1w) {
  if (icut == 0) return true;
  if (icut < 0) ntcut = 1 - ntcut;
  return (n1u + n1w) * icut >= ntcut;
}

}  // namespace

int t2_ladd_basis(const HostSystem &S, const HostFrags &F, int J, int dT, int pi, int icut, int ntcut,
                  const int *n1p, const int *n1h, T2Job &job) {
  job.kind = 0; job.J = J; job.dT = dT; job.pi = pi;
  job.a1.clear(); job.a2.clear(); job.b1.clear(); job.b2.clear(); job.ea.clear(); job.eb.clear();
  for (int side = 0; side < 2; ++side) {
    const int n = F.n[side];
    const int *n1 = side ? n1h : n1p;
    if (icut && !n1) return BCD_ERR_ARG;
    std::vector<int> &x1 = side ? job.b1 : job.a1, &x2 = side ? job.b2 : job.a2;
    std::vector<double> &e = side ? job.eb : job.ea;
    for (int u = 0; u < n; ++u)
      for (int w = u; w < n; ++w) {
        if (icut && !cut_keep(icut, ntcut, n1[u], n1[w])) continue;
        const int su = F.sub[side][u], sw = F.sub[side][w];
        if ((pi + S.ip[su] + S.ip[sw]) % 2) continue;
        if (!tri2(S.j2[su], S.j2[sw], 2 * J)) continue;
        if (u == w && (std::abs((S.j2[su] + S.j2[sw]) / 2 - J + 1) % 2)) continue;
        if (S.ch[su] + S.ch[sw] != dT) continue;
        x1.push_back(u);
        x2.push_back(w);
        e.push_back(F.e[side][u] + F.e[side][w]);
      }
  }
  return BCD_OK;
}

int t2_ph_basis(const HostSystem &S, const HostFrags &F, int J, int dT, int pi, int icut, int ntcut,
                const int *n1p, const int *n1h, T2Job &job) {
  job.kind = 1; job.J = J; job.dT = dT; job.pi = pi;
  job.a1.clear(); job.a2.clear(); job.b1.clear(); job.b2.clear(); job.ea.clear(); job.eb.clear();
  if (icut && (!n1p || !n1h)) return BCD_ERR_ARG;
  for (int u = 0; u < F.n[0]; ++u)
    for (int w = 0; w < F.n[1]; ++w) {
      if (icut && !cut_keep(icut, ntcut, n1p[u], n1h[w])) continue;
      const int sa = F.sub[0][u], sb = F.sub[1][w];
      if ((pi + S.ip[sa] + S.ip[sb]) % 2) continue;
      if (!tri2(S.j2[sa], S.j2[sb], 2 * J)) continue;
      const double e = F.e[0][u] - F.e[1][w];
      if (S.ch[sa] - S.ch[sb] == dT) { job.a1.push_back(u); job.a2.push_back(w); job.ea.push_back(e); }
      if (S.ch[sb] - S.ch[sa] == dT) { job.b1.push_back(u); job.b2.push_back(w); job.eb.push_back(e); }
    }
  return BCD_OK;
}

int t2_job_dims(T2Job &j) {
  const int na = (int)j.ea.size(), nb = (int)j.eb.size();
  const double untouched = std::nan("");
  if (j.mode < 0 || j.mode > 4 || (j.kind == 1 && j.mode == 2)) return -100;
  if (j.kind == 0) {
    j.nf = j.mode == 2 ? 0 : na;
    j.nb = j.mode == 1 ? 0 : nb;
    j.free_poles.assign(j.nf + j.nb, untouched);
    if (j.mode != 4) {
      if (j.mode != 2)
        for (int k = 0; k < na; ++k) j.free_poles[k] = j.ea[k];
      if (j.mode != 1)
        for (int k = 0; k < nb; ++k) j.free_poles[j.nf + k] = j.eb[k];
    }
  } else {
    j.nf = na;
    j.nb = j.mode == 1 ? 0 : nb;
    j.free_poles.assign(j.nf + j.nb, untouched);
    if (j.mode != 4) {
      const bool copied = j.dT == 0 && j.mode != 1, direct = j.dT != 0 && j.mode != 1;
      for (int k = 0; k < na; ++k) {
        j.free_poles[k] = j.ea[k];
        if (copied) j.free_poles[j.nf + k] = j.ea[k];
      }
      if (direct)
        for (int k = 0; k < nb; ++k) j.free_poles[j.nf + k] = j.eb[k];
    }
  }
  return j.nf + j.nb;
}

// --------------------------------------------------------------------- plan

namespace {

// Staging of one plan: a dry pass counts, the real pass hands out device
// addresses and fills the host copies of the uploaded arrays.
struct Bump {
  bool dry = true;
  int *di = nullptr;
  double *dd = nullptr, *dws = nullptr, *dout = nullptr;
  std::vector<int> hi;
  std::vector<double> hd;
  long ni = 0, nd = 0, nws = 0, nout = 0;
  static long up(long n) { return (n + 31) & ~31L; }
  const int *ints(const int *src, long n) {
    const long o = ni;
    ni += up(n);
    if (dry) return nullptr;
    std::copy(src, src + n, hi.begin() + o);
    return di + o;
  }
  const double *dbls(const double *src, long n) {
    const long o = nd;
    nd += up(n);
    if (dry) return nullptr;
    std::copy(src, src + n, hd.begin() + o);
    return dd + o;
  }
  double *ws(long n) {
    const long o = nws;
    nws += up(n);
    return dry ? nullptr : dws + o;
  }
  double *out(long n) {
    const long o = nout;
    nout += up(n);
    return dry ? nullptr : dout + o;
  }
};

struct FragSubs {
  const std::vector<int> &p, &h;
};

struct Emit {
  const T2System &S;
  const HostSystem &H;
  Bump &B;
  std::vector<AmpTask> &amp;
  std::vector<PandyaTask> &pan;
  std::vector<GemmTask> &g1, &g2;
  std::vector<DiagTask> &diag;
  std::vector<ZeroTask> &zero;
  std::vector<int> &sixj_J;
  double &f1, &f2, &nout;
  long area = 0;                   // output entries the current job's epilogues write

  int slot(int J) {
    for (size_t k = 0; k < sixj_J.size(); ++k)
      if (sixj_J[k] == J) return (int)k;
    sixj_J.push_back(J);
    return (int)sixj_J.size() - 1;
  }

  // Operand padding: T [r64(rows) x r16(P)], K [r64(P) x r16(P)],
  // G [r64(rows) x r64(P)]; the workspace is zeroed once, and the padding is
  // never written except by GEMM1, which writes exact zeros there.
  static int r16(int n) { return (n + kTK - 1) / kTK * kTK; }          // K padding
  static int r64(int n) {                                              // row padding: the block tile
    const int bt = kUseBigTiles ? kBigTile : kBigTile / 2;
    return (n + bt - 1) / bt * bt;
  }

  double *amp_task(int kind, int J, const std::vector<int> &u, const std::vector<int> &w, int ukind, int wkind,
                   const int *pair, int npair) {
    AmpTask t;
    std::memset(&t, 0, sizeof t);
    t.kind = kind; t.J = J; t.nrows = (int)u.size(); t.npair = npair; t.ld = r16(npair);
    t.ukind = ukind; t.wkind = wkind;
    t.u = B.ints(u.data(), (long)u.size());
    t.w = B.ints(w.data(), (long)w.size());
    t.pair = pair;
    t.T = B.ws((long)r64(t.nrows) * t.ld);
    amp.push_back(t);
    return t.T;
  }

  double *kmat(int P) { return B.ws((long)r64(P) * r16(P)); }

  double *gemm1(const double *A, int M, const double *K, int P) {
    GemmTask t;
    std::memset(&t, 0, sizeof t);
    t.M = r64(M); t.N = r64(P); t.K = r16(P); t.lda = r16(P); t.ldb = r16(P); t.ldc = r64(P);
    t.Mv = t.M; t.Nv = t.N;
    t.A = A; t.B = K;
    t.C = B.ws((long)t.M * t.ldc);
    g1.push_back(t);
    f1 += 2.0 * M * P * P;
    return t.C;
  }

  void gemm2(const double *G, int M, const double *Tc, int N, int P, double *out, int ld, const int *rs,
             const int *cs, const double *x3, int nw, const int (&w)[2][4], bool sym = false) {
    GemmTask t;
    std::memset(&t, 0, sizeof t);
    t.M = r64(M); t.N = r64(N); t.K = r16(P); t.lda = r64(P); t.ldb = r16(P); t.sym = sym;
    t.Mv = M; t.Nv = N;
    t.A = G; t.B = Tc; t.out = out; t.ld = ld; t.nw = nw;
    for (int k = 0; k < nw; ++k) {
      t.c0[k] = w[k][0]; t.r0[k] = w[k][1]; t.tr[k] = w[k][2]; t.sgn[k] = w[k][3];
    }
    t.rs = rs; t.cs = cs; t.x3 = x3;
    g2.push_back(t);
    f2 += (sym ? 1.0 : 2.0) * M * N * P;
    nout += (double)nw * M * N;
    area += (long)nw * M * N;
  }

  void energies(double *out, int ld, const std::vector<double> &e, int k0, int sgn) {
    if (e.empty()) return;
    DiagTask d;
    d.out = out; d.ld = ld; d.k0 = k0; d.n = (int)e.size(); d.sgn = sgn;
    d.e = B.dbls(e.data(), (long)e.size());
    diag.push_back(d);
  }

  const int *eq_flags(const std::vector<int> &u, const std::vector<int> &w) {
    std::vector<int> f(u.size());
    for (size_t k = 0; k < u.size(); ++k) f[k] = u[k] == w[k];
    return B.ints(f.data(), (long)f.size());
  }

  // x3 of PolProp_b.cpp:330-331: -1 when (|2j_g - 2j_d| / 2) is odd, for the
  // (qp g, qh d) of each column entry
  const double *x3_of(const std::vector<int> &g, const std::vector<int> &d, const FragSubs &F) {
    std::vector<double> x(g.size());
    for (size_t k = 0; k < g.size(); ++k) {
      const int sg = F.p[g[k]], sd = F.h[d[k]];
      x[k] = ((std::abs(H.j2[sg] - H.j2[sd]) / 2) % 2) ? -1.0 : 1.0;
    }
    return B.dbls(x.data(), (long)x.size());
  }


  void ring_pairs(int J, int pi, int dT, std::vector<int> &out) const {
    out.clear();
    for (int i = 0; i < H.norb; ++i)
      for (int j = 0; j < H.norb; ++j) {
        const int si = H.orb_sh[i], sj = H.orb_sh[j];
        if ((pi + H.ip[si] + H.ip[sj]) % 2) continue;
        if (!tri2(H.j2[si], H.j2[sj], 2 * J)) continue;
        if (H.ch[si] - H.ch[sj] != dT) continue;
        out.push_back((i << 16) | j);
      }
  }
};

}  // namespace

void T2Plan::release() {
  for (void *p : bufs) (void)bcd::sync_free(p);
  bufs.clear();
  d_out = nullptr;
  d_sixj = nullptr;
}

int T2Plan::prepare(const T2System &S, const HostSystem &H, const HostFrags &F, hipStream_t st) {
  release();
  for (T2Job &j : jobs)
    if (t2_job_dims(j) < 0) return BCD_ERR_ARG;
  Bump B;
  std::vector<int> ring_list;
  auto emit_all = [&](std::vector<int> &pairs_host_out) {
    amp.clear(); pan.clear(); g1.clear(); g2.clear(); diag.clear(); zero.clear(); sixj_J.clear(); vcopy.clear();
    flops_g1 = flops_g2 = out_doubles = 0.0;
    pairs_host_out.clear();
    Emit E{S, H, B, amp, pan, g1, g2, diag, zero, sixj_J, flops_g1, flops_g2, out_doubles};
    for (T2Job &j : jobs) {
      const int ntot = j.nf + j.nb;
      if (j.ld < std::max(ntot, 1)) j.ld = std::max(ntot, 1);
      double *out = B.out((long)std::max(ntot, 1) * j.ld);
      j.out_off = B.dry ? 0 : (long)(out - B.dout);
      E.area = 0;
      const int nf = j.nf, ld = j.ld;
      const bool add = j.add || j.mode == 4;
      if (j.kind == 0) {
        const bool make_pp = j.mode != 2, make_hh = j.mode != 1, make_b = j.mode == 3;
        if (j.mode != 4) {
          const size_t key = ((size_t)j.J * 2 + j.pi) * 8 + j.dT;
          const int c = (j.J <= H.jmax && j.dT >= 0 && j.dT < 8) ? S.chan_of[key] : -1;
          const int P = c >= 0 ? H.npair[c] : 0;
          if (P > 0) {
            const int *pairs = S.d_pairs + S.pair_off[c];
            double *K = E.kmat(P);
            vcopy.push_back(VCopy{K, S.I.V + H.voff[c], P, Emit::r16(P)});
            const int npp = (int)j.ea.size(), nhh = (int)j.eb.size();
            const double *Tpp = nullptr, *Thh = nullptr, *Gpp = nullptr, *Ghh = nullptr;
            const int *eqp = nullptr, *eqh = nullptr;
            if ((make_pp || make_b) && npp) {
              Tpp = E.amp_task(0, j.J, j.a1, j.a2, 0, 0, pairs, P);
              eqp = E.eq_flags(j.a1, j.a2);
            }
            if ((make_hh || make_b) && nhh) {
              Thh = E.amp_task(0, j.J, j.b1, j.b2, 1, 1, pairs, P);
              eqh = E.eq_flags(j.b1, j.b2);
            }
            if (make_pp && npp) Gpp = E.gemm1(Tpp, npp, K, P);
            if ((make_hh || make_b) && nhh) Ghh = E.gemm1(Thh, nhh, K, P);
            if (make_pp && npp)
              E.gemm2(Gpp, npp, Tpp, npp, P, out, ld, eqp, eqp, nullptr, 1, {{0, 0, 0, 1}, {0, 0, 0, 0}}, true);
            if (make_b && npp && nhh)
              E.gemm2(Ghh, nhh, Tpp, npp, P, out, ld, eqh, eqp, nullptr, 2, {{0, nf, 0, -1}, {nf, 0, 1, 1}});
            if (make_hh && nhh)
              E.gemm2(Ghh, nhh, Thh, nhh, P, out, ld, eqh, eqh, nullptr, 1, {{nf, nf, 0, -1}, {0, 0, 0, 0}}, true);
          }
        }
        if (add) {
          if (make_pp) E.energies(out, ld, j.ea, 0, 1);
          if (make_hh) E.energies(out, ld, j.eb, nf, 1);
        }
      } else {
        const bool direct = j.dT != 0 && j.mode != 1, copied = j.dT == 0 && j.mode != 1;
        const bool make_b = j.mode == 3, t0 = copied && make_b;
        const int nph = (int)j.ea.size(), nhp = (int)j.eb.size();
        FragSubs FS{F.sub[0], F.sub[1]};
        if (j.mode != 4) {
          std::vector<int> qp, qm;
          E.ring_pairs(j.J, j.pi, j.dT, qp);
          const int Pp = (int)qp.size();
          if (Pp > 0 && nph > 0) {
            const int *dqp = B.ints(qp.data(), Pp);
            PandyaTask pt{j.J, Pp, Emit::r16(Pp), E.slot(j.J), dqp, E.kmat(Pp)};
            pan.push_back(pt);
            const double *Tph = E.amp_task(1, j.J, j.a1, j.a2, 0, 1, dqp, Pp);
            const double *Gph = E.gemm1(Tph, nph, pt.K, Pp);
            if (copied)
              E.gemm2(Gph, nph, Tph, nph, Pp, out, ld, nullptr, nullptr, nullptr, 2, {{0, 0, 0, 1}, {nf, nf, 0, -1}},
                      true);
            else
              E.gemm2(Gph, nph, Tph, nph, Pp, out, ld, nullptr, nullptr, nullptr, 1, {{0, 0, 0, 1}, {0, 0, 0, 0}},
                      true);
            if (t0) {
              const double *Tt = E.amp_task(2, j.J, j.a1, j.a2, 0, 1, dqp, Pp);
              E.gemm2(Gph, nph, Tt, nph, Pp, out, ld, nullptr, nullptr, E.x3_of(j.a1, j.a2, FS), 2,
                      {{0, nf, 0, -1}, {nf, 0, 0, 1}});
            }
            if (direct && make_b && nhp > 0) {
              const double *Tt = E.amp_task(2, j.J, j.b1, j.b2, 0, 1, dqp, Pp);
              E.gemm2(Gph, nph, Tt, nhp, Pp, out, ld, nullptr, nullptr, E.x3_of(j.b1, j.b2, FS), 2,
                      {{nf, 0, 0, 1}, {0, nf, 1, -1}});
            }
          }
          if (direct && nhp > 0) {
            E.ring_pairs(j.J, j.pi, -j.dT, qm);
            const int Pm = (int)qm.size();
            if (Pm > 0) {
              const int *dqm = B.ints(qm.data(), Pm);
              PandyaTask pt{j.J, Pm, Emit::r16(Pm), E.slot(j.J), dqm, E.kmat(Pm)};
              pan.push_back(pt);
              const double *Thp = E.amp_task(1, j.J, j.b1, j.b2, 0, 1, dqm, Pm);
              const double *Ghp = E.gemm1(Thp, nhp, pt.K, Pm);
              E.gemm2(Ghp, nhp, Thp, nhp, Pm, out, ld, nullptr, nullptr, nullptr, 1, {{nf, nf, 0, -1}, {0, 0, 0, 0}},
                      true);
            }
          }
        }
        if (add) {
          E.energies(out, ld, j.ea, 0, 1);
          if (j.dT == 0 && j.mode != 1) E.energies(out, ld, j.ea, nf, -1);
          if (direct) E.energies(out, ld, j.eb, nf, -1);
        }
      }
      // the extent Fill_*_mtx clears, unless the products write all of it
      // (every block present and no padding between rows); the clearing
      // kernel runs before the products
      if (ntot > 0 && !(E.area == (long)ntot * ntot && ld == ntot))
        zero.push_back(ZeroTask{out, (long)(ntot - 1) * ld + ntot});
    }
  };
  std::vector<int> unused;
  B.dry = true;
  emit_all(unused);
  // allocate
  auto dmalloc = [&](size_t bytes, void **p) {
    if (bcd::sync_malloc(p, std::max<size_t>(bytes, 8)) != hipSuccess) return BCD_ERR_NOMEM;
    bufs.push_back(*p);
    return BCD_OK;
  };
  void *pi, *pd, *pw, *po;
  int rc;
  if ((rc = dmalloc(sizeof(int) * B.ni, &pi)) || (rc = dmalloc(sizeof(double) * B.nd, &pd)) ||
      (rc = dmalloc(sizeof(double) * B.nws, &pw)) || (rc = dmalloc(sizeof(double) * B.nout, &po)))
    return rc;
  const long ni = B.ni, nd = B.nd;
  B = Bump();
  B.dry = false;
  B.di = (int *)pi; B.dd = (double *)pd; B.dws = (double *)pw; B.dout = (double *)po;
  B.hi.assign(ni, 0);
  B.hd.assign(nd, 0.0);
  emit_all(unused);
  d_out = B.dout;
  out_len = B.nout;
  if (hipMemcpyAsync(pi, B.hi.data(), sizeof(int) * B.hi.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(pd, B.hd.data(), sizeof(double) * B.hd.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(pw, 0, sizeof(double) * std::max<long>(B.nws, 1), st) != hipSuccess)
    return BCD_ERR_DEVICE;
  for (const VCopy &v : vcopy)
    if (hipMemcpy2DAsync(v.dst, sizeof(double) * v.ld, v.src, sizeof(double) * v.P, sizeof(double) * v.P, v.P,
                         hipMemcpyDeviceToDevice, st) != hipSuccess)
      return BCD_ERR_DEVICE;
  // 6j tables: the Racah sum of arguments up to 2j = jmax2 (orbits) and
  // 2J, 2J' <= 2 jmax2 reaches binomial rows 2 jmax2 + 1
  if (!sixj_J.empty() && 2 * ((int)S.jidx.size() - 1) + 2 >= kBin) return BCD_ERR_CAPACITY;
  if (!sixj_J.empty()) {
    std::vector<double> bn(kBin * kBin, 0.0);
    for (int n = 0; n < kBin; ++n) { bn[n * kBin] = 1.0; bn[n * kBin + n] = 1.0; }
    for (int n = 1; n < kBin - 1; ++n)
      for (int k = 1; k <= n; ++k) bn[(n + 1) * kBin + k] = bn[n * kBin + k - 1] + bn[n * kBin + k];
    std::vector<int> jv(S.nj);
    for (size_t v = 0; v < S.jidx.size(); ++v)
      if (S.jidx[v] >= 0) jv[S.jidx[v]] = (int)v;
    double *dbn;
    int *djv, *dJs;
    std::vector<void *> tmp;
    if ((rc = dev_copy(bufs, bn.data(), bn.size(), &dbn)) || (rc = dev_copy(bufs, jv.data(), jv.size(), &djv)) ||
        (rc = dev_copy(bufs, sixj_J.data(), sixj_J.size(), &dJs)))
      return rc;
    const int nJp = H.jmax + 1;
    const long n = (long)sixj_J.size() * S.nj * S.nj * S.nj * S.nj * nJp;
    void *p6;
    if ((rc = dmalloc(sizeof(double) * n, &p6))) return rc;
    d_sixj = (double *)p6;
    hipLaunchKernelGGL(k_t2_sixj, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, st, d_sixj,
                       dJs, (int)sixj_J.size(), S.nj, djv, nJp, dbn);
    if (hipGetLastError() != hipSuccess) return BCD_ERR_DEVICE;
  }
  // task lists and launch tables
  auto up_tasks = [&](auto &vec, auto **dst) {
    using T = typename std::remove_reference<decltype(vec)>::type::value_type;
    T *p = nullptr;
    int r = dev_copy(bufs, vec.data(), vec.size(), &p);
    *dst = p;
    return r;
  };
  std::vector<int2> arows, prows, zchunks;
  for (size_t k = 0; k < amp.size(); ++k)
    for (int r = 0; r < amp[k].nrows; ++r) arows.push_back(int2{(int)k, r});
  for (size_t k = 0; k < pan.size(); ++k)
    for (int r = 0; r < pan[k].npair; ++r) prows.push_back(int2{(int)k, r});
  for (size_t k = 0; k < zero.size(); ++k)
    for (long c = 0; c * kZeroChunk < zero[k].n; ++c) zchunks.push_back(int2{(int)k, (int)c});
  // tiles of the 64 (small) and 128 (big) kernels; big tiles for tasks whose
  // both dimensions reach 128
  // Blocks are handed to the XCDs round-robin by blockIdx.  Tiles are
  // grouped into patches of kPatch x kPatch; each patch goes whole to one XCD
  // (least loaded first), so the operand bands a patch reads stay in that
  // XCD's L2.  The per-XCD queues are interleaved, padded with idle entries
  // (task -1).
  auto tiles = [](const std::vector<GemmTask> &g, bool big) {
    std::vector<std::vector<int4>> q(kXcd);
    std::vector<long> load(kXcd, 0);
    std::vector<int4> patch;
    for (size_t k = 0; k < g.size(); ++k) {
      if ((kUseBigTiles && std::min(g[k].M, g[k].N) >= kBigTile) != big) continue;
      const int bt = big ? kBigTile : kBigTile / 2;
      const int mt = (g[k].M + bt - 1) / bt, nt = (g[k].N + bt - 1) / bt;
      for (int pm = 0; pm < mt; pm += kPatch)
        for (int pn = g[k].sym ? pm : 0; pn < nt; pn += kPatch) {
          patch.clear();
          for (int m = pm; m < std::min(mt, pm + kPatch); ++m)
            for (int n = pn; n < std::min(nt, pn + kPatch); ++n)
              if (!g[k].sym || n >= m) patch.push_back(int4{(int)k, m, n, 0});
          if (patch.empty()) continue;
          const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
          q[x].insert(q[x].end(), patch.begin(), patch.end());
          load[x] += (long)patch.size();
        }
    }
    size_t len = 0;
    for (const auto &v : q) len = std::max(len, v.size());
    std::vector<int4> t(len * kXcd, int4{-1, 0, 0, 0});
    for (int x = 0; x < kXcd; ++x)
      for (size_t i = 0; i < q[x].size(); ++i) t[i * kXcd + x] = q[x][i];
    return t;
  };
  std::vector<int4> t1 = tiles(g1, false), t2 = tiles(g2, false), t1b = tiles(g1, true), t2b = tiles(g2, true);
  n_t1b = (long)t1b.size();
  n_t2b = (long)t2b.size();
  t1.insert(t1.end(), t1b.begin(), t1b.end());
  t2.insert(t2.end(), t2b.begin(), t2b.end());
  int2 *dar, *dpr, *dzc;
  int4 *dt1, *dt2;
  if ((rc = up_tasks(amp, &d_amp)) || (rc = up_tasks(pan, &d_pan)) || (rc = up_tasks(g1, &d_g1)) ||
      (rc = up_tasks(g2, &d_g2)) || (rc = up_tasks(diag, &d_diag)) || (rc = up_tasks(zero, &d_zero)) ||
      (rc = dev_copy(bufs, arows.data(), arows.size(), &dar)) ||
      (rc = dev_copy(bufs, prows.data(), prows.size(), &dpr)) ||
      (rc = dev_copy(bufs, zchunks.data(), zchunks.size(), &dzc)) ||
      (rc = dev_copy(bufs, t1.data(), t1.size(), &dt1)) || (rc = dev_copy(bufs, t2.data(), t2.size(), &dt2)))
    return rc;
  d_amp_rows = (int *)dar; amp_rows = (long)arows.size();
  d_pan_rows = (int *)dpr; pan_rows = (long)prows.size();
  d_zero_chunks = dzc; zero_chunks = (long)zchunks.size();
  d_t1 = (int *)dt1; n_t1 = (long)t1.size() - n_t1b;
  d_t2 = (int *)dt2; n_t2 = (long)t2.size() - n_t2b;
  if (hipStreamSynchronize(st) != hipSuccess) return BCD_ERR_DEVICE;
  return BCD_OK;
}

int T2Plan::launch(const T2System &S, const FragDev &frag, hipStream_t st, hipEvent_t *ev) const {
  auto mark = [&](int k) {
    if (ev) (void)hipEventRecord(ev[k], st);
  };
  mark(0);
  if (zero_chunks)
    hipLaunchKernelGGL(k_t2_zero, dim3((unsigned)zero_chunks), dim3(256), 0, st, d_zero, d_zero_chunks);
  mark(1);
  if (amp_rows)
    hipLaunchKernelGGL(k_t2_amp, dim3((unsigned)amp_rows), dim3(128), 0, st, d_amp, (const int2 *)d_amp_rows, S.I,
                       frag);
  mark(2);
  if (pan_rows)
    hipLaunchKernelGGL(k_t2_pandya, dim3((unsigned)pan_rows), dim3(128), 0, st, d_pan, (const int2 *)d_pan_rows,
                       S.I, d_sixj, S.d_jidx, S.nj, S.I.jmax + 1);
  mark(3);
  const int4 *t1 = (const int4 *)d_t1, *t2 = (const int4 *)d_t2;
  if (n_t1b) hipLaunchKernelGGL((k_t2_gemm<false, 4, kTK, kNBuf>), dim3((unsigned)n_t1b), dim3(256), 0, st, d_g1, t1 + n_t1);
  if (n_t1) hipLaunchKernelGGL((k_t2_gemm<false, 2, kTK, kNBuf>), dim3((unsigned)n_t1), dim3(256), 0, st, d_g1, t1);
  mark(4);
  if (n_t2b) hipLaunchKernelGGL((k_t2_gemm<true, 4, kTK, kNBuf>), dim3((unsigned)n_t2b), dim3(256), 0, st, d_g2, t2 + n_t2);
  if (n_t2) hipLaunchKernelGGL((k_t2_gemm<true, 2, kTK, kNBuf>), dim3((unsigned)n_t2), dim3(256), 0, st, d_g2, t2);
  mark(5);
  if (!diag.empty())
    hipLaunchKernelGGL(k_t2_diag, dim3((unsigned)std::min<size_t>(diag.size(), 1024)), dim3(128), 0, st, d_diag,
                       (int)diag.size());
  mark(6);
  return hipGetLastError() == hipSuccess ? BCD_OK : BCD_ERR_DEVICE;
}

}  // namespace bcd
