// bcd_arrow.hpp -- the Dyson eigenproblem of one channel by successive
// arrowhead updates.
//
// The Dyson matrix of DysonMatrix.cpp:269-390,
//
//       [ h   C^T ]      h: nI x nI static part (U1 + Sigma^inf),
//   A = [         ],     C: couplings of the m self-energy poles,
//       [ C   L   ]      L: diagonal pole energies,
//
// holds the poles in diagonal form: for a Lanczos-reduced side the fishbone
// block D is diagonalised once per solve (D = U L U^T, C = U^T M), which is a
// similarity transform that leaves the first nI components of every
// eigenvector -- the quasiparticle amplitudes -- unchanged.  The eigenpairs of
// A then follow from nI arrowhead problems: head k (orbit k) joins the group
// of already-diagonal directions through its coupling vector, the arrowhead
// [alpha z^T; z diag(d)] is solved by its secular equation
//
//   f(x) = x - alpha + sum_i z_i^2 / (d_i - x) = 0,
//
// one root per interval of the sorted d, the eigenvectors come from couplings
// recomputed from the roots (Loewner's formula; Gu and Eisenstat, SIAM J.
// Matrix Anal. Appl. 16 (1995) 172), and only the nI head rows of the
// eigenvectors plus the couplings of the heads still to come are carried
// along.  Work per channel is O(nI^2 (m+nI)^2) instead of O((m+nI)^3), and
// the roots are independent, so every thread of a block owns a few.
//
// Deflation: couplings below tol = 8 eps max(|alpha|, max|d|, |z|) drop out
// (the direction is an eigenvector to within tol); two directions whose
// energies differ by less than tol after a Givens rotation that moves the
// whole coupling onto one of them are merged the same way.
//
// The body is a template over the block it runs on (thread index, thread
// count, barrier, failure flag) so that the same code runs as a kernel and,
// in the tests, on host threads.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

namespace bcd {

constexpr int kArrowNI = 16;       // largest nI (orbits per subshell)

// The smallest accumulator bound of the compiled ones (4, 8, 16) that holds nI.
__host__ __device__ inline int arrow_ni_bound(int nI) { return nI <= 4 ? 4 : (nI <= 8 ? 8 : 16); }
constexpr int kArrowMaxIt = 100;   // secular-equation iterations per root
constexpr int kArrowMaxThreads = 256;
constexpr int kArrowRed = 40;      // reduction scratch (doubles): ten slots per wave of a 256-thread block

// One channel's Dyson problem in arrowhead form.  Poles come from up to two
// sources: energy i of source s is E[s][i*es[s]], its coupling to orbit r is
// C[s][i*ci[s] + r*cr[s]].
struct ArrowJob {
  int nI, j2, n, inst;             // n = nI + np[0] + np[1]; n <= 0: nothing to do
  int np[2], es[2], ci[2], cr[2];
  const double *E[2], *C[2];
  const double *h, *u1;            // static part = u1 + h (nI x nI, row-major)
  double efermi;
  double *ws;                      // workspace when it does not sit in LDS
  double *out_e, *out_x;           // records: qp (descending e) then qh (descending e)
  int *out_nqp, *out_nqh;
  double *kolt;                    // Calc_En partial sums
  // Roots of the previous sweep (may be null): warm[k * wld + j] is root j
  // of head k's problem and warm_p[k] that problem's pole count (-1: none).
  // A root starts from its previous value when the count is unchanged and
  // the value still lies in the root's interval; the roots are stored back.
  double *warm;
  int *warm_p;
  int wld, _wpad;
};

// Rows per column of the amplitude buffers kept for each of H and G: nI,
// padded to an even count so that a column is read in 16-byte pairs.
__host__ __device__ inline int arrow_h2(int nI) { return (nI + 1) & ~1; }

__host__ __device__ inline size_t arrow_ws_doubles(int n, int nI) {
  return (size_t)(6 + 4 * arrow_h2(nI)) * n + kArrowRed;
}

__host__ __device__ inline size_t arrow_ws_bytes(int n, int nI) {
  return arrow_ws_doubles(n, nI) * sizeof(double) + (3 * (size_t)n + 8) * sizeof(int);
}

// Root j (0..p) of the secular equation with sorted poles d[0..p-1] and
// squared couplings zz (all non-zero).  guess (NaN: none) is a starting
// value, used when it lies strictly inside the root's interval: the origin is
// then the nearer end and the first evaluation is at the guess.  The root is
// returned as an offset tau from the
// pole org it is closest to, so that x - d_i = tau - (d_i - d_org) keeps full
// relative accuracy near the poles.
//
// Iteration: the poles on the origin's side of the root (psi) are modelled by
// the origin pole alone and those on the far side plus the slope of the
// linear term (phi) by the far pole, each weight matching its part's slope at
// the current point (the "middle way" of R.-C. Li, LAPACK Working Note 89,
// 1993); the extreme roots have no far pole and keep the linear term exact.
// For an interior root the first evaluation, at the middle of the interval,
// both picks the origin (the nearer end) and supplies the first step.  A
// bracket on tau falls back to bisection when a step leaves it.  Stops when
// |f| is at the rounding level of its terms or the step is below an ulp of
// tau.  B.rcp(x) is 1/x to within an ulp.  Returns the iterations used
// (kArrowMaxIt if it did not settle).
// Sums of zz_i r_i and zz_i r_i^2 over i in [i0, i1), r_i = 1/((d_i - dofs) -
// tau), in four interleaved partial sums so that the reciprocal chains of
// consecutive poles overlap.
template <class Blk>
__host__ __device__ inline void arrow_sums(const Blk &B, const double *d, const double *zz, int i0, int i1,
                                           double dofs, double tau, double &s, double &ds) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
  int i = i0;
  for (; i + 3 < i1; i += 4) {
    const double r0 = B.rcp((d[i] - dofs) - tau), r1 = B.rcp((d[i + 1] - dofs) - tau);
    const double r2 = B.rcp((d[i + 2] - dofs) - tau), r3 = B.rcp((d[i + 3] - dofs) - tau);
    const double t0 = zz[i] * r0, t1 = zz[i + 1] * r1, t2 = zz[i + 2] * r2, t3 = zz[i + 3] * r3;
    s0 += t0; q0 += t0 * r0;
    s1 += t1; q1 += t1 * r1;
    s2 += t2; q2 += t2 * r2;
    s3 += t3; q3 += t3 * r3;
  }
  for (; i < i1; ++i) {
    const double r0 = B.rcp((d[i] - dofs) - tau), t0 = zz[i] * r0;
    s0 += t0; q0 += t0 * r0;
  }
  s = (s0 + s1) + (s2 + s3);
  ds = (q0 + q1) + (q2 + q3);
}

template <class Blk>
__host__ __device__ inline int arrow_root(const Blk &B, int j, int p, const double *d, const double *zz,
                                          double alpha, double znorm, double guess, int *org_out,
                                          double *tau_out) {
  const double eps = 2.2204460492503131e-16;
  int o = 0, far = -1;
  double lo = 0.0, hi = 0.0, tau = 0.0;
  // poles on the origin's side: below sp (psi_left) or from sp on; an extreme
  // root has them all on one side (sp = p)
  int sp = p;
  bool psi_left = true, eval0 = false, ready = false;   // ready: origin and start chosen
  double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0, mag = 0.0, f = 0.0;
  if (j == 0 || j == p) {
    o = j == 0 ? 0 : p - 1; far = -1;
    if (j == 0) { lo = ((fmin(alpha, d[0]) - d[0]) - znorm) * (1.0 + 1e-7); hi = 0.0; }
    else { lo = 0.0; hi = ((fmax(alpha, d[p - 1]) - d[p - 1]) + znorm) * (1.0 + 1e-7); }
    tau = 0.5 * (lo + hi);
    const double tg = guess - d[o];
    if (tg > lo && tg < hi) tau = tg;
    ready = true;
  } else if (guess > d[j - 1] && guess < d[j]) {
    // start from the guess: origin at the nearer pole, the whole interval as
    // bracket, first evaluation at the guess
    const double a = d[j - 1], b = d[j];
    if (guess - a <= b - guess) { o = j - 1; far = j; lo = 0.0; hi = b - a; tau = guess - a; }
    else { o = j; far = j - 1; lo = a - b; hi = 0.0; tau = guess - b; psi_left = false; }
    sp = j;
    ready = tau > lo && tau < hi;
    eval0 = ready;
  }
  if (!ready) {
    o = 0; far = -1; sp = p; psi_left = true;
    const double a = d[j - 1], b = d[j], mid = 0.5 * (a + b);
    double sl, dsl, sr, dsr;
    arrow_sums(B, d, zz, 0, j, mid, 0.0, sl, dsl);
    arrow_sums(B, d, zz, j, p, mid, 0.0, sr, dsr);
    // terms left of the root are all negative, right of it all positive
    mag = fabs(sl) + fabs(sr);
    f = (mid - alpha) + sl + sr;
    sp = j;
    if (f >= 0.0) {
      o = j - 1; far = j; lo = 0.0; hi = mid - a; tau = hi;
      psi = sl; dpsi = dsl; phi = sr; dphi = dsr;
    } else {
      o = j; far = j - 1; lo = mid - b; hi = 0.0; tau = lo;
      psi_left = false; psi = sr; dpsi = dsr; phi = sl; dphi = dsl;
    }
    mag += fabs(d[o] - alpha) + fabs(tau);
  }
  const double dofs = d[o];
  int it = 0;
  for (; it < kArrowMaxIt; ++it) {
    if (it > 0 || far < 0 || eval0) {
      double sl, dsl, sr, dsr;
      arrow_sums(B, d, zz, 0, sp, dofs, tau, sl, dsl);
      arrow_sums(B, d, zz, sp, p, dofs, tau, sr, dsr);
      mag = fabs(sl) + fabs(sr);
      if (psi_left) { psi = sl; dpsi = dsl; phi = sr; dphi = dsr; }
      else { psi = sr; dpsi = dsr; phi = sl; dphi = dsl; }
      f = (dofs - alpha) + tau + psi + phi;
      mag += fabs(dofs - alpha) + fabs(tau);
      if (fabs(f) <= 8.0 * eps * mag) break;
      if (f < 0.0) lo = tau; else hi = tau;
      if (hi - lo <= 2.0 * eps * fmax(fabs(lo), fabs(hi))) break;
    }
    const double so = dpsi * tau * tau;
    double tn;
    if (far < 0) {
      // a + t - so / t, with a + tau - so / tau = f
      const double a = f - tau + so / tau;
      const double sq = sqrt(a * a + 4.0 * so);
      if (tau > 0.0) tn = a > 0.0 ? 2.0 * so / (a + sq) : 0.5 * (sq - a);
      else tn = a < 0.0 ? -2.0 * so / (sq - a) : -0.5 * (a + sq);
    } else {
      // c + so / (0 - t) + sf / (dfar - t)
      const double dfar = d[far] - dofs, gap = dfar - tau;
      const double sf = (dphi + 1.0) * gap * gap;
      const double c = f + so / tau - sf / gap;
      const double Bq = c * dfar + so + sf, C0 = so * dfar;
      if (c == 0.0) {
        tn = C0 / Bq;
      } else {
        const double disc = Bq * Bq - 4.0 * c * C0;
        if (disc < 0.0) {
          tn = 0.5 * (lo + hi);
        } else {
          const double sq = sqrt(disc), qq = Bq >= 0.0 ? Bq + sq : Bq - sq;
          const double t1 = qq / (2.0 * c), t2 = qq != 0.0 ? 2.0 * C0 / qq : t1;
          tn = (t1 > lo && t1 < hi) ? t1 : t2;
        }
      }
    }
    if (!(tn > lo && tn < hi)) {
      // a step out of the bracket with |f| already near rounding is noise
      if (fabs(f) <= 1e3 * eps * mag) break;
      tn = 0.5 * (lo + hi);
    } else if (fabs(tn - tau) <= 4.0 * eps * fabs(tau)) {
      tau = tn;
      break;
    }
    tau = tn;
  }
  *org_out = o;
  *tau_out = tau;
  return it;
}

// Block-wide sum and maximum of one value per thread; every thread gets the
// result.  The block type supplies them (wave shuffles on the GPU, a serial
// loop on host threads); both use a fixed order.
template <class Blk>
__host__ __device__ inline double arrow_block_sum(const Blk &B, double *red, double v) {
  return B.sum(red, v);
}

template <class Blk>
__host__ __device__ inline double arrow_block_max(const Blk &B, double *red, double v) {
  return B.max(red, v);
}

// N block-wide sums at once (each in the order arrow_block_sum uses).
template <int N, class Blk>
__host__ __device__ inline void arrow_block_sums(const Blk &B, double *red, double *v) {
  B.template sums<N>(red, v);
}

// The whole channel: poles, nI arrowhead updates, split at the Fermi energy,
// descending sort per side and the Calc_En partial sums (DysonMatrix.cpp:
// 436-484, 542-594).  ws holds arrow_ws_bytes(J.n, J.nI) bytes; NIMAX >= J.nI
// bounds the per-thread accumulators.
template <int NIMAX, class Blk>
__host__ __device__ void arrow_body(const Blk &B, const ArrowJob &J, double *ws) {
  const int nI = J.nI, N = J.n, tid = B.tid(), nt = B.nt();
  const int m = J.np[0] + J.np[1];
  // amplitude buffers A and B, column-major: column c holds its nI H rows
  // (the eigenvector amplitudes on the orbits), padded to H2, then its nI G
  // rows (the couplings still to be added), padded likewise -- S contiguous
  // doubles, so that the eigenvector sums read a column in 16-byte pairs
  const int H2 = arrow_h2(nI);
  const long S = 2 * H2;
  double *dA = ws, *dB = dA + N;
  double *HGA = dB + N, *HGB = HGA + S * N;
  double *dn = HGB + S * N, *zn = dn + N, *tau = zn + N, *zh = tau + N, *red = zh + N;
  int *org = (int *)(red + kArrowRed), *nd = org + N, *dl = nd + N, *cnt = dl + N;
  const double eps = 2.2204460492503131e-16;

  if (tid == 0) cnt[2] = 0;
  // phase timers (profiling only): 0 load and sort, 7 norms, 1 deflation,
  // 2 roots, 3 Loewner, 4 eigenvectors, 5 reorder, 6 split and output
  unsigned long long tk = B.lap(-1, 0);
  // algorithmic FLOPs of this thread's share: 9 per pole per secular-equation
  // evaluation, 4 per Loewner factor, and 5 + 2 per carried row per pole for
  // each eigenvector
  double flops = 0.0;
  // poles (unsorted) into the B buffers
  for (int i = tid; i < m; i += nt) {
    const int s = i < J.np[0] ? 0 : 1, ii = s ? i - J.np[0] : i;
    dB[i] = J.E[s][(long)ii * J.es[s]];
    double *cb = HGB + i * S;
    for (int r = 0; r < S; ++r) cb[r] = 0.0;
    for (int r = 0; r < nI; ++r) cb[H2 + r] = J.C[s][(long)ii * J.ci[s] + (long)r * J.cr[s]];
  }
  B.sync();
  // ascending by energy, ties by position, into the A buffers
  for (int i = tid; i < m; i += nt) {
    const double v = dB[i];
    int rk = 0;
#pragma unroll 4
    for (int k = 0; k < m; ++k) { const double w = dB[k]; rk += (w < v) || (w == v && k < i); }
    dA[rk] = v;
    for (int r = 0; r < S; ++r) HGA[rk * S + r] = HGB[i * S + r];
  }
  B.sync();

  int g = m;
  tk = B.lap(0, tk);
  for (int k = 0; k < nI; ++k) {
    const double alpha = J.u1[k * nI + k] + J.h[k * nI + k];
    // z[i * S]: coupling k of column i
    double *z = HGA + H2 + k;
    double s2 = 0.0, dmx = 0.0;
    for (int i = tid; i < g; i += nt) { s2 += z[i * S] * z[i * S]; dmx = fmax(dmx, fabs(dA[i])); }
    const double znorm = sqrt(arrow_block_sum(B, red, s2));
    const double dmax = arrow_block_max(B, red, dmx);
    const double tol = 8.0 * eps * fmax(fmax(fabs(alpha), dmax), znorm);
    tk = B.lap(7, tk);
    // deflation, in energy order: couplings below tol leave at once (kept
    // indices compacted to nd, the others to dl, by a prefix scan); then
    // neighbouring kept directions whose energies differ by less than tol
    // after moving the whole coupling onto one of them merge.  Merges chain
    // (a merge changes the next pair's values), so they run on one thread,
    // but only when some neighbouring pair qualifies on the original values
    // -- if none does, the chain makes no merge either.
    {
      int kept = 0, gone = 0;
      for (int b0 = 0; b0 < g; b0 += nt) {
        const int i = b0 + tid;
        const int keep = (i < g && fabs(z[i * S]) > tol) ? 1 : 0;
        int tot = 0;
        const int pre = B.scan01((int *)red, keep, tot);
        if (i < g) {
          if (keep) nd[kept + pre] = i;
          else dl[gone + (tid - pre)] = i;
        }
        kept += tot;
        gone += (g - b0 < nt ? g - b0 : nt) - tot;
      }
      if (tid == 0) { cnt[0] = kept; cnt[1] = gone; cnt[3] = 0; }
      B.sync();
      for (int a = tid; a + 1 < kept; a += nt) {
        const int i0 = nd[a], i1 = nd[a + 1];
        const double za = z[i0 * S], zb = z[i1 * S], t = hypot(za, zb);
        if (fabs((dA[i1] - dA[i0]) * (za / t) * (zb / t)) <= tol) cnt[3] = 1;
      }
      B.sync();
      if (cnt[3] && tid == 0) {
        int p = 0, q = gone;
        for (int a = 0; a < kept; ++a) {
          const int i = nd[a], prev = p > 0 ? nd[p - 1] : -1;
          if (prev >= 0) {
            const double za = z[prev * S], zb = z[i * S], t = hypot(za, zb), c = za / t, sn = zb / t;
            if (fabs((dA[i] - dA[prev]) * c * sn) <= tol) {
              // rotate (prev, i): i carries the whole coupling, prev leaves
              double *cp = HGA + prev * S, *cq = HGA + i * S;
              for (int r = 0; r < nI; ++r) {
                const double x = cp[r], y = cq[r];
                cp[r] = sn * x - c * y;
                cq[r] = c * x + sn * y;
              }
              for (int r = k + 1; r < nI; ++r) {
                const double x = cp[H2 + r], y = cq[H2 + r];
                cp[H2 + r] = sn * x - c * y;
                cq[H2 + r] = c * x + sn * y;
              }
              const double da = dA[prev], db = dA[i];
              dA[prev] = sn * sn * da + c * c * db;
              dA[i] = c * c * da + sn * sn * db;
              z[prev * S] = 0.0;
              z[i * S] = t;
              --p;
              dl[q++] = prev;
            }
          }
          nd[p++] = i;
        }
        cnt[0] = p;
        cnt[1] = q;
      }
      B.sync();
    }
    const int p = cnt[0], q = cnt[1];
    // zh holds the squared couplings for the roots until the Loewner step
    // overwrites it
    for (int i = tid; i < p; i += nt) { dn[i] = dA[nd[i]]; zn[i] = z[nd[i] * S]; zh[i] = zn[i] * zn[i]; }
    B.sync();
    tk = B.lap(1, tk);
    // roots, started from the previous sweep's where the problem kept its size
    const int pw = J.warm_p ? J.warm_p[k] : -1;
    for (int j = tid; j <= p; j += nt) {
      if (p == 0) { org[j] = -1; tau[j] = 0.0; continue; }
      double *wj = J.warm ? J.warm + (long)k * J.wld + j : nullptr;
      const double guess = (wj && pw == p) ? *wj : __builtin_nan("");
      const int its = arrow_root(B, j, p, dn, zh, alpha, znorm, guess, &org[j], &tau[j]);
      if (wj) *wj = dn[org[j]] + tau[j];
      if (its >= kArrowMaxIt) cnt[2] = 1;
      flops += 9.0 * p * (its + 1);
    }
    B.sync();
    if (tid == 0 && J.warm_p) J.warm_p[k] = p;
    tk = B.lap(2, tk);
    // couplings recomputed from the roots: zh_i^2 = -prod_j (d_i - x_j) /
    // prod_{l != i} (d_i - d_l), paired along the interlacing x_i < d_i < x_{i+1}
    for (int i = tid; i < p; i += nt) {
      const double di = dn[i];
      double v = -((di - dn[org[i]]) - tau[i]) * ((di - dn[org[i + 1]]) - tau[i + 1]);
      // two interleaved partial products (pairs of ratios near 1, no overflow)
      double v1 = 1.0;
      {
        int jj = 0;
#pragma unroll 2
        for (; jj + 1 < i; jj += 2) {
          v *= ((di - dn[org[jj]]) - tau[jj]) * B.rcp(di - dn[jj]);
          v1 *= ((di - dn[org[jj + 1]]) - tau[jj + 1]) * B.rcp(di - dn[jj + 1]);
        }
        if (jj < i) v *= ((di - dn[org[jj]]) - tau[jj]) * B.rcp(di - dn[jj]);
        jj = i + 2;
#pragma unroll 2
        for (; jj + 1 <= p; jj += 2) {
          v *= ((di - dn[org[jj]]) - tau[jj]) * B.rcp(di - dn[jj - 1]);
          v1 *= ((di - dn[org[jj + 1]]) - tau[jj + 1]) * B.rcp(di - dn[jj]);
        }
        if (jj <= p) v *= ((di - dn[org[jj]]) - tau[jj]) * B.rcp(di - dn[jj - 1]);
      }
      v *= v1;
      zh[i] = copysign(sqrt(fmax(v, 0.0)), zn[i]);
      flops += 4.0 * (p + 1);
    }
    B.sync();
    tk = B.lap(3, tk);
    // eigenvectors: roots to positions [0, p], deflated directions to [p+1, p+q]
    for (int j = tid; j <= p; j += nt) {
      double ah[NIMAX], ag[NIMAX];
#pragma unroll
      for (int r = 0; r < NIMAX; ++r) {
        ah[r] = (r == k) ? 1.0 : 0.0;
        ag[r] = (r > k && r < nI) ? J.u1[r * nI + k] + J.h[r * nI + k] : 0.0;
      }
      double lam = alpha, nu2 = 1.0;
      if (p > 0) {
        const int o = org[j];
        const double t = tau[j], dofs = dn[o];
        lam = dofs + t;
#pragma unroll 2
        for (int i = 0; i < p; ++i) {
          const double w = zh[i] * B.rcp(t - (dn[i] - dofs));
          nu2 += w * w;
          const double *col = HGA + nd[i] * S;
          // pairs of rows; a pad row holds 0 and rows of G up to k are not
          // used, so the pair's other entry may be added freely
#pragma unroll
          for (int q = 0; q < NIMAX; q += 2) {
            if (q < nI) {
              const double h0 = col[q], h1 = col[q + 1];
              ah[q] += w * h0;
              ah[q + 1] += w * h1;
              if (q + 1 > k) {
                const double g0 = col[H2 + q], g1 = col[H2 + q + 1];
                ag[q] += w * g0;
                ag[q + 1] += w * g1;
              }
            }
          }
        }
      }
      const double inv = 1.0 / sqrt(nu2);
      flops += (5.0 + 2.0 * (2 * nI - k - 1)) * p;
      dB[j] = lam;
      double *cb = HGB + j * S;
#pragma unroll
      for (int r = 0; r < NIMAX; ++r) {
        if (r < H2) {
          cb[r] = r < nI ? ah[r] * inv : 0.0;
          cb[H2 + r] = (r > k && r < nI) ? ag[r] * inv : 0.0;
        }
      }
    }
    for (int t = tid; t < q; t += nt) {
      const long c = dl[t], pos = p + 1 + t;
      dB[pos] = dA[c];
      for (int r = 0; r < S; ++r) HGB[pos * S + r] = HGA[c * S + r];
    }
    B.sync();
    tk = B.lap(4, tk);
    g += 1;
    for (int i = tid; i < g; i += nt) {
      const double v = dB[i];
      int rk = 0;
#pragma unroll 4
      for (int l = 0; l < g; ++l) { const double w = dB[l]; rk += (w < v) || (w == v && l < i); }
      dA[rk] = v;
      for (int r = 0; r < S; ++r) HGA[rk * S + r] = HGB[i * S + r];
    }
    B.sync();
    tk = B.lap(5, tk);
  }

  tk = B.lap(6, tk);
  // dA ascending, amplitudes in the H rows of HGA: split at E_F, both sides
  // descending
  const double ef = J.efermi;
  int c_le = 0;
  for (int i = tid; i < N; i += nt) c_le += (dA[i] <= ef);
  const int nqh = (int)(arrow_block_sum(B, red, (double)c_le) + 0.5), nqp = N - nqh;
  double kb[9];
  for (int b = 0; b < 9; ++b) kb[b] = 0.0;
  for (int i = tid; i < N; i += nt) {
    const double di = dA[i];
    const long slot = di > ef ? (N - 1 - i) : nqp + (nqh - 1 - i);
    J.out_e[slot] = di;
    double x1 = 0.0;
    for (int r = 0; r < nI; ++r) {
      const double xr = HGA[i * S + r];
      J.out_x[slot * nI + r] = xr;
      x1 += xr * xr;
    }
    if (di <= ef) {
      double x2 = 0.0;
      for (int r = 0; r < nI; ++r)
        for (int c = 0; c < nI; ++c) x2 += J.u1[r * nI + c] * HGA[i * S + r] * HGA[i * S + c];
      const int bk = x1 > 0.01 ? 0 : (x1 > 1.e-5 ? 1 : 2);
      kb[3 * bk] += x2; kb[3 * bk + 1] += x1 * di; kb[3 * bk + 2] += x1;
    }
  }
  double tail[10];
  for (int b = 0; b < 9; ++b) tail[b] = kb[b];
  tail[9] = flops;
  arrow_block_sums<10>(B, red, tail);
  for (int b = 0; b < 9; ++b) kb[b] = tail[b];
  flops = tail[9];
  if (tid == 0) {
    *J.out_nqp = nqp;
    *J.out_nqh = nqh;
    const double f = (double)J.j2 + 1.0;
    J.kolt[0] = (kb[0] + kb[3] + kb[6]) * f;
    J.kolt[1] = (kb[1] + kb[4] + kb[7]) * f;
    J.kolt[2] = (kb[2] + kb[5] + kb[8]) * f;
    if (cnt[2]) B.fail();
    B.add_work(flops);
  }
  B.lap(6, tk);
}


// The Dyson problem of a channel without dynamic self-energy (the HF stage):
// the Dyson matrix is the static part u1 + h alone (lower triangle read, as
// arrow_body reads it).  One thread: cyclic Jacobi rotations (Golub and Van
// Loan, Matrix Computations, 4th ed., 8.5.2) on the matrix held in registers
// until the off-diagonal part is below 1e-17 of the matrix norm, then the
// records, counts and Calc_En sums arrow_body writes for such a channel.  Each
// eigenvector is signed so that its largest component (the first of equal
// ones) is positive.  N is the compiled bound of nI.  Returns 0, or 1 when nI
// lies outside [1, N] or the rotations do not converge.
constexpr int kStaticNI = 4;
constexpr int kStaticSweeps = 32;
template <int N>
__host__ __device__ inline int static_dyson(int nI, int j2, const double *h, const double *u1, double ef,
                                            double *out_e, double *out_x, int *out_nqp, int *out_nqh,
                                            double *kolt) {
  if (nI < 1 || nI > N) return 1;
  double A[N][N], V[N][N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      A[r][c] = 0.0;
      V[r][c] = r == c ? 1.0 : 0.0;
    }
  double nrm = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c)
      if (r < nI) {
        const double v = u1[r * nI + c] + h[r * nI + c];
        A[r][c] = v;
        A[c][r] = v;
        nrm += (r == c ? 1.0 : 2.0) * v * v;
      }
  bool done = false;
  for (int sw = 0; sw < kStaticSweeps && !done; ++sw) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += A[p][q] * A[p][q];
    if (!(off > 1e-34 * nrm)) { done = true; break; }
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double apq = A[p][q];
        if (q >= nI || apq == 0.0) continue;
        // rotation J in the (p, q) plane with (J^T A J)_pq = 0
        const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        A[p][q] = 0.0;
        A[q][p] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  if (!done) return 1;
  double d[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    d[i] = A[i][i];
    // sign: largest component positive
    double big = 0.0, sg = 1.0;
#pragma unroll
    for (int r = 0; r < N; ++r)
      if (r < nI && fabs(V[r][i]) > big) { big = fabs(V[r][i]); sg = V[r][i] < 0.0 ? -1.0 : 1.0; }
#pragma unroll
    for (int r = 0; r < N; ++r) V[r][i] *= sg;
  }
  int nqh = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) nqh += (i < nI && d[i] <= ef) ? 1 : 0;
  const int nqp = nI - nqh;
  double kb[9];
#pragma unroll
  for (int b = 0; b < 9; ++b) kb[b] = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i >= nI) continue;
    // ascending rank, ties by position; both sides stored descending
    int rk = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) rk += (k < nI && (d[k] < d[i] || (d[k] == d[i] && k < i))) ? 1 : 0;
    const int slot = d[i] > ef ? (nI - 1 - rk) : nqp + (nqh - 1 - rk);
    out_e[slot] = d[i];
    double x1 = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r)
      if (r < nI) {
        out_x[slot * nI + r] = V[r][i];
        x1 += V[r][i] * V[r][i];
      }
    if (d[i] <= ef) {
      double x2 = 0.0;
#pragma unroll
      for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c)
          if (r < nI && c < nI) x2 += u1[r * nI + c] * V[r][i] * V[c][i];
      const int bk = x1 > 0.01 ? 0 : (x1 > 1.e-5 ? 1 : 2);
      kb[3 * bk] += x2;
      kb[3 * bk + 1] += x1 * d[i];
      kb[3 * bk + 2] += x1;
    }
  }
  *out_nqp = nqp;
  *out_nqh = nqh;
  const double f = (double)j2 + 1.0;
  kolt[0] = (kb[0] + kb[3] + kb[6]) * f;
  kolt[1] = (kb[1] + kb[4] + kb[7]) * f;
  kolt[2] = (kb[2] + kb[5] + kb[8]) * f;
  return 0;
}

}  // namespace bcd
