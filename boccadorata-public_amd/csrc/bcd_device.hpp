// bcd_device.hpp -- device-side data layout of the SCGF hot path.
//
// One batch holds several independent systems ("instances").  Every subshell
// of every instance is a "channel" (global index gc).  All device buffers are
// flat; the structs below carry offsets into them and are built on the host
// once per batch.
#pragma once

#include <hip/hip_runtime.h>

namespace bcd {

constexpr int kKeyAB = 1024;
constexpr int kKeyJ = 128;
constexpr double kSqrt2 = 1.4142135623730950488016887;

// Static tables of one instance (model space + dense interaction blocks).
struct InstDev {
  int nsubsh, norb, nspb, jmax, nchan2b;
  int gc0;                       // first global channel of this instance
  int nq;                        // length of the flattened density rho (sum_v no_v^2)
  int _pad;
  const int *l, *j2, *ip, *ch, *no, *orb0, *orb_sh;   // [nsubsh] / [norb]
  const int *qoff;               // [nsubsh] offset of rho_v inside rho
  const int *pidx;               // [(jmax+1)*norb*norb]: (chan2b << 16) | row, or -1
  const long *voff;              // [nchan2b] start of each dense J-block
  const int *npair;              // [nchan2b]
  const double *V;               // dense symmetric blocks, row-major
  long rho_off;                  // start of this instance's rho in the density buffer
};

// <ab;J|V|cd;J> from the dense J-blocks, with the exchange phase
// (-1)^(ja+jb-J+1) for each pair given in reverse order and 0 for pairs that
// are not stored (VppInt_t::get, Vpp_potential.cpp:557-607).
__device__ inline double vget(const InstDev &I, int a, int b, int c, int d, int J) {
  int ph = 0;
  if (J > I.jmax) return 0.0;
  if (a > b) { int t = a; a = b; b = t; ph += (I.j2[I.orb_sh[a]] + I.j2[I.orb_sh[b]]) / 2 - J + 1; }
  if (c > d) { int t = c; c = d; d = t; ph += (I.j2[I.orb_sh[c]] + I.j2[I.orb_sh[d]]) / 2 - J + 1; }
  const long base = (long)J * I.norb;
  const int pa = I.pidx[(base + a) * I.norb + b];
  const int pc = I.pidx[(base + c) * I.norb + d];
  if (pa < 0 || pc < 0) return 0.0;
  const int ca = pa >> 16, cc = pc >> 16;
  if (ca != cc) return 0.0;
  const double v = I.V[I.voff[ca] + (long)(pa & 0xffff) * I.npair[ca] + (pc & 0xffff)];
  return (abs(ph) & 1) ? -v : v;
}

// One channel (subshell) of one instance.
struct ChanDev {
  int inst, ish, nI, j2, ch, orb0;
  int cap;                       // propagator slots of this channel (Dyson dimension bound)
  int blk_cnt;                   // Sigma(2) blocks of this channel
  long blk_off;
  long u1_off, h_off;            // nI*nI each
  long w_off;                    // W rows nI(nI+1)/2 x nq
  long rfw_off, rbk_off;         // unreduced pole rows, width 2+nI
  int rfw_cap, rbk_cap;
  long lfw_off, lbk_off;         // reduced rows, width (nI+1)+nI
  int lfw_cap, lbk_cap;          // rows
  long lws_fw_off, lws_bk_off;   // Lanczos vectors (lcap x M)
  long piv_off;                  // pivots [nI x nI]
  long tri_off;                  // d[cap], e[cap], Qhead[nI*cap]
  long mat_off;                  // dense Dyson matrix scratch (cap*cap) for the global-memory path
  long pole_off;                 // diagonalised reduced sides: lam_fw[lfw_cap], Ct_fw[nI x lfw_cap],
                                 // lam_bk[lbk_cap], Ct_bk[nI x lbk_cap]
  long aws_off;                  // arrowhead workspace (doubles) for the global-memory path
  long warm_off;                 // previous sweep's arrowhead roots, nI x cap doubles
  int trip_off, ntrip;           // (a, b, g) triples of the Sigma(2) blocks: starts in the triple table
  int lanc_n;                    // Lanczos length of the channel's instance (read with lanc_kind)
  int ef_set;                    // 1: the Fermi energy of this subshell is forced to ef_force
  int merge;                     // 1: every (a, b, g) triple has at most kMergeMaxJ J blocks (k_merge_poles)
  int _pad2;
  double ef_force;               // (Set_external_EFermi, DysonMatrix.cpp:601-617)
};

// One (Jab, a <= b, g) block of the Sigma(2) build of channel gc; blocks of a
// channel are stored in the reference's loop order.
struct BlockDev {
  int gc, a, b, g, J, nJab, nJgd;
  int c2b;                       // two-body channel of the block's (alpha beta; J) pairs, -1: none
};

// The subshells a, b, g (below 256) and nJab of a block in one int, for the
// passes of the Sigma(2) plan, which read nothing else of it.
__host__ __device__ inline int block_pack(const BlockDev &B) {
  return B.a | (B.b << 8) | (B.g << 16) | ((B.nJab & 1) << 24);
}

// One non-empty Sigma(2) block in the compact work list, with everything the
// build needs before its first dependent load: the block, its channel and
// instance, the fragment counts of both sides (2p1h: qp(a), qp(b), qh(g);
// 2h1p: qh(a), qh(b), qp(g)) packed two per int, and the first row of each
// side in the channel's rows.
struct Sig2Work {
  int bidx, gc, inst;
  int n01, n23, n45;             // counts 0..5 as (lo | hi << 16)
  int off_fw, off_bk;
  int nvt;                       // entries of the block's V table, nI * pairs * n_g (its LDS size class)
  int _pad;
};

// Per-channel dynamic state (fragment counts live in separate arrays).
struct ChanState {
  int mfw, mbk;                  // unreduced pole counts (Sigma(2) build)
  int lfw, lbk;                  // reduced pole counts (Lanczos)
  int npiv_tot;                  // pivots handed to the Lanczos (nt_pvts)
  int den;                       // leading columns of the reduced rows
  int ipiv_fw[10], ipiv_bk[10];  // [n_pvt, 1, starts...]
  int itr_fw[8], itr_bk[8];      // iteration counts per pivot (after rescaling)
  int err;
  int cfw, cbk;                  // pole counts after merging the J-degenerate groups
  int lmfw, lmbk;                // pole counts the Lanczos jobs run on (merged or not)
  int _pad;
  double kolt[3];                // Calc_En partial sums of the last sweep
};

struct InstState;

// Per-instance dynamic state.
struct InstState {
  int cur;                       // propagator buffer holding the current iterate
  int active;                    // 1 while the current stage still iterates
  int its;                       // sweeps done in the current stage
  int itr_max;                   // ItrMax of the current stage (< 0: unbounded)
  int hf_its, sc_its;
  int err;
  int mix_k;                     // density-mixing steps of the current stage
  int mix_s;                     // first iterate of the current mixing history
  int failed;                    // 1: dropped from this solve (a channel exceeds a kernel limit)
  double mix_fn;                 // norm of the previous step's density residual
  double conv;
  double diff;
  double ef[2];                  // Fermi energy per charge for the next sweep
  double e_kolt, u1_tot, a_tot;
};

// Fermi energy a Dyson sweep of channel C splits at: the forced value, or
// the midpoint k_finalize took from the sweep's input propagator
// (Seek_EFermi, DysonMatrix.cpp:619-635).
__device__ inline double chan_efermi(const ChanDev &C, const InstState *S) {
  return C.ef_set ? C.ef_force : (S ? S->ef[C.ch & 1] : 0.0);
}

}  // namespace bcd

namespace bcd {

// One Lanczos reduction (Reduce_Mtx_Lncz) of the pole rows of one side.
struct LancJob {
  const double *Min;             // [M x Tin], column 0 = E, columns NA_in.. = couplings
  int Tin, NA_in, NB, M;
  const double *piv;             // [nt_pvts x NB] pivots (normalised by the kernel)
  int nt_pvts, nLanc;            // nLanc: vectors to build (already clamped)
  const int *itr;                // iterations per pivot [nt_pvts]
  double *Mout;                  // [nLanc x Tout]
  int Tout, NA_out;
  int *i_piv_out;                // n_pvt + 2 entries
  int *n_out;                    // rows produced, or < 0 on error
  double *V;                     // workspace [nLanc x M]
  int inst;                      // instance whose active flag gates the job (-1: none)
  int _pad;
};

// One Dyson eigenproblem of one channel.
struct DysonJob {
  int nI, j2, nfw, nbk, den_fw, den_bk, npiv_fw, npiv_bk;
  int n, inst;                   // dimension; gating instance (-1: none)
  int ch;
  int path;                      // pole blocks: 0 one-wave kernels (L <= 64, nI <= 8), 1 the general ones
  double efermi;                 // used when inst < 0
  const double *h, *u1;          // static part = u1 + h, both nI x nI
  const double *rfw, *rbk;       // pole rows
  const int *ipiv_fw, *ipiv_bk;
  double *tri;                   // d[n], e[n], Qhead[nI*n]
  double *mat;                   // n x n scratch for the global-memory path
  double *out_e, *out_x;         // records sorted: qp (descending) then qh (descending)
  int *out_nqp, *out_nqh;
  double *kolt;                  // 3 Calc_En partial sums
  // poles != 0: diagonalise one reduced side alone, D = U diag(lam) U^T (its
  // fishbone block in the fw slot, no static part).  The tracked rows start
  // as the couplings M^T, so they end as M^T U: out_lam[k] = lam_k and
  // out_ct[r * ct_ld + k] = (U^T M)[k][r], the coupling of pole k to orbit r.
  int poles, ct_ld;
  double *out_lam, *out_ct;
};

}  // namespace bcd

namespace bcd {

// Read access to one propagator buffer pair: fragments of channel gc are
// records [0, nqp) (quasiparticles, descending e) then [nqp, nqp+nqh)
// (quasiholes, descending e) at e[e_off[gc]..], x[x_off[gc] + rec*nI..].
// Which of the two buffers is current is per instance (InstState::cur);
// without instance state buffer 0 is used.
struct PropView {
  double *e[2], *x[2];
  int *nqp[2], *nqh[2];
  const long *e_off, *x_off;
  const int *cap;
  InstState *is;
};

}  // namespace bcd
