// bcd_twobody.hpp -- two-body propagator matrices on the GPU: the pp/hh
// ladder of gII_prop_t (gII_b.cpp:62-563) and the ph ring of Pol_prop_t
// (PolProp_b.cpp:43-614), as used by CoupledCluster_t (CCD_a.cpp:158-240).
//
// Each matrix element of the reference is a fourfold orbit sum
//   x(r, c) = sum_{ia ib ig id} A_r(ia, ib) K_J(ia ib, ig id) A_c(ig, id)
// with A_r the product of the two fragment amplitudes of basis entry r and
// K_J the pp interaction (ladder) or its Pandya transform (ring).  Here the
// orbit pairs of one two-body channel are indexed once (p), so a whole block
// of the matrix is  T_rows K T_cols^T  with T[r][p] the folded amplitudes:
// two FP64 MFMA GEMMs per block (G = T K, then G T^T), with the reference's
// signs, 1/sqrt2 factors and block placement applied in the second GEMM's
// epilogue.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "bcd_device.hpp"
#include "bcd_system.hpp"

namespace bcd {

// Kernel families of one build, in launch order.
constexpr int kT2Kernels = 6;   // zero, amp, pandya, gemm1, gemm2, diag

// Fragments of one propagator on the device (SpProp_t fqp/fqh rows, N_SPB_MX wide).
struct FragDev {
  const double *fq[2];            // [0] quasiparticles, [1] quasiholes: [n x nspb]
  const int *sub[2];              // subshell of each fragment (n_clj_p / n_clj_h)
  int nspb, _pad;
};

// T[r][p] for the rows of one basis block and the pairs of one channel.
//   kind 0: ladder, p = canonical pair a <= b of a dense V block; both
//           orderings of the orbits fold into p with the exchange phase,
//           equal orbits carry sqrt2.
//   kind 1: ring, p = ordered pair (i, j): X_u(i) Y_w(j).
//   kind 2: ring, the same amplitudes at the swapped pair (j, i).
struct AmpTask {
  int kind, J, nrows, npair, ld, _pad;
  int ukind, wkind;               // 0 = quasiparticle, 1 = quasihole
  const int *u, *w;               // fragments of each row
  const int *pair;                // (a << 16) | b, [npair]
  double *T;                      // [nrows x ld]
};

// Pandya transform of the pp interaction over one ordered pair list:
// K[p][q] = -sum_J' (2J'+1) <i l;J'|V|k j;J'> {ji jj J; jk jl J'} * sqrt2 (i == l) * sqrt2 (k == j)
// for p = (i, j), q = (k, l) (PolProp_b.cpp:569-614).
struct PandyaTask {
  int J, npair, ld, slot;         // slot of the (2J'+1)*6j table for this J
  const int *pair;
  double *K;                      // [npair x ld]
};

// C = A B^T (A [M x K], B [N x K], both K-contiguous).  M, N are multiples
// of the block tile and K of the K step; the operands are zero-padded to
// them, so the loads need no bounds.  Mv, Nv bound what is written.  With
// C != nullptr a plain store; otherwise the epilogue: y = x3[c] * C(r, c), then y /= sqrt2
// for rs[r] and for cs[c], written as sgn * y at entry (Cidx, Ridx) =
// (c0 + c, r0 + r), or (c0 + r, r0 + c) when transposed, of the output
// matrix (stored as out[Cidx * ld + Ridx]); up to two writes per element.
// sym: the product is symmetric (B rows = A's rows through the same
// symmetric K); only tiles with tile_n >= tile_m run, and each element with
// r <= c is also written at the mirrored position (c, r).
struct GemmTask {
  int M, N, K, lda, ldb, ldc;
  int Mv, Nv, sym, _pad;
  const double *A, *B;
  double *C;
  double *out;
  int ld, nw;
  int c0[2], r0[2], tr[2], sgn[2];
  const int *rs, *cs;
  const double *x3;
};

// out[k * (ld + 1)] += sgn * e[k - k0] for k in [k0, k0 + n).
struct DiagTask {
  double *out;
  const double *e;
  int ld, k0, n, sgn;
};

struct ZeroTask {
  double *out;
  long n;
};

// One matrix of the plan: its channel, mode and basis, as the reference's
// objects hold them.
struct T2Job {
  int kind;                       // 0 ladder (gII_prop_t), 1 ring (Pol_prop_t)
  int J, pi, dT, mode, add;
  std::vector<int> a1, a2, b1, b2;          // ladder: n1_pp n2_pp k1_hh k2_hh; ring: n_ph k_ph n_hp k_hp
  std::vector<double> ea, eb;               // e_pp e_hh / e_ph e_hp
  int nf = 0, nb = 0, ld = 0;               // block dimensions of the filled matrix, leading dimension
  long out_off = 0;                         // first double of the matrix in the plan's output buffer
  std::vector<double> free_poles;           // what Fill_*_mtx writes to Free_poles (ndim entries, NaN = untouched)
};

// Fragment subshells and energies on the host (basis enumeration).
struct HostFrags {
  int n[2] = {0, 0};               // quasiparticles, quasiholes
  std::vector<int> sub[2];
  std::vector<double> e[2];
};

// Device tables of one system for the two-body kernels.
struct T2System {
  InstDev I{};                     // model space + dense V blocks (device pointers)
  std::vector<void *> bufs;
  std::vector<int> chan_of;        // (J * 2 + parity) * 8 + charge -> dense V block, or -1
  std::vector<std::vector<int>> pairs;   // per V block: (a << 16) | b in row order
  int *d_pairs = nullptr;          // all V-block pair lists, back to back
  std::vector<long> pair_off;
  int nj = 0;                      // distinct 2j values
  std::vector<int> jidx;           // 2j -> index
  int *d_jidx = nullptr;
  int upload(const HostSystem &S);
  void release();
  ~T2System() { release(); }
};

// A set of matrices built together (one launch per kernel for all of them).
struct T2Plan {
  std::vector<T2Job> jobs;
  double *d_out = nullptr;         // all matrices
  long out_len = 0;
  std::vector<void *> bufs;
  std::vector<AmpTask> amp;
  std::vector<PandyaTask> pan;
  std::vector<GemmTask> g1, g2;
  std::vector<DiagTask> diag;
  std::vector<ZeroTask> zero;
  std::vector<int> sixj_J;         // J of each 6j-table slot
  struct VCopy { double *dst; const double *src; int P, ld; };
  std::vector<VCopy> vcopy;        // dense V blocks copied into zero-padded GEMM operands
  double *d_sixj = nullptr;
  // device copies of the task lists and their launch tables
  AmpTask *d_amp = nullptr; int *d_amp_rows = nullptr; long amp_rows = 0;
  PandyaTask *d_pan = nullptr; int *d_pan_rows = nullptr; long pan_rows = 0;
  GemmTask *d_g1 = nullptr, *d_g2 = nullptr;
  int *d_t1 = nullptr, *d_t2 = nullptr; long n_t1 = 0, n_t2 = 0;   // 64-tiles first, then n_t1b / n_t2b 128-tiles
  long n_t1b = 0, n_t2b = 0;
  DiagTask *d_diag = nullptr;
  ZeroTask *d_zero = nullptr;
  int2 *d_zero_chunks = nullptr; long zero_chunks = 0;
  double flops_g1 = 0.0;           // 2*M*N*K of the G = T K products (unpadded)
  double flops_g2 = 0.0;           // 2*M*N*K of the G T^T products (unpadded)
  double out_doubles = 0.0;        // matrix entries the epilogues write
  ~T2Plan() { release(); }
  void release();
  // Allocates everything and stages the task lists of the jobs (whose bases
  // index the fragments F).
  int prepare(const T2System &S, const HostSystem &H, const HostFrags &F, hipStream_t st);
  // Enqueues the whole build on st.  With ev (kT2Kernels + 1 events) an event
  // is recorded before each kernel family and after the last one.
  int launch(const T2System &S, const FragDev &frag, hipStream_t st, hipEvent_t *ev = nullptr) const;
};

// Basis enumeration (host), as Build_LaddRPA_basis / Build_phRPA_basis.
int t2_ladd_basis(const HostSystem &S, const HostFrags &F, int J, int dT, int pi, int icut, int ntcut,
                  const int *n1p, const int *n1h, T2Job &job);
int t2_ph_basis(const HostSystem &S, const HostFrags &F, int J, int dT, int pi, int icut, int ntcut,
                const int *n1p, const int *n1h, T2Job &job);
// Sets nf, nb and free_poles of a job from its mode (Fill_*_mtx bookkeeping);
// returns the matrix dimension or a negative reference error code.
int t2_job_dims(T2Job &job);

}  // namespace bcd
