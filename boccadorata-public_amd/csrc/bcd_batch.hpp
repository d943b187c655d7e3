// bcd_batch.hpp -- a device-resident batch of SCGF instances.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "bcd_system.hpp"

namespace bcd {

struct Mlp;

constexpr int kLdsBytes = 160 * 1024 - 1024;   // dynamic LDS ceiling used by the Dyson kernels
constexpr int kMaxSweeps = 2000;
constexpr int kSweepChunk = 4;                 // sweeps queued between host reads of the convergence flags

// Starting propagator of one instance, fragments as SpProp_t stores them.
struct HostProp {
  int nspb = 0;
  std::vector<int> np, nh;
  std::vector<double> ep, fqp, eh, fqh;
};

struct DevArena {
  void *ptr = nullptr;
  size_t cap = 0;
  int grow(size_t bytes);
  void release();
};

enum KernelTimer {
  KT_SIGMA2, KT_LANCZOS, KT_TRIDIAG, KT_QL, KT_DENSITY, KT_SIGINF, KT_FINAL, KT_WBUILD, KT_PLAN, KT_ARROW, KT_POLES,
  KT_MERGE, KT_N
};

// Extra streams the Lanczos buckets run on beside the batch's own stream, so
// that the tail of one bucket's launch overlaps the others.
constexpr int kLancAux = 2;
// Streams the Lanczos buckets are spread over (the batch's own and the first
// kLancStreams - 1 auxiliary ones).  The device runs four hardware queues per
// process by default (GPU_MAX_HW_QUEUES); with more streams than the queues
// left beside the copy traffic, two streams share a queue and their buckets
// run one after the other.
constexpr int kLancStreams = 3;
// Launch order of the Lanczos buckets: by estimated time (true) or by job
// size, largest first (false).
constexpr bool kLancOrderByEst = true;
// The autotuner gives a bucket the shape of the next larger bucket when that
// shape is within this fraction of the bucket's fastest.
constexpr float kLancMergeTol = 0.05f;
static_assert(kLancStreams >= 1 && kLancStreams <= kLancAux + 1, "Lanczos streams exist");
// Estimates for spreading the Lanczos buckets over the streams when a bucket
// has no autotuned time: the device's streaming rate and one CU's (bytes per
// ms), and the pole blocks that follow each bucket (ms).
constexpr double kLancChipBytesPerMs = 5.0e9;
constexpr double kLancCuBytesPerMs = 1.1e7;
constexpr double kLancPoleMs = 5.0;
// Measured placement: after a forked Lanczos stage, the last launch of the
// stream that ended last moves to the stream that ended first when the
// measured times predict the stage to end at least kLancMoveGainMs earlier;
// at most kLancMoves such moves for one set of launches, then it is kept.
constexpr int kLancMoves = 2;
constexpr float kLancMoveGainMs = 1.0f;

struct Batch {
  struct PoleCfg {
    int lmax = 1;                 // largest reduced side (rows)
    int lw = 1, ldw = 1;          // largest side of the one-wave path and its LDS leading dimension
    size_t wlds = 0;              // LDS of k_tridiag_poles
    bool in_lds = false;          // the general path's largest side fits LDS
    bool any_wave = false, any_general = false;
  };
  std::vector<std::pair<int, int>> arrow_classes;   // (arrow_threads, arrow_ni_bound) present among the channels
  std::vector<int> hf_classes;      // arrow_ni_bound of the channels above kStaticNI orbits
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t lstream[kLancAux] = {};
  hipEvent_t lfork = nullptr, ljoin[kLancAux] = {};
  // Placement of the Lanczos launches over the streams, measured: lt0 at the
  // fork and lbt[i] after launch i's pole blocks on its stream give each
  // launch's end in the last forked stage (lp_timed).  lp_key: the launches'
  // first buckets in launch order; lp_used: the streams they ran on; lp_fixed:
  // the placement comes from a measurement; lp_frozen: no more measuring;
  // lp_moves: moves made for this set of launches (lanczos()).
  hipEvent_t lt0 = nullptr, lbt[kNLancBuckets] = {};
  std::vector<int> lp_key, lp_used, lp_best;   // lp_best: the placement that ended earliest
  float lp_best_end = 0.f;
  bool lp_timed = false, lp_fixed = false, lp_frozen = false;
  int lp_moves = 0;
  bcd_solve_params par{};
  std::vector<const HostSystem *> sys;
  std::vector<int> inst_lanc_n;      // Lanczos length per instance (empty: par.lanc_n for all); set before setup
  int set_efermi(int inst, int ish, double ef);
  int set_params(const bcd_solve_params &p);
  std::string err_detail;            // what the last solve reported (limits hit, instances dropped)
  void note_error(const std::string &what);
  int n_inst = 0, n_gc = 0, nImax = 1;
  long n_blk = 0, capmax = 0, mat_len = 0;
  std::vector<InstDev> h_insts;
  std::vector<ChanDev> h_chans;
  std::vector<BlockDev> h_blocks;
  std::vector<int> h_trip_start, h_trip_blk;
  std::vector<ChanState> h_cs;
  std::vector<InstState> h_is;
  std::vector<int> hf_its, sc_its, hf_ok, sc_ok;
  std::vector<long> n_fw_tot, n_bk_tot;

  DevArena a_tab_i, a_tab_l, a_tab_d, a_inst, a_chan, a_blocks, a_work, a_prop, a_rows, a_lws, a_mat, a_jobs, a_aws,
      a_prof, a_warm, a_mix;
  double *d_warm = nullptr;                          // previous sweep's arrowhead roots (warm start)
  double *d_mix = nullptr;                           // density-mixing history (accel > 0)
  long rho_len = 0;                                  // doubles of the density buffer
  int *d_warm_p = nullptr;
  InstDev *d_insts = nullptr;
  InstState *d_is = nullptr;
  ChanDev *d_chans = nullptr;
  ChanState *d_cs = nullptr;
  long *d_row_fw_base = nullptr, *d_row_bk_base = nullptr, *d_e_off = nullptr, *d_x_off = nullptr;
  int *d_cap = nullptr;
  BlockDev *d_blocks = nullptr;
  int *d_bpack = nullptr;                            // block_pack() of every block (Sigma(2) plan)
  int *d_blk_fw_off = nullptr, *d_blk_bk_off = nullptr, *d_work_count = nullptr;
  Sig2Work *d_work_list = nullptr;
  int *d_trip_start = nullptr, *d_trip_blk = nullptr;
  double *d_crows_fw = nullptr, *d_crows_bk = nullptr;   // J-merged pole rows (Lanczos path)
  int *d_status = nullptr;
  double *d_work = nullptr;                          // arrowhead FLOP counter of the current solve
  unsigned long long *d_prof = nullptr;              // arrowhead phase timers (profiling)
  double *d_u1 = nullptr, *d_h = nullptr, *d_w = nullptr, *d_rho = nullptr, *d_tri = nullptr, *d_piv = nullptr;
  double *d_lrows = nullptr, *d_kolt = nullptr, *d_rows_fw = nullptr, *d_rows_bk = nullptr, *d_poles = nullptr;
  long l_u1 = 0, l_h = 0, l_w = 0, l_rho = 0, l_tri = 0, l_piv = 0, l_lr = 0, l_kolt = 0, l_pole = 0, l_end = 0;
  double *pe_[3] = {nullptr, nullptr, nullptr}, *px_[3] = {nullptr, nullptr, nullptr};
  int *pnqp_[3] = {nullptr, nullptr, nullptr}, *pnqh_[3] = {nullptr, nullptr, nullptr};
  long prop_e_len = 0, prop_x_len = 0, prop_n_len = 0;
  DysonJob *d_djobs = nullptr, *d_pjobs = nullptr;   // per-sweep dense jobs; side diagonalisation jobs
  LancJob *d_ljobs = nullptr;
  int *d_lorder = nullptr;                           // Lanczos jobs, longest first
  int *d_cspack = nullptr;                           // kCsPack counts per channel (pull_counts)
  int *d_porder = nullptr;                           // channels by Sigma(2) block count, most first
  std::vector<int> h_cspack;
  long aws_len = 0;                                  // global arrowhead workspace (doubles), 0: LDS

  hipEvent_t ev[6] = {};
  double stage_ms[6] = {0, 0, 0, 0, 0, 0};
  bool profile = false;
  bool use_graphs = true;            // full sweep chunks replayed from a captured graph (off while profiling)
  std::pair<hipEvent_t, hipEvent_t> kev[KT_N] = {};
  double ktime[KT_N] = {};
  long kcount[KT_N] = {};

  ~Batch();
  std::vector<DevArena *> arenas();
  // sigma_only: the batch serves stage calls (Sigma(2), Sigma^inf) and never
  // runs a Dyson sweep, so no Dyson storage is sized for the poles.
  bool sigma_only = false;
  int setup(int dev, const std::vector<const HostSystem *> &systems, const bcd_solve_params &p,
            const std::vector<HostProp> &g0, bool sigma_only_ = false);
  PropView view(bool with_state) const;
  size_t tridiag_lds(int n, bool in_lds) const;
  void kt_begin(int k);
  void kt_end(int k);
  int sweep(int src);
  int check(int &n_active);
  int stage(int src, int itr_max, double conv, bool hf);
  int build_sigma2(int only_ish);
  int tune_bucket(int bk, int mmax, int nj, const int *ord, int prev);
  // Lanczos tile shape per pole-count bucket; tune_lanczos: pick them by
  // timing during the next solve
  int lanc_shape[kNLancBuckets] = {kLancDefault[0], kLancDefault[1], kLancDefault[2], kLancDefault[3],
                                   kLancDefault[4], kLancDefault[5], kLancDefault[6], kLancDefault[7],
                                   kLancDefault[8], kLancDefault[9]};
  static_assert(kNLancBuckets == 10, "every bucket has a default shape");
  float lanc_tune_ms[kNLancBuckets][kNLancShapes] = {};
  bool tune_lanczos = false;
  bool poles_done = false;       // lanczos() ran the pole-block kernels bucket by bucket
  bool w_built = false;          // d_w holds the Sigma^inf tensor of this batch (k_wbuild)
  int prepare();
  int work(double *w, int n);
  double coupling_flops = -1.0;
  int lanczos();
  int decompose_poles();
  int pull_counts();
  int pull_cs();
  PoleCfg pole_cfg() const;
  void launch_poles(const PoleCfg &c, const int *order, int njobs, hipStream_t s);
  void general_poles(const PoleCfg &c, hipStream_t s);
  // learned start and the fallback to the plain iteration
  Mlp *mlp = nullptr;                 // network seeding the sc0 iteration (not owned; nullptr: off)
  bool learn_record = false;          // compute the network's inputs and targets during solve()
  bool learn_valid = false;           // d_feat / d_target hold the last solve's
  bool seed_pending = false;          // the next sweep applies the predicted densities
  double fallback_tol = 1e-6;         // residual above which a seeded or mixed solve is redone plainly
  DevArena a_learn;
  float *d_feat = nullptr, *d_target = nullptr, *d_pred = nullptr, *d_inst_feat = nullptr;
  double *d_hf_e = nullptr, *d_hf_x = nullptr;
  int *d_hf_n = nullptr, *d_seed_on = nullptr, *d_flag = nullptr;
  std::vector<float> h_inst_feat;     // per instance: hbar*omega estimate, rms of V, Nmax, 0
  std::vector<int> seeded, fell_back, first_its;
  std::vector<double> resid;          // plain-sweep residual after a seeded or mixed stage (-1: not checked)
  std::vector<char> stage_mask;       // non-empty: the instances the next stage iterates
  bool learn_hook = false;            // the stage runs after_first_sweep() after its first sweep
  int after_first_sweep();
  int learn_alloc();
  int learn_features(bool features);   // false: the HF snapshot only
  int verify_and_fallback(int src);
  int solve();
  int solve_impl();
  int result(int inst, int nspb, int *np, int *nh, int *n_qp, int *n_qh, double *ep, double *zp, double *fqp,
             double *eh, double *zh, double *fqh);
};

}  // namespace bcd
