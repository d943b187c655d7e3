// bcd_batch.hip -- device layout and orchestration of the batched SCGF solve.
//
// The driver mirrors Run_itrs (Main.cpp:105-479) for the chain HF=-1 ->
// MakeSelfEn 2nd [Lanczos] -> ExtSE=-1: every stage is a short sequence of
// kernel launches over all channels of all instances on one stream; the host
// only checks the per-instance convergence flags between sweeps.
#include "bcd_hipsync.hpp"
#include "bcd_batch.hpp"
#include "bcd_learn.hpp"
#include "bcd_mlp.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace bcd {

std::mutex &hip_sync_lock() {
  static std::mutex m;
  return m;
}

static long pad8(long x) { return (x + 7) & ~7L; }

int DevArena::grow(size_t bytes) {
  if (bytes <= cap) return BCD_OK;
  if (ptr) (void)bcd::sync_free(ptr);
  ptr = nullptr;
  cap = 0;
  if (bcd::sync_malloc(&ptr, bytes) != hipSuccess) { ptr = nullptr; return BCD_ERR_NOMEM; }
  cap = bytes;
  return BCD_OK;
}

void DevArena::release() {
  if (ptr) (void)bcd::sync_free(ptr);
  ptr = nullptr;
  cap = 0;
}

Batch::~Batch() {
  for (auto *a : arenas()) a->release();
  for (auto &e : ev) { if (e) (void)hipEventDestroy(e); }
  for (auto &kv : kev) {
    if (kv.first) (void)hipEventDestroy(kv.first);
    if (kv.second) (void)hipEventDestroy(kv.second);
  }
  if (lfork) (void)hipEventDestroy(lfork);
  if (lt0) (void)hipEventDestroy(lt0);
  for (auto &e : lbt) { if (e) (void)hipEventDestroy(e); }
  for (int k = 0; k < kLancAux; ++k) {
    if (ljoin[k]) (void)hipEventDestroy(ljoin[k]);
    if (lstream[k]) (void)hipStreamDestroy(lstream[k]);
  }
  if (stream) (void)hipStreamDestroy(stream);
}

std::vector<DevArena *> Batch::arenas() {
  return {&a_tab_i, &a_tab_l, &a_tab_d, &a_inst, &a_chan, &a_blocks, &a_work, &a_prop, &a_rows,
          &a_lws, &a_mat, &a_jobs, &a_aws, &a_prof, &a_warm, &a_mix, &a_learn};
}

// The last HIP error a batch call of this thread met, with its source line;
// bcd_batch_solve appends it to the batch's error detail.
static thread_local std::string t_hip_error;

static void note_hip(hipError_t e, int line) {
  t_hip_error = std::string(hipGetErrorName(e)) + " at bcd_batch.hip:" + std::to_string(line);
}

#define HTRY(x)                                   \
  do {                                            \
    hipError_t e_ = (x);                          \
    if (e_ != hipSuccess) {                       \
      note_hip(e_, __LINE__);                     \
      return BCD_ERR_DEVICE;                      \
    }                                             \
  } while (0)

#define BTRY(x)                 \
  do {                          \
    int r_ = (x);               \
    if (r_ != BCD_OK) return r_; \
  } while (0)

int Batch::setup(int dev, const std::vector<const HostSystem *> &systems, const bcd_solve_params &p,
                 const std::vector<HostProp> &g0, bool sigma_only_) {
  device = dev;
  sigma_only = sigma_only_;
  par = p;
  sys = systems;
  n_inst = (int)sys.size();
  h_inst_feat.assign((size_t)kLearnInstFeat * n_inst, 0.f);
  for (int i = 0; i < n_inst; ++i) {
    h_inst_feat[(size_t)kLearnInstFeat * i + 0] = (float)sys[i]->hw_est;
    h_inst_feat[(size_t)kLearnInstFeat * i + 1] = (float)sys[i]->v_rms;
    h_inst_feat[(size_t)kLearnInstFeat * i + 2] = (float)sys[i]->nmax_est;
  }
  learn_valid = false;
  HTRY(hipSetDevice(device));
  // non-blocking streams: synchronous copies other host threads issue on the
  // null stream neither wait for this batch nor break its graph captures
  // (every synchronous copy of this batch follows a synchronisation of its
  // stream)
  HTRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  for (auto &e : ev) HTRY(hipEventCreate(&e));
  HTRY(hipEventCreateWithFlags(&lfork, hipEventDisableTiming));
  HTRY(hipEventCreate(&lt0));
  for (auto &e : lbt) HTRY(hipEventCreate(&e));
  for (int k = 0; k < kLancAux; ++k) {
    HTRY(hipStreamCreateWithFlags(&lstream[k], hipStreamNonBlocking));
    HTRY(hipEventCreateWithFlags(&ljoin[k], hipEventDisableTiming));
  }

  // ---- static tables, concatenated per element type
  std::vector<int> ti;
  std::vector<long> tl;
  std::vector<double> td;
  struct Off { long l, j2, ip, ch, no, orb0, orb_sh, qoff, pidx, npair, voff, V; };
  std::vector<Off> off(n_inst);
  auto put_i = [&](const std::vector<int> &v) { long o = (long)ti.size(); ti.insert(ti.end(), v.begin(), v.end()); ti.resize(pad8(ti.size())); return o; };
  long rho_tot = 0;
  for (int i = 0; i < n_inst; ++i) {
    const HostSystem &S = *sys[i];
    Off &o = off[i];
    o.l = put_i(S.l); o.j2 = put_i(S.j2); o.ip = put_i(S.ip); o.ch = put_i(S.ch); o.no = put_i(S.no);
    o.orb0 = put_i(S.orb0); o.orb_sh = put_i(S.orb_sh); o.qoff = put_i(S.qoff); o.pidx = put_i(S.pidx);
    o.npair = put_i(S.npair);
    o.voff = (long)tl.size(); tl.insert(tl.end(), S.voff.begin(), S.voff.end()); tl.resize(pad8(tl.size()));
    // an interaction already on the device (uploaded when the instance was
    // added) stays where it is
    o.V = -1;
    if (!S.dV) { o.V = (long)td.size(); td.insert(td.end(), S.V.begin(), S.V.end()); td.resize(pad8(td.size())); }
  }
  BTRY(a_tab_i.grow(sizeof(int) * std::max<size_t>(ti.size(), 1)));
  BTRY(a_tab_l.grow(sizeof(long) * std::max<size_t>(tl.size(), 1)));
  BTRY(a_tab_d.grow(sizeof(double) * std::max<size_t>(td.size(), 1)));
  HTRY(bcd::sync_memcpy(a_tab_i.ptr, ti.data(), sizeof(int) * ti.size(), hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(a_tab_l.ptr, tl.data(), sizeof(long) * tl.size(), hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(a_tab_d.ptr, td.data(), sizeof(double) * td.size(), hipMemcpyHostToDevice));
  const int *Ti = (const int *)a_tab_i.ptr;
  const long *Tl = (const long *)a_tab_l.ptr;
  const double *Td = (const double *)a_tab_d.ptr;

  // ---- instances and channels
  h_insts.resize(n_inst);
  h_chans.clear();
  h_blocks.clear();
  h_trip_start.clear();
  h_trip_blk.clear();
  nImax = 1;
  const bool dyn = par.mode >= 1;
  long o_u1 = 0, o_h = 0, o_w = 0, o_tri = 0, o_mat = 0, o_piv = 0, o_lr = 0, o_e = 0, o_x = 0, o_blk = 0;
  long o_pole = 0, o_aws = 0, o_warm = 0;
  std::vector<long> e_off, x_off;
  std::vector<int> cap;
  for (int i = 0; i < n_inst; ++i) {
    const HostSystem &S = *sys[i];
    const Off &o = off[i];
    InstDev &I = h_insts[i];
    I.nsubsh = S.nsubsh; I.norb = S.norb; I.nspb = S.nspb; I.jmax = S.jmax; I.nchan2b = S.nchan2b;
    I.gc0 = (int)h_chans.size(); I.nq = S.nq; I._pad = 0;
    I.l = Ti + o.l; I.j2 = Ti + o.j2; I.ip = Ti + o.ip; I.ch = Ti + o.ch; I.no = Ti + o.no;
    I.orb0 = Ti + o.orb0; I.orb_sh = Ti + o.orb_sh; I.qoff = Ti + o.qoff; I.pidx = Ti + o.pidx;
    I.voff = Tl + o.voff; I.npair = Ti + o.npair; I.V = S.dV ? S.dV : Td + o.V;
    I.rho_off = rho_tot;
    rho_tot += pad8(S.nq);
    for (int s = 0; s < S.nsubsh; ++s) {
      ChanDev C;
      std::memset(&C, 0, sizeof C);
      const int nI = S.no[s];
      nImax = std::max(nImax, nI);
      C.inst = i; C.ish = s; C.nI = nI; C.j2 = S.j2[s]; C.ch = S.ch[s]; C.orb0 = S.orb0[s];
      // Lanczos vectors per side and the resulting Dyson dimension bound
      const int ln = inst_lanc_n.empty() || inst_lanc_n[i] <= 0 ? par.lanc_n : inst_lanc_n[i];
      C.lanc_n = ln;
      C.ef_set = 0;
      C.ef_force = 0.0;
      long lside = par.lanc_kind == 0 ? (long)ln * std::min(nI, kMaxPiv) : ln;
      long lfw = std::min<long>(lside, S.mbound_fw[s]), lbk = std::min<long>(lside, S.mbound_bk[s]);
      long cp = nI;
      if (dyn && !sigma_only) cp += par.lanczos ? lfw + lbk : S.mbound_fw[s] + S.mbound_bk[s];
      if (!g0.empty()) cp = std::max<long>(cp, g0[i].np[s] + g0[i].nh[s]);
      C.cap = (int)cp;
      C.blk_off = o_blk; C.blk_cnt = (int)S.blocks[s].size();
      for (BlockDev B : S.blocks[s]) { B.gc = I.gc0 + s; h_blocks.push_back(B); }
      C.trip_off = (int)h_trip_start.size();
      C.ntrip = (int)S.trip_start[s].size() - 1;
      {
        const int tb0 = (int)h_trip_blk.size();
        for (int v : S.trip_start[s]) h_trip_start.push_back(tb0 + v);
        for (int v : S.trip_blk[s]) h_trip_blk.push_back((int)(o_blk + v));
        C.merge = 1;
        for (size_t k = 0; k + 1 < S.trip_start[s].size(); ++k)
          if (S.trip_start[s][k + 1] - S.trip_start[s][k] > kMergeMaxJ) C.merge = 0;
      }
      o_blk += C.blk_cnt;
      C.u1_off = o_u1; o_u1 += pad8((long)nI * nI);
      C.h_off = o_h; o_h += pad8((long)nI * nI);
      C.w_off = o_w; o_w += pad8((long)nI * (nI + 1) / 2 * S.nq);
      C.tri_off = o_tri; o_tri += pad8(cp * (2 + nI));
      C.mat_off = o_mat; o_mat += pad8(cp * cp);
      C.piv_off = o_piv; o_piv += pad8((long)kMaxPiv * nI);
      const int width = std::max(2, std::min(nI, kMaxPiv) + 1) + nI;
      C.lfw_cap = par.lanczos ? (int)lfw : 0; C.lbk_cap = par.lanczos ? (int)lbk : 0;
      C.lfw_off = o_lr; o_lr += pad8((long)C.lfw_cap * width);
      C.lbk_off = o_lr; o_lr += pad8((long)C.lbk_cap * width);
      C.pole_off = o_pole; o_pole += pad8((long)(1 + nI) * (C.lfw_cap + C.lbk_cap));
      C.aws_off = o_aws; o_aws += pad8((long)((arrow_ws_bytes((int)cp, nI) + 7) / 8));
      C.warm_off = o_warm; o_warm += pad8((long)nI * cp);
      e_off.push_back(o_e); o_e += pad8(cp);
      x_off.push_back(o_x); o_x += pad8(cp * nI);
      cap.push_back((int)cp);
      h_chans.push_back(C);
      capmax = std::max<long>(capmax, cp);
    }
  }
  n_gc = (int)h_chans.size();
  n_blk = (long)h_blocks.size();
  mat_len = o_mat;
  arrow_classes.clear();
  hf_classes.clear();
  for (const ChanDev &C : h_chans) {
    const std::pair<int, int> cls(arrow_threads(C.cap), arrow_ni_bound(C.nI));
    if (std::find(arrow_classes.begin(), arrow_classes.end(), cls) == arrow_classes.end()) arrow_classes.push_back(cls);
    if (C.nI > kStaticNI && std::find(hf_classes.begin(), hf_classes.end(), cls.second) == hf_classes.end())
      hf_classes.push_back(cls.second);
  }
  std::sort(arrow_classes.begin(), arrow_classes.end());
  std::sort(hf_classes.begin(), hf_classes.end());

  // ---- warm-start roots of the arrowhead sweeps
  BTRY(a_warm.grow(sizeof(double) * pad8(std::max<long>(o_warm, 1)) + sizeof(int) * kArrowNI * (size_t)n_gc + 64));
  d_warm = (double *)a_warm.ptr;
  d_warm_p = (int *)(d_warm + pad8(std::max<long>(o_warm, 1)));

  // ---- one-body parts
  std::vector<double> u1(std::max<long>(o_u1, 1), 0.0);
  for (const ChanDev &C : h_chans) sys[C.inst]->u1(C.ish, par.hw, par.t_fac, par.ext_u1, u1.data() + C.u1_off);

  // ---- device buffers (fixed size)
  BTRY(a_inst.grow(pad8(sizeof(InstDev) * n_inst) + pad8(sizeof(InstState) * n_inst) + 64));
  d_insts = (InstDev *)a_inst.ptr;
  d_is = (InstState *)((char *)a_inst.ptr + pad8(sizeof(InstDev) * n_inst));
  HTRY(bcd::sync_memcpy(d_insts, h_insts.data(), sizeof(InstDev) * n_inst, hipMemcpyHostToDevice));

  size_t chan_bytes = pad8(sizeof(ChanDev) * n_gc) + pad8(sizeof(ChanState) * n_gc) + 2 * pad8(sizeof(long) * n_gc) +
                      2 * pad8(sizeof(long) * n_gc) + pad8(sizeof(int) * n_gc) + 64;
  BTRY(a_chan.grow(chan_bytes));
  char *cp8 = (char *)a_chan.ptr;
  d_chans = (ChanDev *)cp8; cp8 += pad8(sizeof(ChanDev) * n_gc);
  d_cs = (ChanState *)cp8; cp8 += pad8(sizeof(ChanState) * n_gc);
  d_row_fw_base = (long *)cp8; cp8 += pad8(sizeof(long) * n_gc);
  d_row_bk_base = (long *)cp8; cp8 += pad8(sizeof(long) * n_gc);
  d_e_off = (long *)cp8; cp8 += pad8(sizeof(long) * n_gc);
  d_x_off = (long *)cp8; cp8 += pad8(sizeof(long) * n_gc);
  d_cap = (int *)cp8;
  HTRY(bcd::sync_memcpy(d_chans, h_chans.data(), sizeof(ChanDev) * n_gc, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memset(d_cs, 0, sizeof(ChanState) * n_gc));
  HTRY(bcd::sync_memset(d_row_fw_base, 0, sizeof(long) * n_gc * 2));
  HTRY(bcd::sync_memcpy(d_e_off, e_off.data(), sizeof(long) * n_gc, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(d_x_off, x_off.data(), sizeof(long) * n_gc, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(d_cap, cap.data(), sizeof(int) * n_gc, hipMemcpyHostToDevice));

  const long nts = std::max<long>((long)h_trip_start.size(), 1), ntb = std::max<long>((long)h_trip_blk.size(), 1);
  BTRY(a_blocks.grow(pad8(sizeof(BlockDev) * std::max<long>(n_blk, 1)) + 3 * pad8(sizeof(int) * std::max<long>(n_blk, 1)) +
                     pad8(sizeof(Sig2Work) * std::max<long>(n_blk, 1)) + pad8(sizeof(int) * nts) +
                     pad8(sizeof(int) * ntb) + 64));
  cp8 = (char *)a_blocks.ptr;
  d_blocks = (BlockDev *)cp8; cp8 += pad8(sizeof(BlockDev) * std::max<long>(n_blk, 1));
  d_bpack = (int *)cp8; cp8 += pad8(sizeof(int) * std::max<long>(n_blk, 1));
  d_blk_fw_off = (int *)cp8; cp8 += pad8(sizeof(int) * std::max<long>(n_blk, 1));
  d_blk_bk_off = (int *)cp8; cp8 += pad8(sizeof(int) * std::max<long>(n_blk, 1));
  d_trip_start = (int *)cp8; cp8 += pad8(sizeof(int) * nts);
  d_trip_blk = (int *)cp8; cp8 += pad8(sizeof(int) * ntb);
  d_work_list = (Sig2Work *)cp8;
  if (n_blk) HTRY(bcd::sync_memcpy(d_blocks, h_blocks.data(), sizeof(BlockDev) * n_blk, hipMemcpyHostToDevice));
  if (n_blk) {
    // the fields of each block the Sigma(2) plan reads, one int per block
    std::vector<int> bp((size_t)n_blk);
    for (long k = 0; k < n_blk; ++k) {
      const BlockDev &B = h_blocks[k];
      if (B.a > 255 || B.b > 255 || B.g > 255) {
        note_error("Sigma(2) plan: a model space with more than 256 subshells");
        return BCD_ERR_CAPACITY;
      }
      bp[k] = block_pack(B);
    }
    HTRY(bcd::sync_memcpy(d_bpack, bp.data(), sizeof(int) * n_blk, hipMemcpyHostToDevice));
  }
  if (!h_trip_start.empty())
    HTRY(bcd::sync_memcpy(d_trip_start, h_trip_start.data(), sizeof(int) * h_trip_start.size(), hipMemcpyHostToDevice));
  if (!h_trip_blk.empty())
    HTRY(bcd::sync_memcpy(d_trip_blk, h_trip_blk.data(), sizeof(int) * h_trip_blk.size(), hipMemcpyHostToDevice));

  // work doubles: u1, h, W, rho, tri, piv, lrows, kolt ; ints: status, work_count
  l_u1 = 0; l_h = l_u1 + pad8(o_u1); l_w = l_h + pad8(o_h); l_rho = l_w + pad8(o_w);
  l_tri = l_rho + pad8(rho_tot); l_piv = l_tri + pad8(o_tri); l_lr = l_piv + pad8(o_piv);
  l_kolt = l_lr + pad8(o_lr); l_pole = l_kolt + pad8(3L * n_gc); l_end = l_pole + pad8(o_pole) + 32;
  BTRY(a_work.grow(sizeof(double) * l_end));
  double *W0 = (double *)a_work.ptr;
  d_u1 = W0 + l_u1; d_h = W0 + l_h; d_w = W0 + l_w; d_rho = W0 + l_rho; d_tri = W0 + l_tri;
  d_piv = W0 + l_piv; d_lrows = W0 + l_lr; d_kolt = W0 + l_kolt; d_poles = W0 + l_pole;
  d_status = (int *)(W0 + l_end - 16);
  d_work_count = d_status + 2;
  d_work = W0 + l_end - 8;
  rho_len = rho_tot;
  if (par.accel > 0) {
    BTRY(a_mix.grow(sizeof(double) * 2 * (size_t)(par.accel + 1) * std::max<long>(rho_tot, 1) + 64));
    d_mix = (double *)a_mix.ptr;
  }
  BTRY(a_prof.grow(sizeof(unsigned long long) * 8 * (size_t)n_gc));
  d_prof = (unsigned long long *)a_prof.ptr;
  HTRY(bcd::sync_memset(a_work.ptr, 0, sizeof(double) * l_end));
  w_built = false;
  HTRY(bcd::sync_memcpy(d_u1, u1.data(), sizeof(double) * o_u1, hipMemcpyHostToDevice));

  // propagators: 3 buffers (initial, ping, pong), counts
  const long pe = pad8(o_e), px = pad8(o_x), pn = pad8(n_gc);
  BTRY(a_prop.grow(sizeof(double) * 3 * (pe + px) + sizeof(int) * 6 * pn + 64));
  double *P0 = (double *)a_prop.ptr;
  for (int b = 0; b < 3; ++b) { pe_[b] = P0 + b * (pe + px); px_[b] = pe_[b] + pe; }
  int *C0 = (int *)(P0 + 3 * (pe + px));
  for (int b = 0; b < 3; ++b) { pnqp_[b] = C0 + 2 * b * pn; pnqh_[b] = pnqp_[b] + pn; }
  prop_e_len = pe; prop_x_len = px; prop_n_len = pn;
  HTRY(bcd::sync_memset(a_prop.ptr, 0, a_prop.cap));

  // initial propagator
  std::vector<double> he(pe, 0.0), hx(px, 0.0);
  std::vector<int> hqp(pn, 0), hqh(pn, 0);
  for (int i = 0; i < n_inst; ++i) {
    const HostSystem &S = *sys[i];
    const HostProp &G = g0[i];
    long kp = 0, kh = 0;
    for (int s = 0; s < S.nsubsh; ++s) {
      const int gc = h_insts[i].gc0 + s, nI = S.no[s];
      hqp[gc] = G.np[s]; hqh[gc] = G.nh[s];
      for (int k = 0; k < G.np[s]; ++k, ++kp) {
        he[e_off[gc] + k] = G.ep[kp];
        for (int r = 0; r < nI; ++r) hx[x_off[gc] + (long)k * nI + r] = G.fqp[kp * G.nspb + r];
      }
      for (int k = 0; k < G.nh[s]; ++k, ++kh) {
        he[e_off[gc] + G.np[s] + k] = G.eh[kh];
        for (int r = 0; r < nI; ++r) hx[x_off[gc] + (long)(G.np[s] + k) * nI + r] = G.fqh[kh * G.nspb + r];
      }
    }
  }
  HTRY(bcd::sync_memcpy(pe_[2], he.data(), sizeof(double) * pe, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(px_[2], hx.data(), sizeof(double) * px, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(pnqp_[2], hqp.data(), sizeof(int) * pn, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(pnqh_[2], hqh.data(), sizeof(int) * pn, hipMemcpyHostToDevice));

  // jobs
  BTRY(a_jobs.grow(pad8(sizeof(DysonJob) * 3 * n_gc) + pad8(sizeof(LancJob) * 2 * n_gc) +
                   pad8(sizeof(int) * 2 * n_gc) + pad8(sizeof(int) * kCsPack * n_gc) + pad8(sizeof(int) * n_gc) + 64));
  d_djobs = (DysonJob *)a_jobs.ptr;
  d_pjobs = d_djobs + n_gc;
  d_ljobs = (LancJob *)((char *)a_jobs.ptr + pad8(sizeof(DysonJob) * 3 * n_gc));
  d_lorder = (int *)((char *)d_ljobs + pad8(sizeof(LancJob) * 2 * n_gc));
  d_cspack = (int *)((char *)d_lorder + pad8(sizeof(int) * 2 * n_gc));
  d_porder = (int *)((char *)d_cspack + pad8(sizeof(int) * kCsPack * n_gc));
  h_cspack.resize((size_t)kCsPack * n_gc);
  HTRY(bcd::sync_memset(a_jobs.ptr, 0, a_jobs.cap));
  {
    // the Sigma(2) plan takes the channels with the most blocks first, so
    // that the longest plan blocks do not start last
    std::vector<int> po(n_gc);
    for (int gc = 0; gc < n_gc; ++gc) po[gc] = gc;
    std::stable_sort(po.begin(), po.end(), [&](int x, int y) { return h_chans[x].blk_cnt > h_chans[y].blk_cnt; });
    HTRY(bcd::sync_memcpy(d_porder, po.data(), sizeof(int) * n_gc, hipMemcpyHostToDevice));
  }

  // dense matrices in HBM when the largest one does not fit in LDS: the
  // whole Dyson matrix (dense solver) or one reduced side (arrowhead solver)
  int lside_max = 1;
  for (const ChanDev &C : h_chans) lside_max = std::max(lside_max, std::max(C.lfw_cap, C.lbk_cap));
  if ((par.dyson_solver == 1 && tridiag_lds((int)capmax, true) > (size_t)kLdsBytes) ||
      (par.dyson_solver == 0 && par.lanczos && tridiag_lds(lside_max, true) > (size_t)kLdsBytes))
    BTRY(a_mat.grow(sizeof(double) * (mat_len + 16)));
  // arrowhead workspaces in HBM when the largest one does not fit in LDS
  aws_len = 0;
  if (par.dyson_solver == 0 && arrow_ws_bytes((int)capmax, nImax) > (size_t)kLdsBytes) {
    aws_len = o_aws;
    BTRY(a_aws.grow(sizeof(double) * (aws_len + 16)));
  }

  // kernel attributes for large dynamic LDS
  HTRY(hipFuncSetAttribute((const void *)k_tridiag<true>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  HTRY(hipFuncSetAttribute((const void *)k_tridiag<false>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  HTRY(hipFuncSetAttribute((const void *)k_ql, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  if (arrow_set_lds_limit(kLdsBytes)) return BCD_ERR_DEVICE;
  h_is.assign(n_inst, InstState());
  return BCD_OK;
}

// New iteration settings for later solves.  The fields that size the device
// state (mode, Lanczos settings, Dyson solver, pivot rounding, kinetic term)
// must stay; precision, density mixing, ItrMax and the stopping rule may
// change.
int Batch::set_params(const bcd_solve_params &p) {
  if (p.mode != par.mode || p.lanczos != par.lanczos || p.lanc_kind != par.lanc_kind || p.lanc_n != par.lanc_n ||
      p.dyson_solver != par.dyson_solver || p.piv_round != par.piv_round || p.t_fac != par.t_fac ||
      p.hw != par.hw || p.ext_u1 != par.ext_u1)
    return BCD_ERR_ARG;
  if (p.precision < 0 || p.precision > 3 || p.accel < 0 || p.accel > kMixMax) return BCD_ERR_ARG;
  HTRY(hipSetDevice(device));
  if (p.accel > 0) {
    BTRY(a_mix.grow(sizeof(double) * 2 * (size_t)(p.accel + 1) * std::max<long>(rho_len, 1) + 64));
    d_mix = (double *)a_mix.ptr;
  }
  par = p;
  coupling_flops = -1.0;
  return BCD_OK;
}

// Set_external_EFermi (DysonMatrix.cpp:601-617): the first value given for
// a subshell is the one Seek_EFermi returns for it from then on.
int Batch::set_efermi(int inst, int ish, double ef) {
  if (inst < 0 || inst >= n_inst || ish < 0 || ish >= h_insts[inst].nsubsh) return BCD_ERR_ARG;
  ChanDev &C = h_chans[h_insts[inst].gc0 + ish];
  if (C.ef_set) return BCD_OK;
  C.ef_set = 1;
  C.ef_force = ef;
  HTRY(hipSetDevice(device));
  HTRY(bcd::sync_memcpy(d_chans + h_insts[inst].gc0 + ish, &C, sizeof(ChanDev), hipMemcpyHostToDevice));
  return BCD_OK;
}

PropView Batch::view(bool with_state) const {
  PropView P;
  for (int b = 0; b < 2; ++b) { P.e[b] = pe_[b]; P.x[b] = px_[b]; P.nqp[b] = pnqp_[b]; P.nqh[b] = pnqh_[b]; }
  P.e_off = d_e_off; P.x_off = d_x_off; P.cap = d_cap;
  P.is = with_state ? d_is : nullptr;
  return P;
}

size_t Batch::tridiag_lds(int n, bool in_lds) const {
  size_t d = (size_t)(2 + nImax) * n + nImax + 16;
  if (in_lds) d += (size_t)n * n;
  return d * sizeof(double);
}

void Batch::kt_begin(int k) {
  if (!profile) return;
  auto &p = kev[k];
  // profiling only: a failed timer call leaves that kernel group's time at 0
  if (!p.first) { (void)hipEventCreate(&p.first); (void)hipEventCreate(&p.second); }
  (void)hipEventRecord(p.first, stream);
}

void Batch::kt_end(int k) {
  if (!profile) return;
  (void)hipEventRecord(kev[k].second, stream);
  (void)hipEventSynchronize(kev[k].second);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, kev[k].first, kev[k].second);
  ktime[k] += ms;
  kcount[k] += 1;
}

// Queue one sweep (Sigma^inf from the current propagator, Dyson, convergence
// bookkeeping) for every instance that still iterates; instances that have
// stopped are skipped inside every kernel, so queued sweeps past an
// instance's end change nothing.
int Batch::sweep(int src) {
  PropView P = view(true);
  const int gb = (n_gc + 127) / 128;
  kt_begin(KT_DENSITY);
  hipLaunchKernelGGL(k_density, dim3(n_gc), dim3(64), 0, stream, d_insts, d_chans, 0, n_gc, P, d_rho);
  if (seed_pending) {
    // second sc0 sweep of a learned start: the predicted densities replace
    // those of the first sweep's output
    hipLaunchKernelGGL(k_learn_seed, dim3((n_gc + 63) / 64), dim3(64), 0, stream, d_insts, d_chans, n_gc, d_seed_on,
                       d_hf_e, d_hf_x, d_hf_n, d_feat, d_pred, d_rho);
    seed_pending = false;
  }
  // density mixing in the sc0 stage only: the HF propagator, and with it
  // Sigma(2) and its Lanczos reduction, stay those of the plain iteration
  if (par.accel > 0 && src != 0)
    hipLaunchKernelGGL(k_mix, dim3(n_inst), dim3(256), 0, stream, d_insts, d_is, d_rho, d_mix, rho_len, par.accel,
                       par.mix_start > 0.0 ? par.mix_start : kMixStart);
  kt_end(KT_DENSITY);
  kt_begin(KT_SIGINF);
  hipLaunchKernelGGL(k_sigma_inf, dim3(n_gc), dim3(256), 0, stream, d_insts, d_chans, 0, n_gc, d_is, d_w, d_rho, d_h);
  kt_end(KT_SIGINF);
  if (par.dyson_solver == 0) {
    ArrowSweep a;
    a.chans = d_chans; a.n_gc = n_gc; a.src = src; a.cs = d_cs; a.P = P; a.h = d_h; a.u1 = d_u1;
    a.row_fw_base = d_row_fw_base; a.row_bk_base = d_row_bk_base; a.rows_fw = d_rows_fw; a.rows_bk = d_rows_bk;
    a.poles = d_poles; a.aws = aws_len ? (double *)a_aws.ptr : nullptr; a.kolt = d_kolt; a.status = d_status;
    a.work = d_work;
    a.prof = profile ? d_prof : nullptr;
    a.warm = d_warm;
    a.warm_p = d_warm_p;
    a.skip_ni = 0;
    a.nt_class = 0;
    a.nb_class = 0;
    kt_begin(KT_ARROW);
    if (src == 0) {
      // HF sweeps: channels up to kStaticNI orbits by one thread each, the
      // larger ones (if any) by the arrowhead kernel, one launch per
      // accumulator bound
      hipLaunchKernelGGL(k_dyson_static, dim3((n_gc + 63) / 64), dim3(64), 0, stream, a);
      a.skip_ni = kStaticNI;
      for (int nb : hf_classes) {
        a.nb_class = nb;
        launch_arrow_sweep(a, nb, !aws_len, 64, aws_len ? 0 : arrow_ws_bytes(nImax, std::min(nb, nImax)), stream);
      }
    } else {
      // one launch per (block size, accumulator bound) class: each channel
      // runs in the launch of its own class, so its reductions and its
      // unrolled code do not depend on the other channels of the batch.  The
      // classes touch disjoint channels and run concurrently, spread over
      // the batch's stream and the auxiliary ones.
      const int ncls = (int)arrow_classes.size();
      const int naux = std::min(kLancAux, ncls - 1);
      if (naux > 0) {
        HTRY(hipEventRecord(lfork, stream));
        for (int k = 0; k < naux; ++k) HTRY(hipStreamWaitEvent(lstream[k], lfork, 0));
      }
      for (int c = 0; c < ncls; ++c) {
        const auto &cls = arrow_classes[c];
        a.nt_class = cls.first;
        a.nb_class = cls.second;
        const int cmax = cls.first == kArrowMaxThreads ? (int)capmax : std::min<int>(cls.first, (int)capmax);
        const hipStream_t sc = (naux > 0 && c % (naux + 1)) ? lstream[c % (naux + 1) - 1] : stream;
        launch_arrow_sweep(a, cls.second, !aws_len, cls.first,
                           aws_len ? 0 : arrow_ws_bytes(cmax, std::min(cls.second, nImax)), sc);
      }
      for (int k = 0; k < naux; ++k) {
        HTRY(hipEventRecord(ljoin[k], lstream[k]));
        HTRY(hipStreamWaitEvent(stream, ljoin[k], 0));
      }
    }
    kt_end(KT_ARROW);
  } else {
    hipLaunchKernelGGL(k_dyson_prep, dim3(gb), dim3(128), 0, stream, d_chans, 0, n_gc, d_cs, P, src, d_h, d_u1,
                       d_row_fw_base, d_row_bk_base, d_rows_fw, d_rows_bk, d_lrows, d_tri,
                       (double *)a_mat.ptr, d_kolt, d_djobs);
    const int nmax = src == 0 ? nImax : (int)capmax;
    const bool in_lds = tridiag_lds(nmax, true) <= (size_t)kLdsBytes;
    kt_begin(KT_TRIDIAG);
    if (in_lds)
      hipLaunchKernelGGL(k_tridiag<true>, dim3(n_gc), dim3(256), tridiag_lds(nmax, true), stream, d_djobs, n_gc);
    else
      hipLaunchKernelGGL(k_tridiag<false>, dim3(n_gc), dim3(256), tridiag_lds(nmax, false), stream, d_djobs, n_gc);
    kt_end(KT_TRIDIAG);
    kt_begin(KT_QL);
    hipLaunchKernelGGL(k_ql, dim3(n_gc), dim3(64), (size_t)(2 + nImax) * nmax * sizeof(double), stream, d_djobs,
                       n_gc, d_status);
    kt_end(KT_QL);
  }
  kt_begin(KT_FINAL);
  hipLaunchKernelGGL(k_finalize, dim3(n_inst), dim3(256), 0, stream, d_insts, d_chans, P, d_cs, d_kolt,
                     par.conv_kind, 0);
  kt_end(KT_FINAL);
  HTRY(hipGetLastError());
  return BCD_OK;
}

// Per-instance flags after the queued sweeps: how many instances still
// iterate, capacity and eigensolver failures.
int Batch::check(int &n_active) {
  HTRY(hipMemcpyAsync(h_is.data(), d_is, sizeof(InstState) * n_inst, hipMemcpyDeviceToHost, stream));
  HTRY(hipStreamSynchronize(stream));
  n_active = 0;
  for (int i = 0; i < n_inst; ++i) {
    const InstState &s = h_is[i];
    if (s.err) {
      note_error("Dyson sweep: a channel of instance " + std::to_string(i) + " exceeds its propagator slots");
      return BCD_ERR_CAPACITY;
    }
    n_active += s.active;
  }
  int st = 0;
  HTRY(bcd::sync_memcpy(&st, d_status, sizeof(int), hipMemcpyDeviceToHost));
  if (st) return BCD_ERR_EIGEN;
  return BCD_OK;
}

int Batch::stage(int src, int itr_max, double conv, bool hf) {
  // stage start: keep the current buffer, reset the sweep controls
  HTRY(hipMemcpyAsync(h_is.data(), d_is, sizeof(InstState) * n_inst, hipMemcpyDeviceToHost, stream));
  HTRY(hipStreamSynchronize(stream));
  for (int i = 0; i < n_inst; ++i) {
    InstState &s = h_is[i];
    s.active = s.failed ? 0 : (stage_mask.empty() || stage_mask[i] ? 1 : 0);
    s.its = 0; s.itr_max = itr_max; s.conv = conv; s.err = 0;
    s.mix_k = 0; s.mix_s = 0; s.mix_fn = 0.0;
  }
  HTRY(hipMemcpyAsync(d_is, h_is.data(), sizeof(InstState) * n_inst, hipMemcpyHostToDevice, stream));
  // no arrowhead roots to start from in a stage's first sweep
  HTRY(hipMemsetAsync(d_warm_p, 0xff, sizeof(int) * kArrowNI * (size_t)n_gc, stream));
  hipLaunchKernelGGL(k_finalize, dim3(n_inst), dim3(256), 0, stream, d_insts, d_chans, view(true), d_cs, d_kolt,
                     par.conv_kind, 1);
  HTRY(hipGetLastError());
  const int hard_cap = itr_max >= 0 ? itr_max + 1 : kMaxSweeps;
  int n_active = n_inst, done = 0;
  hipGraphExec_t gexec = nullptr;
  while (done < hard_cap && n_active > 0) {
    // sweeps are queued kSweepChunk at a time between reads of the flags,
    // twice as many in the HF stage of the arrowhead path, whose sweeps cost
    // about as much as one read
    const int chunk = (hf && par.dyson_solver == 0) ? 2 * kSweepChunk : kSweepChunk;
    const int k = std::min(chunk, hard_cap - done);
    const bool hook = learn_hook && done == 0;
    if (use_graphs && !profile && !hook && !seed_pending && k == chunk && par.dyson_solver == 0) {
      // a full chunk: captured once per stage into a graph, replayed after
      if (!gexec) {
        hipGraph_t g = nullptr;
        // relaxed: the region holds only this batch's launches and its own
        // event records and waits; other host threads may keep issuing
        // synchronous calls meanwhile
        std::unique_lock<std::mutex> cap_lock(hip_sync_lock());
        HTRY(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        int rc = BCD_OK;
        for (int q = 0; q < k && rc == BCD_OK; ++q) rc = sweep(src);
        const hipError_t e = hipStreamEndCapture(stream, &g);
        cap_lock.unlock();
        if (rc == BCD_OK && e == hipSuccess) {
          const hipError_t ei = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
          if (ei != hipSuccess) { note_hip(ei, __LINE__); rc = BCD_ERR_DEVICE; }
        } else if (rc == BCD_OK) {
          note_hip(e, __LINE__);
          rc = BCD_ERR_DEVICE;
        }
        if (g) (void)hipGraphDestroy(g);
        if (rc != BCD_OK) {
          if (gexec) (void)hipGraphExecDestroy(gexec);
          return rc;
        }
      }
      const hipError_t el = hipGraphLaunch(gexec, stream);
      if (el != hipSuccess) {
        note_hip(el, __LINE__);
        (void)hipGraphExecDestroy(gexec);
        return BCD_ERR_DEVICE;
      }
    } else {
      for (int q = 0; q < k; ++q) {
        BTRY(sweep(src));
        if (hook && q == 0) BTRY(after_first_sweep());
      }
    }
    done += k;
    const int rc = check(n_active);
    if (rc != BCD_OK) {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      return rc;
    }
  }
  if (gexec) HTRY(hipGraphExecDestroy(gexec));
  for (int i = 0; i < n_inst; ++i) {
    (hf ? hf_its : sc_its)[i] = h_is[i].its;
    (hf ? hf_ok : sc_ok)[i] = h_is[i].diff < conv && !h_is[i].failed;
  }
  return BCD_OK;
}

// The per-channel counts the host needs between stages (pole counts, error
// flag, reduced sizes), gathered on the device and copied back compactly;
// the rest of h_cs is refreshed only by pull_cs().
int Batch::pull_counts() {
  hipLaunchKernelGGL(k_cs_pack, dim3((n_gc + 255) / 256), dim3(256), 0, stream, d_cs, n_gc, d_cspack);
  HTRY(hipGetLastError());
  HTRY(hipMemcpyAsync(h_cspack.data(), d_cspack, sizeof(int) * kCsPack * n_gc, hipMemcpyDeviceToHost, stream));
  HTRY(hipStreamSynchronize(stream));
  h_cs.resize(n_gc);
  for (int gc = 0; gc < n_gc; ++gc) {
    const int *q = &h_cspack[(size_t)kCsPack * gc];
    ChanState &S = h_cs[gc];
    S.mfw = q[0]; S.mbk = q[1]; S.lfw = q[2]; S.lbk = q[3]; S.err = q[4]; S.lmfw = q[5]; S.lmbk = q[6];
  }
  return BCD_OK;
}

int Batch::pull_cs() {
  h_cs.resize(n_gc);
  HTRY(hipMemcpyAsync(h_cs.data(), d_cs, sizeof(ChanState) * n_gc, hipMemcpyDeviceToHost, stream));
  HTRY(hipStreamSynchronize(stream));
  return BCD_OK;
}

int Batch::build_sigma2(int only_ish) {
  PropView P = view(true);
  HTRY(hipMemsetAsync(d_work_count, 0, sizeof(int), stream));
  kt_begin(KT_PLAN);
  hipLaunchKernelGGL(k_sigma2_plan, dim3(n_gc), dim3(256), 0, stream, d_chans, d_bpack, 0, n_gc, P, d_insts, d_cs,
                     d_blk_fw_off, d_blk_bk_off, d_work_list, d_work_count, only_ish, d_porder);
  kt_end(KT_PLAN);
  HTRY(hipGetLastError());
  BTRY(pull_counts());
  // row buffers and Lanczos workspace sized by the actual pole counts
  for (int gc = 0; gc < n_gc; ++gc)
    if (h_cs[gc].err == 3) {   // a fragment count above 65535
      note_error("Sigma(2): a fragment count of channel " + std::to_string(gc) + " exceeds 65535");
      return BCD_ERR_CAPACITY;
    }
  std::vector<long> bfw(n_gc), bbk(n_gc);
  long tfw = 0, tbk = 0, tws = 0;
  bool relayout = false;
  for (int gc = 0; gc < n_gc; ++gc) {
    ChanDev &C = h_chans[gc];
    const long w = 2 + C.nI;
    bfw[gc] = tfw; tfw += pad8((long)h_cs[gc].mfw * w);
    bbk[gc] = tbk; tbk += pad8((long)h_cs[gc].mbk * w);
    if (par.lanczos) {
      const long lws_fw = tws, lws_bk = tws + pad8(lanczos_ws_vectors(C.lfw_cap, C.nI) * lanczos_stride(h_cs[gc].mfw));
      tws = lws_bk + pad8(lanczos_ws_vectors(C.lbk_cap, C.nI) * lanczos_stride(h_cs[gc].mbk));
      if (C.lws_fw_off != lws_fw || C.lws_bk_off != lws_bk) relayout = true;
      C.lws_fw_off = lws_fw; C.lws_bk_off = lws_bk;
    }
    if (!sigma_only && !par.lanczos && h_cs[gc].mfw + h_cs[gc].mbk + C.nI > C.cap) {
      note_error("Dyson: channel " + std::to_string(gc) + " has more poles than its propagator slots");
      return BCD_ERR_CAPACITY;
    }
  }
  n_fw_tot.assign(n_inst, 0); n_bk_tot.assign(n_inst, 0);
  for (int gc = 0; gc < n_gc; ++gc) { n_fw_tot[h_chans[gc].inst] += h_cs[gc].mfw; n_bk_tot[h_chans[gc].inst] += h_cs[gc].mbk; }
  // the merged rows of the Lanczos path sit behind the rows, same layout
  const long crow = par.lanczos ? tfw + tbk : 0;
  BTRY(a_rows.grow(sizeof(double) * (tfw + tbk + crow + 16)));
  d_rows_fw = (double *)a_rows.ptr;
  d_rows_bk = d_rows_fw + tfw;
  // (merging needs every (a, b, g) triple's J blocks in LDS; channels with
  // longer triples keep their rows, ChanDev::merge)
  d_crows_fw = par.lanczos ? d_rows_bk + tbk : nullptr;
  d_crows_bk = par.lanczos ? d_crows_fw + tfw : nullptr;
  if (par.lanczos) BTRY(a_lws.grow(sizeof(double) * (tws + 16)));
  HTRY(hipMemcpyAsync(d_row_fw_base, bfw.data(), sizeof(long) * n_gc, hipMemcpyHostToDevice, stream));
  HTRY(hipMemcpyAsync(d_row_bk_base, bbk.data(), sizeof(long) * n_gc, hipMemcpyHostToDevice, stream));
  if (relayout) HTRY(hipMemcpyAsync(d_chans, h_chans.data(), sizeof(ChanDev) * n_gc, hipMemcpyHostToDevice, stream));
  if (coupling_flops < 0.0) {
    // 2 * poles * K * n per block, K = (alpha,beta) pairs * gamma, from the
    // fragment counts the build used (reference formulation, Op_c.cpp:454-468)
    std::vector<int> cq(pnqp_[0] ? n_gc : 0), chh(n_gc), cq1(n_gc), ch1(n_gc);
    HTRY(bcd::sync_memcpy(cq.data(), pnqp_[0], sizeof(int) * n_gc, hipMemcpyDeviceToHost));
    HTRY(bcd::sync_memcpy(chh.data(), pnqh_[0], sizeof(int) * n_gc, hipMemcpyDeviceToHost));
    HTRY(bcd::sync_memcpy(cq1.data(), pnqp_[1], sizeof(int) * n_gc, hipMemcpyDeviceToHost));
    HTRY(bcd::sync_memcpy(ch1.data(), pnqh_[1], sizeof(int) * n_gc, hipMemcpyDeviceToHost));
    double f = 0.0;
    for (const BlockDev &B : h_blocks) {
      const ChanDev &C = h_chans[B.gc];
      const HostSystem &S = *sys[C.inst];
      const int g0 = h_insts[C.inst].gc0, cur = h_is[C.inst].cur;
      const int *NQ = cur ? cq1.data() : cq.data(), *NH = cur ? ch1.data() : chh.data();
      if (only_ish >= 0 && C.ish != only_ish) continue;
      const int na = S.no[B.a], nb = S.no[B.b], ng = S.no[B.g];
      long npair = 0;
      for (int al = 0; al < na; ++al) npair += nb - ((B.a == B.b) ? al + B.nJab : 0);
      const double K = (double)npair * ng;
      for (int side = 0; side < 2; ++side) {
        const int cA = side ? NH[g0 + B.a] : NQ[g0 + B.a], cB = side ? NH[g0 + B.b] : NQ[g0 + B.b];
        const int cG = side ? NQ[g0 + B.g] : NH[g0 + B.g];
        long poles;
        if (B.a == B.b) { int nb2 = cA + 1; if (B.nJab == 1) nb2 -= 2; poles = (long)(cA * nb2 / 2) * cG; }
        else poles = (long)cA * cB * cG;
        f += 2.0 * poles * K * C.nI;
      }
    }
    coupling_flops = f;
  }
  // the kernel reads the work-list length from the device and strides over
  // it, so a fixed grid needs no copy of the count back to the host
  kt_begin(KT_SIGMA2);
  int vt_cap = 2;
  int c_cap = 2;
  for (const HostSystem *S : sys) {
    vt_cap = std::max(vt_cap, (S->max_vt + 1) & ~1);
    c_cap = std::max(c_cap, (S->max_dg + 1) & ~1);
  }
  // precision 1 / 2: the contraction on the matrix cores with FP32 / BF16
  // operands and FP32 accumulation; 3: FP32 operands and products on the
  // vector units, FP64 sums
  auto kern = par.precision == 1 ? k_sigma2<float, double, 1>
              : par.precision == 2 ? k_sigma2<float, double, 2>
              : par.precision == 3 ? k_sigma2<float, double, 0> : k_sigma2<double, double, 0>;
  // one launch per V-table size class, each with the LDS of its largest
  // table, the classes on concurrent streams (they write disjoint rows)
  // (blocks without (alpha, beta) pairs have nvt = 0 and still write their
  // rows: the first class starts below 0)
  std::vector<std::pair<int, int>> classes;
  int lo = -1;
  for (int cap : kSigVtClass) {
    if (lo >= vt_cap) break;
    classes.push_back({lo, std::min(cap, vt_cap)});
    lo = cap;
  }
  if (lo < vt_cap) classes.push_back({lo, vt_cap});
  const int naux = std::min(kLancAux, (int)classes.size() - 1);
  if (naux > 0) {
    HTRY(hipEventRecord(lfork, stream));
    for (int k = 0; k < naux; ++k) HTRY(hipStreamWaitEvent(lstream[k], lfork, 0));
  }
  for (int c = 0; c < (int)classes.size(); ++c) {
    const int cap = (classes[c].second + 1) & ~1;
    const size_t lds = sigma2_lds(cap, c_cap, par.precision == 1 || par.precision == 2);
    if (lds > 64 * 1024)
      (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const hipStream_t sc = c ? lstream[(c - 1) % kLancAux] : stream;
    hipLaunchKernelGGL(kern, dim3(kSigGrid), dim3(64), lds, sc, d_insts, d_chans, d_blocks, P, d_blk_fw_off,
                       d_blk_bk_off, d_work_list, d_work_count, d_row_fw_base, d_row_bk_base, d_rows_fw, d_rows_bk,
                       classes[c].first, classes[c].second, c_cap);
  }
  for (int k = 0; k < naux; ++k) {
    HTRY(hipEventRecord(ljoin[k], lstream[k]));
    HTRY(hipStreamWaitEvent(stream, ljoin[k], 0));
  }
  kt_end(KT_SIGMA2);
  HTRY(hipGetLastError());
  return BCD_OK;
}

// Time every tile shape that holds the bucket's largest job (and is at most
// four times wider than the bucket) on the bucket's own jobs, keep the
// fastest.  The jobs are idempotent: each run rebuilds the same outputs.
int Batch::tune_bucket(int bk, int mmax, int nj, const int *ord, int prev) {
  hipEvent_t e0, e1;
  HTRY(hipEventCreate(&e0));
  HTRY(hipEventCreate(&e1));
  int best = lanc_shape[bk];
  float best_ms = 0.f;
  for (int sh = 0; sh < kNLancShapes; ++sh) {
    const long cap = (long)kLancShapes[sh].nt * kLancShapes[sh].per;
    if (!lanczos_shape_fits(sh, mmax) || cap > 4L * kLancBucket[bk]) continue;
    HTRY(hipEventRecord(e0, stream));
    launch_lanczos(sh, nj, d_ljobs, ord, stream);
    HTRY(hipEventRecord(e1, stream));
    HTRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HTRY(hipEventElapsedTime(&ms, e0, e1));
    lanc_tune_ms[bk][sh] = ms;
    if (best_ms == 0.f || ms < best_ms) { best_ms = ms; best = sh; }
  }
  // the shape of the bucket of the next larger jobs, when it is nearly as
  // fast here: the two buckets then run as one launch (one queue of jobs,
  // largest first), which fills the device better than two launches
  if (prev >= 0 && lanc_tune_ms[bk][prev] > 0.f && lanc_tune_ms[bk][prev] <= (1.f + kLancMergeTol) * best_ms)
    best = prev;
  lanc_shape[bk] = best;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return BCD_OK;
}

int Batch::lanczos() {
  PropView P = view(true);
  // merge the J-degenerate poles, then the jobs pick merged or plain rows;
  // the (channel, side) blocks of the merge run largest first
  {
    std::vector<std::pair<int, int>> mo;
    for (int gc = 0; gc < n_gc; ++gc)
      for (int side = 0; side < 2; ++side)
        mo.push_back({h_chans[gc].merge ? (side ? h_cs[gc].mbk : h_cs[gc].mfw) : 0, 2 * gc + side});
    std::stable_sort(mo.begin(), mo.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first > b.first; });
    std::vector<int> ord(mo.size());
    for (size_t k = 0; k < mo.size(); ++k) ord[k] = mo[k].second;
    HTRY(hipMemcpyAsync(d_lorder, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice, stream));
  }
  kt_begin(KT_MERGE);
  hipLaunchKernelGGL(k_merge_poles, dim3(2 * n_gc), dim3(256), 0, stream, d_chans, n_gc, d_cs, d_blocks, d_trip_start,
                     d_trip_blk, P, d_insts, d_blk_fw_off, d_blk_bk_off, d_row_fw_base, d_row_bk_base, d_rows_fw,
                     d_rows_bk, d_crows_fw, d_crows_bk, d_lorder);
  kt_end(KT_MERGE);
  HTRY(hipGetLastError());
  hipLaunchKernelGGL(k_lanczos_prep, dim3((n_gc + 63) / 64), dim3(64), 0, stream, d_chans, 0, n_gc, P, d_cs, d_piv,
                     par.lanc_kind, par.piv_round, d_row_fw_base, d_row_bk_base, d_rows_fw, d_rows_bk, d_crows_fw,
                     d_crows_bk, d_lrows, (double *)a_lws.ptr, d_ljobs);
  HTRY(hipGetLastError());
  BTRY(pull_counts());
  // jobs longest first, grouped into pole-count buckets; one launch per
  // bucket with the bucket's tile shape, the largest jobs first
  // a channel with more poles than the largest Lanczos tile holds drops its
  // instance from this solve (reported as not converged); the others go on
  std::vector<char> drop(n_inst, 0);
  for (int gc = 0; gc < n_gc; ++gc)
    for (int side = 0; side < 2; ++side) {
      const int M = side ? h_cs[gc].lmbk : h_cs[gc].lmfw;
      if (M > 0 && lanczos_bucket(M) < 0) {
        drop[h_chans[gc].inst] = 1;
        note_error("Lanczos: instance " + std::to_string(h_chans[gc].inst) + " subshell " +
                   std::to_string(h_chans[gc].ish) + " has " + std::to_string(M) + " poles on one side, above " +
                   std::to_string(kLancBucket[kNLancBuckets - 1]) + "; the instance is dropped from the solve");
      }
    }
  if (std::count(drop.begin(), drop.end(), 1)) {
    HTRY(hipMemcpyAsync(h_is.data(), d_is, sizeof(InstState) * n_inst, hipMemcpyDeviceToHost, stream));
    HTRY(hipStreamSynchronize(stream));
    for (int i = 0; i < n_inst; ++i)
      if (drop[i]) { h_is[i].failed = 1; h_is[i].active = 0; }
    HTRY(hipMemcpyAsync(d_is, h_is.data(), sizeof(InstState) * n_inst, hipMemcpyHostToDevice, stream));
  }
  std::vector<std::pair<int, int>> jm;
  for (int gc = 0; gc < n_gc; ++gc)
    for (int side = 0; side < 2; ++side) {
      const int M = side ? h_cs[gc].lmbk : h_cs[gc].lmfw;
      jm.push_back({drop[h_chans[gc].inst] ? 0 : M, 2 * gc + side});
    }
  std::stable_sort(jm.begin(), jm.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
    return a.first > b.first;
  });
  std::vector<int> order(jm.size());
  for (size_t k = 0; k < jm.size(); ++k) order[k] = jm[k].second;
  HTRY(hipMemcpyAsync(d_lorder, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, stream));
  kt_begin(KT_LANCZOS);
  // the buckets: [k0, k0 + nj) of the job order, largest jobs first
  struct LBucket { int bk, k0, nj; double est; };
  std::vector<LBucket> lb;
  for (size_t k0 = 0; k0 < jm.size();) {
    const int bk = jm[k0].first > 0 ? lanczos_bucket(jm[k0].first) : 0;
    size_t k1 = k0;
    while (k1 < jm.size() && (jm[k1].first > 0 ? lanczos_bucket(jm[k1].first) : 0) == bk) ++k1;
    lb.push_back({bk, (int)k0, (int)(k1 - k0), 0.0});
    k0 = k1;
  }
  // Tuning runs the buckets one at a time, largest first, so that each
  // shape is timed alone.  Otherwise the buckets run concurrently on
  // kLancStreams streams (the batch's and the first auxiliary ones), joined
  // back before the stage ends.  The buckets are still launched largest
  // jobs first: the longest single jobs set the stage's critical path, and
  // launched after a bucket of thousands of blocks they would wait for free
  // CUs.  Each goes to the stream with the least estimated work queued (list
  // scheduling), so that no stream runs several large buckets back to back
  // after the others have drained.  A bucket's estimate is its autotuned
  // time alone on the device where there is one, else its bytes at the
  // device's streaming rate or its largest job's bytes at one CU's rate,
  // whichever is longer; plus its pole blocks.
  const bool fork = !tune_lanczos;
  if (tune_lanczos)
    for (size_t k = 0; k < lb.size(); ++k)
      BTRY(tune_bucket(lb[k].bk, std::max(jm[lb[k].k0].first, 1), lb[k].nj, d_lorder + lb[k].k0,
                       k ? lanc_shape[lb[k - 1].bk] : -1));
  if (fork) {
    // neighbouring buckets with the same tile shape run as one launch
    std::vector<LBucket> merged;
    for (const LBucket &b : lb)
      if (!merged.empty() && lanc_shape[merged.back().bk] == lanc_shape[b.bk]) merged.back().nj += b.nj;
      else merged.push_back(b);
    lb.swap(merged);
    for (LBucket &b : lb) {
      const float tuned = lanc_tune_ms[b.bk][lanc_shape[b.bk]];
      if (tuned > 0.f) {
        b.est = tuned;
      } else {
        double bytes = 0.0, big = 0.0;
        for (int k = b.k0; k < b.k0 + b.nj; ++k) {
          const int gc = jm[k].second / 2, side = jm[k].second % 2;
          const double M = jm[k].first, nI = h_chans[gc].nI;
          const double L = std::min<double>(side ? h_chans[gc].lbk_cap : h_chans[gc].lfw_cap, M);
          const double jb = 8.0 * M * (0.5 * L * L + L * (2.0 + nI));
          bytes += jb;
          big = std::max(big, jb);
        }
        b.est = std::max(bytes / kLancChipBytesPerMs, big / kLancCuBytesPerMs);
      }
      b.est += kLancPoleMs;
    }
    if (kLancOrderByEst)
      std::stable_sort(lb.begin(), lb.end(), [](const LBucket &a, const LBucket &b) { return a.est > b.est; });
  }
  hipStream_t ls[kLancStreams] = {stream};
  for (int a = 1; a < kLancStreams; ++a) ls[a] = lstream[a - 1];
  // the stream of each launch: list scheduling by estimate
  std::vector<int> place(lb.size(), 0);
  {
    double load[kLancStreams] = {};
    for (size_t i = 0; i < lb.size(); ++i) {
      int a = 0;
      if (fork)
        for (int k = 1; k < kLancStreams; ++k)
          if (load[k] < load[a]) a = k;
      load[a] += lb[i].est;
      place[i] = a;
    }
  }
  // The estimates are times alone on the device; run together, the launches
  // share it unevenly (a launch of a few long jobs holds its stream for twice
  // its time alone while taking little of the device), so the streams can
  // end far apart.  The launches' ends measured in the last forked stage
  // correct that: the last launch of the stream that ended last moves to
  // the stream that ended first when that ends the stage earlier by the
  // measured times.  The placement changes no result.
  if (fork) {
    std::vector<int> key(lb.size());
    for (size_t i = 0; i < lb.size(); ++i) key[i] = lb[i].bk;
    if (key != lp_key) {
      lp_key = key;
      lp_used.clear();
      lp_best.clear();
      lp_timed = lp_fixed = lp_frozen = false;
      lp_moves = 0;
    }
    if (lp_fixed) place = lp_used;
    if (lp_timed && !lp_frozen && lp_used.size() == lb.size() && lb.size() <= (size_t)kNLancBuckets) {
      float t[kNLancBuckets] = {};
      bool ok = hipEventQuery(lt0) == hipSuccess;
      for (size_t i = 0; ok && i < lb.size(); ++i)
        ok = hipEventQuery(lbt[i]) == hipSuccess && hipEventElapsedTime(&t[i], lt0, lbt[i]) == hipSuccess;
      bool moved = false;
      if (ok) {
        float send[kLancStreams] = {};
        int last[kLancStreams], cnt[kLancStreams] = {};
        for (int s = 0; s < kLancStreams; ++s) last[s] = -1;
        for (size_t i = 0; i < lb.size(); ++i) {
          const int s = lp_used[i];
          ++cnt[s];
          if (last[s] < 0 || t[i] >= send[s]) { send[s] = t[i]; last[s] = (int)i; }
        }
        int L = 0, F = 0;
        for (int s = 1; s < kLancStreams; ++s) {
          if (send[s] > send[L]) L = s;
          if (send[s] < send[F]) F = s;
        }
        if (lp_best.empty() || send[L] < lp_best_end) { lp_best = lp_used; lp_best_end = send[L]; }
        if (lp_moves < kLancMoves && L != F && cnt[L] >= 2) {
          float before = 0.f, other = 0.f;   // end of L's launch before its last; the third stream's end
          for (size_t i = 0; i < lb.size(); ++i)
            if (lp_used[i] == L && (int)i != last[L]) before = std::max(before, t[i]);
          for (int s = 0; s < kLancStreams; ++s)
            if (s != L && s != F) other = std::max(other, send[s]);
          const float to_f = send[F] + (send[L] - before);
          if (std::max(std::max(before, to_f), other) + kLancMoveGainMs < send[L]) {
            place = lp_used;
            place[last[L]] = F;
            ++lp_moves;
            moved = true;
          }
        }
      }
      if (!moved) {
        // keep the placement that ended earliest of those measured
        if (!lp_best.empty()) place = lp_best;
        lp_frozen = true;
      }
      lp_fixed = true;
    }
  } else {
    lp_key.clear();
  }
  bool used[kLancStreams] = {};
  if (fork) HTRY(hipEventRecord(lfork, stream));
  // outside tuning and profiling, each bucket's pole blocks are diagonalised
  // on its stream right after its Lanczos jobs (they need nothing else), so
  // that they overlap the other buckets' reductions
  const PoleCfg pc = pole_cfg();
  poles_done = fork && !profile && par.dyson_solver == 0;
  const bool timed = poles_done && lb.size() <= (size_t)kNLancBuckets;
  if (timed) HTRY(hipEventRecord(lt0, stream));
  for (size_t i = 0; i < lb.size(); ++i) {
    const LBucket &b = lb[i];
    const int a = place[i];
    if (a > 0 && !used[a]) HTRY(hipStreamWaitEvent(ls[a], lfork, 0));
    used[a] = true;
    launch_lanczos(lanc_shape[b.bk], b.nj, d_ljobs, d_lorder + b.k0, ls[a]);
    if (poles_done) launch_poles(pc, d_lorder + b.k0, b.nj, ls[a]);
    if (timed) HTRY(hipEventRecord(lbt[i], ls[a]));
  }
  if (fork) {
    lp_used = place;
    lp_timed = timed;
  }
  for (int a = 1; a < kLancStreams; ++a)
    if (used[a]) {
      HTRY(hipEventRecord(ljoin[a - 1], ls[a]));
      HTRY(hipStreamWaitEvent(stream, ljoin[a - 1], 0));
    }
  tune_lanczos = false;
  kt_end(KT_LANCZOS);
  HTRY(hipGetLastError());
  BTRY(pull_counts());
  for (int gc = 0; gc < n_gc; ++gc)
    if (!drop[h_chans[gc].inst] && (h_cs[gc].err || h_cs[gc].lfw < 0 || h_cs[gc].lbk < 0)) {
      note_error("Lanczos: channel " + std::to_string(gc) + " failed (code " +
                 std::to_string(std::min(h_cs[gc].lfw, h_cs[gc].lbk)) + ")");
      return BCD_ERR_LANCZOS;
    }
  return BCD_OK;
}

// Launch parameters of the pole-block kernels for this batch.  Each reduced
// side takes the one-wave path or the general one by its own size
// (pole_path), so both may run in one batch.
Batch::PoleCfg Batch::pole_cfg() const {
  PoleCfg c;
  for (const ChanDev &C : h_chans)
    for (int side = 0; side < 2; ++side) {
      const int cap = side ? C.lbk_cap : C.lfw_cap;
      if (cap <= 0) continue;
      c.lmax = std::max(c.lmax, cap);
      if (pole_path(cap, C.nI) == 0) { c.any_wave = true; c.lw = std::max(c.lw, cap); }
      else c.any_general = true;
    }
  c.ldw = c.lw | 1;
  c.wlds = sizeof(double) * ((size_t)c.ldw * c.ldw + 2 * kPoleWaveN);
  c.in_lds = tridiag_lds(c.lmax, true) <= (size_t)kLdsBytes;
  return c;
}

// Prepares the pole-block jobs of the njobs sides order[0..njobs) and runs
// the one-wave path on those it takes (k_tridiag_poles, k_ql_poles) on
// stream s; the general path runs in general_poles().
void Batch::launch_poles(const PoleCfg &c, const int *order, int njobs, hipStream_t s) {
  hipLaunchKernelGGL(k_poles_prep, dim3((njobs + 127) / 128), dim3(128), 0, s, d_chans, n_gc, d_cs, d_lrows, d_tri,
                     c.in_lds ? nullptr : (double *)a_mat.ptr, d_poles, d_pjobs, order, njobs);
  if (!c.any_wave) return;
  // above 64 KiB the kernels need the larger dynamic-LDS limit
  if (c.wlds > 65536) {
    (void)hipFuncSetAttribute((const void *)k_tridiag_poles, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.wlds);
  }
  hipLaunchKernelGGL(k_tridiag_poles, dim3(njobs), dim3(64), c.wlds, s, d_pjobs, order, njobs, c.ldw);
  hipLaunchKernelGGL(k_ql_poles, dim3((njobs + 7) / 8), dim3(64), (size_t)8 * 2 * c.lw * sizeof(double), s,
                     d_pjobs, order, njobs,
                     d_status, c.lw, std::min(nImax, kQLGroupNI));
}

// The general path (k_tridiag + k_ql) over every prepared side it takes.
void Batch::general_poles(const PoleCfg &c, hipStream_t s) {
  if (!c.any_general) return;
  if (c.in_lds)
    hipLaunchKernelGGL(k_tridiag<true>, dim3(2 * n_gc), dim3(256), tridiag_lds(c.lmax, true), s, d_pjobs, 2 * n_gc);
  else
    hipLaunchKernelGGL(k_tridiag<false>, dim3(2 * n_gc), dim3(256), tridiag_lds(c.lmax, false), s, d_pjobs,
                       2 * n_gc);
  hipLaunchKernelGGL(k_ql, dim3(2 * n_gc), dim3(64), (size_t)(2 + nImax) * c.lmax * sizeof(double), s, d_pjobs,
                     2 * n_gc, d_status);
}

// Diagonalise every Lanczos-reduced side once (its fishbone block and
// couplings are fixed for the whole sc0 stage): pole energies and rotated
// couplings for the arrowhead sweeps.  The one-wave sides are done already
// when lanczos() ran them bucket by bucket.
int Batch::decompose_poles() {
  const PoleCfg c = pole_cfg();
  if (tridiag_lds(c.lmax, false) > (size_t)kLdsBytes || (!c.in_lds && c.any_general && !a_mat.ptr)) {
    note_error("pole blocks: a reduced side of " + std::to_string(c.lmax) + " rows does not fit the kernels");
    return BCD_ERR_CAPACITY;
  }
  kt_begin(KT_POLES);
  if (!poles_done) launch_poles(c, nullptr, 2 * n_gc, stream);
  general_poles(c, stream);
  kt_end(KT_POLES);
  HTRY(hipGetLastError());
  return BCD_OK;
}

// Algorithmic work of the last solve (see bcd_batch_work in bcdor_amd.h).
int Batch::work(double *w, int n) {
  for (int k = 0; k < n; ++k) w[k] = 0.0;
  if (!h_cs.empty()) BTRY(pull_cs());
  if (n >= 15) {
    std::vector<unsigned long long> t(8 * (size_t)n_gc);
    int khz = 0;
    HTRY(bcd::sync_memcpy(t.data(), d_prof, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost));
    HTRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    for (int k = 0; k < 8; ++k) {
      double tk = 0.0;
      for (int gc = 0; gc < n_gc; ++gc) tk += (double)t[8L * gc + k];
      w[7 + k] = khz > 0 ? tk / (double)khz : 0.0;
    }
  }
  w[0] = coupling_flops;
  if (d_work) HTRY(bcd::sync_memcpy(&w[5], d_work, sizeof(double), hipMemcpyDeviceToHost));
  for (int gc = 0; gc < (int)h_cs.size() && par.lanczos && par.dyson_solver == 0; ++gc) {
    const double lf = std::max(h_cs[gc].lfw, 0), lb = std::max(h_cs[gc].lbk, 0);
    w[6] += 9.0 * (lf * lf * lf + lb * lb * lb);
  }
  // Lanczos bytes as k_lanczos moves them (lanczos_job): the rows the job
  // runs on (merged or not, M of them) once -- E and the nI couplings --, the
  // coupling columns written once; per step j the nI columns again when a
  // pivot chain starts, each earlier Krylov vector but the latest once (the
  // latest is on chip), the nI columns and the E*v_last vector of every
  // finished chain for the sums, the new vector written, and E*v_last
  // written at the end of a chain.
  for (int gc = 0; gc < (int)h_cs.size(); ++gc) {
    const ChanDev &C = h_chans[gc];
    const ChanState &Q = h_cs[gc];
    for (int side = 0; side < 2 && par.lanczos; ++side) {
      const double M = side ? Q.lmbk : Q.lmfw;
      const int L = side ? Q.lbk : Q.lfw;
      if (L <= 0 || M <= 0) continue;
      const int *itr = side ? Q.itr_bk : Q.itr_fw;
      const int nb = C.nI;
      double vec = 1.0 + 2.0 * nb;          // E and couplings read, couplings written
      int j = 0, done_chains = 0;
      for (int k = 0; k < Q.npiv_tot && k < kMaxPiv && j < L; ++k) {
        if (itr[k] < 1) continue;
        for (int t = 0; t < itr[k] && j < L; ++t, ++j) {
          if (t == 0) vec += nb;                       // start vector M u
          vec += j > 1 ? j - 1 : 0;                   // Gram-Schmidt reads
          vec += nb + done_chains;                    // couplings and g_i sums
          vec += 1.0;                                 // the new vector
        }
        if (j < L) vec += 1.0;                        // E * v_last of the finished chain
        ++done_chains;
      }
      w[1] += 8.0 * M * vec;
    }
  }
  for (int gc = 0; gc < n_gc; ++gc) {
    const ChanDev &C = h_chans[gc];
    const int i = C.inst;
    const double n0 = C.nI;
    double n1 = C.nI;
    if (par.mode >= 1 && gc < (int)h_cs.size())
      n1 += par.lanczos ? std::max(h_cs[gc].lfw, 0) + std::max(h_cs[gc].lbk, 0) : h_cs[gc].mfw + h_cs[gc].mbk;
    const double hi = hf_its.empty() ? 0 : hf_its[i], si = sc_its.empty() ? 0 : sc_its[i];
    w[3] += hi + si;
    w[4] += hi * n0 * n0 * n0 + si * n1 * n1 * n1;
  }
  w[2] = 9.0 * w[4];
  for (int gc = 0; n >= 17 && gc < (int)h_cs.size() && par.lanczos; ++gc) {
    w[15] += std::max(h_cs[gc].mfw, 0) + std::max(h_cs[gc].mbk, 0);
    w[16] += std::max(h_cs[gc].lmfw, 0) + std::max(h_cs[gc].lmbk, 0);
  }
  return BCD_OK;
}

int Batch::prepare() {
  HTRY(hipSetDevice(device));
  HTRY(hipMemcpyAsync(pe_[0], pe_[2], sizeof(double) * (prop_e_len + prop_x_len), hipMemcpyDeviceToDevice, stream));
  HTRY(hipMemcpyAsync(pnqp_[0], pnqp_[2], sizeof(int) * 2 * prop_n_len, hipMemcpyDeviceToDevice, stream));
  for (InstState &s : h_is) { std::memset(&s, 0, sizeof s); s.cur = 0; s.active = 1; }
  HTRY(hipMemcpyAsync(d_is, h_is.data(), sizeof(InstState) * n_inst, hipMemcpyHostToDevice, stream));
  HTRY(hipStreamSynchronize(stream));
  return BCD_OK;
}

void Batch::note_error(const std::string &what) {
  if (!err_detail.empty()) err_detail += "; ";
  err_detail += what;
}

int Batch::solve() {
  t_hip_error.clear();
  const int rc = solve_impl();
  if (rc == BCD_ERR_DEVICE && !t_hip_error.empty()) note_error(t_hip_error);
  return rc;
}

int Batch::solve_impl() {
  HTRY(hipSetDevice(device));
  err_detail.clear();
  for (int k = 0; k < KT_N; ++k) { ktime[k] = 0.0; kcount[k] = 0; }
  hf_its.assign(n_inst, 0);
  sc_its.assign(n_inst, 0);
  hf_ok.assign(n_inst, 0);
  sc_ok.assign(n_inst, par.mode == 0 ? 1 : 0);
  seeded.assign(n_inst, 0);
  fell_back.assign(n_inst, 0);
  first_its.assign(n_inst, 0);
  resid.assign(n_inst, -1.0);
  learn_valid = false;
  seed_pending = false;
  stage_mask.clear();
  HTRY(hipMemsetAsync(d_status, 0, sizeof(int), stream));
  HTRY(hipMemsetAsync(d_work, 0, sizeof(double), stream));
  HTRY(hipMemsetAsync(d_prof, 0, 8 * sizeof(unsigned long long) * n_gc, stream));
  HTRY(hipEventRecord(ev[0], stream));
  // start from the initial propagator in buffer 0
  HTRY(hipMemcpyAsync(pe_[0], pe_[2], sizeof(double) * (prop_e_len + prop_x_len), hipMemcpyDeviceToDevice, stream));
  HTRY(hipMemcpyAsync(pnqp_[0], pnqp_[2], sizeof(int) * 2 * prop_n_len, hipMemcpyDeviceToDevice, stream));
  for (InstState &s : h_is) { std::memset(&s, 0, sizeof s); s.cur = 0; s.active = 1; }
  HTRY(hipMemcpyAsync(d_is, h_is.data(), sizeof(InstState) * n_inst, hipMemcpyHostToDevice, stream));
  // W depends on the interaction and the channel layout only, both fixed for
  // the batch: it is built by the first solve and kept
  kt_begin(KT_WBUILD);
  if (!w_built) {
    hipLaunchKernelGGL(k_wbuild, dim3(n_gc), dim3(256), 0, stream, d_insts, d_chans, 0, n_gc, d_w);
    HTRY(hipGetLastError());
    w_built = true;
  }
  kt_end(KT_WBUILD);
  HTRY(hipEventRecord(ev[1], stream));
  BTRY(stage(0, par.hf_itr_max, par.hf_conv, true));
  HTRY(hipEventRecord(ev[2], stream));
  if (par.mode >= 1) {
    BTRY(build_sigma2(-1));
    // the learned start's inputs and the HF snapshot its fallback restarts
    // from, taken while the propagator is the HF one and the unreduced rows
    // are in place
    const bool checked = par.mode == 2 && (mlp || par.accel > 0) && fallback_tol > 0.0;
    const bool learned = mlp && par.mode == 2;
    if (learn_record || checked || learned) BTRY(learn_features(learn_record || learned));
    if (learned) {
      BTRY(mlp->reserve(n_gc));
      for (int i = 0; i < n_inst; ++i) seeded[i] = h_is[i].failed ? 0 : 1;
      HTRY(hipMemcpyAsync(d_seed_on, seeded.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice, stream));
    }
    HTRY(hipEventRecord(ev[3], stream));
    if (par.lanczos) BTRY(lanczos());
    if (par.lanczos && par.dyson_solver == 0) BTRY(decompose_poles());
    HTRY(hipEventRecord(ev[4], stream));
    const int src = par.lanczos ? 2 : 1;
    learn_hook = par.mode == 2 && (learn_record || learned);
    const int rc = stage(src, par.mode == 1 ? 0 : par.sc_itr_max, par.sc_conv, false);
    learn_hook = false;
    seed_pending = false;
    BTRY(rc);
    if (checked) BTRY(verify_and_fallback(src));
    else first_its = sc_its;
    if (learn_record) {
      hipLaunchKernelGGL(k_learn_density, dim3(n_gc), dim3(64), 0, stream, d_chans, n_gc, view(true), d_hf_e, d_hf_x,
                         d_hf_n, d_target, kLearnOut, 0);
      HTRY(hipGetLastError());
      learn_valid = true;
    }
  } else {
    HTRY(hipEventRecord(ev[3], stream));
    HTRY(hipEventRecord(ev[4], stream));
  }
  HTRY(hipEventRecord(ev[5], stream));
  HTRY(hipEventSynchronize(ev[5]));
  for (int k = 0; k < 5; ++k) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
    stage_ms[k] = ms;
  }
  float tot = 0.f;
  (void)hipEventElapsedTime(&tot, ev[0], ev[5]);
  stage_ms[5] = tot;
  HTRY(bcd::sync_memcpy(h_is.data(), d_is, sizeof(InstState) * n_inst, hipMemcpyDeviceToHost));
  return BCD_OK;
}

int Batch::learn_alloc() {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t g = (size_t)n_gc, n = (size_t)n_inst;
  const size_t b_feat = al(sizeof(float) * g * kLearnIn), b_out = al(sizeof(float) * g * kLearnOut);
  const size_t b_if = al(sizeof(float) * n * kLearnInstFeat), b_he = al(sizeof(double) * g * kMaxNI);
  const size_t b_hx = al(sizeof(double) * g * kMaxNI * kMaxNI), b_hn = al(sizeof(int) * 2 * g);
  const size_t b_n = al(sizeof(int) * n);
  const size_t tot = b_feat + 2 * b_out + b_if + b_he + b_hx + b_hn + 2 * b_n;
  const bool fresh = a_learn.cap < tot || !d_feat;
  BTRY(a_learn.grow(tot));
  char *q = (char *)a_learn.ptr;
  d_feat = (float *)q; q += b_feat;
  d_target = (float *)q; q += b_out;
  d_pred = (float *)q; q += b_out;
  d_inst_feat = (float *)q; q += b_if;
  d_hf_e = (double *)q; q += b_he;
  d_hf_x = (double *)q; q += b_hx;
  d_hf_n = (int *)q; q += b_hn;
  d_seed_on = (int *)q; q += b_n;
  d_flag = (int *)q;
  if (fresh) {
    HTRY(hipMemcpyAsync(d_inst_feat, h_inst_feat.data(), sizeof(float) * h_inst_feat.size(), hipMemcpyHostToDevice,
                        stream));
    HTRY(hipMemsetAsync(d_target, 0, b_out, stream));   // outputs past the first kLearnTri stay 0
  }
  return BCD_OK;
}

// After the first sc0 sweep: its density D1 into the inputs and, with a
// network, the prediction that the next sweep starts from.
int Batch::after_first_sweep() {
  hipLaunchKernelGGL(k_learn_density, dim3(n_gc), dim3(64), 0, stream, d_chans, n_gc, view(true), d_hf_e, d_hf_x,
                     d_hf_n, d_feat, kLearnIn, kFeatD1);
  HTRY(hipGetLastError());
  if (mlp) {
    BTRY(mlp->forward_dev(n_gc, d_feat, d_pred, stream));
    seed_pending = true;
  }
  return BCD_OK;
}

int Batch::learn_features(bool features) {
  BTRY(learn_alloc());
  hipLaunchKernelGGL(k_learn_features, dim3(n_gc), dim3(256), 0, stream, d_insts, d_chans, n_gc, view(true),
                     d_inst_feat, d_cs, d_row_fw_base, d_row_bk_base, d_rows_fw, d_rows_bk, d_feat, d_hf_e, d_hf_x,
                     d_hf_n, features ? 1 : 0);
  HTRY(hipGetLastError());
  return BCD_OK;
}

// After a seeded or density-mixed sc0 stage: one plain sweep from every
// instance's result measures its fixed-point residual (the stage's
// convergence measure of that sweep); instances whose residual exceeds
// fallback_tol, or that did not converge, are solved again by the plain
// iteration from their HF propagator.
int Batch::verify_and_fallback(int src) {
  const std::vector<int> its0 = sc_its, ok0 = sc_ok;
  first_its = its0;
  const int accel = par.accel;
  par.accel = 0;
  int rc = stage(src, 0, par.sc_conv, false);
  par.accel = accel;
  BTRY(rc);
  std::vector<int> flag(n_inst, 0);
  int nfb = 0;
  for (int i = 0; i < n_inst; ++i) {
    sc_its[i] = its0[i] + (h_is[i].failed ? 0 : 1);
    if (h_is[i].failed) { sc_ok[i] = 0; continue; }
    resid[i] = h_is[i].diff;
    const bool good = ok0[i] && resid[i] <= fallback_tol;
    sc_ok[i] = good ? 1 : 0;
    if (!good) { flag[i] = 1; ++nfb; }
  }
  if (!nfb) return BCD_OK;
  HTRY(hipMemcpyAsync(d_flag, flag.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_learn_restore, dim3(n_gc), dim3(64), 0, stream, d_chans, n_gc, view(true), d_flag, d_hf_e,
                     d_hf_x, d_hf_n);
  HTRY(hipGetLastError());
  const std::vector<int> its1 = sc_its, ok1 = sc_ok;
  stage_mask.assign(flag.begin(), flag.end());
  par.accel = 0;
  rc = stage(src, par.sc_itr_max, par.sc_conv, false);
  par.accel = accel;
  stage_mask.clear();
  BTRY(rc);
  for (int i = 0; i < n_inst; ++i) {
    if (flag[i]) {
      fell_back[i] = 1;
      sc_its[i] = its1[i] + h_is[i].its;
    } else {
      sc_its[i] = its1[i];
      sc_ok[i] = ok1[i];
    }
  }
  return BCD_OK;
}

int Batch::result(int inst, int nspb, int *np, int *nh, int *n_qp, int *n_qh, double *ep, double *zp, double *fqp,
                  double *eh, double *zh, double *fqh) {
  if (inst < 0 || inst >= n_inst) return BCD_ERR_ARG;
  const InstDev &I = h_insts[inst];
  const HostSystem &S = *sys[inst];
  const int cur = h_is[inst].cur;
  std::vector<int> cq(S.nsubsh), ch_(S.nsubsh);
  HTRY(bcd::sync_memcpy(cq.data(), pnqp_[cur] + I.gc0, sizeof(int) * S.nsubsh, hipMemcpyDeviceToHost));
  HTRY(bcd::sync_memcpy(ch_.data(), pnqh_[cur] + I.gc0, sizeof(int) * S.nsubsh, hipMemcpyDeviceToHost));
  int tq = 0, th = 0;
  for (int s = 0; s < S.nsubsh; ++s) { tq += cq[s]; th += ch_[s]; }
  if (n_qp) *n_qp = tq;
  if (n_qh) *n_qh = th;
  if (!ep) return BCD_OK;
  if (nspb < S.nspb) return BCD_ERR_ARG;
  std::vector<long> eo(n_gc), xo(n_gc);
  HTRY(bcd::sync_memcpy(eo.data(), d_e_off, sizeof(long) * n_gc, hipMemcpyDeviceToHost));
  HTRY(bcd::sync_memcpy(xo.data(), d_x_off, sizeof(long) * n_gc, hipMemcpyDeviceToHost));
  int kp = 0, kh = 0;
  for (int s = 0; s < S.nsubsh; ++s) {
    const int gc = I.gc0 + s, nI = S.no[s], tot = cq[s] + ch_[s];
    std::vector<double> e(std::max(tot, 1)), x(std::max(tot * nI, 1));
    if (tot) {
      HTRY(bcd::sync_memcpy(e.data(), pe_[cur] + eo[gc], sizeof(double) * tot, hipMemcpyDeviceToHost));
      HTRY(bcd::sync_memcpy(x.data(), px_[cur] + xo[gc], sizeof(double) * tot * nI, hipMemcpyDeviceToHost));
    }
    if (np) np[s] = cq[s];
    if (nh) nh[s] = ch_[s];
    for (int k = 0; k < tot; ++k) {
      const bool qp = k < cq[s];
      double *E = qp ? ep : eh, *Z = qp ? zp : zh, *F = qp ? fqp : fqh;
      const int j = qp ? kp++ : kh++;
      E[j] = e[k];
      double z = 0.0;
      for (int r = 0; r < nspb; ++r) {
        const double v = r < nI ? x[(long)k * nI + r] : 0.0;
        F[(long)j * nspb + r] = v;
        z += v * v;
      }
      Z[j] = z;
    }
  }
  return BCD_OK;
}

}  // namespace bcd
