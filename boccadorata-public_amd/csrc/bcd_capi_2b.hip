// bcd_capi_2b.hip -- the extern "C" boundary of the two-body propagator
// matrices: the ladder of gII_prop_t and the ring of Pol_prop_t
// (include/bcdor_amd.h, "two-body propagator matrices").  Bases are
// enumerated on the host as the reference does; every matrix element is
// computed on the GPU (bcd_twobody.hip).  Nothing falls back to the CPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "bcd_hipsync.hpp"
#include "bcd_ctx.hpp"
#include "bcdor_amd.h"

using namespace bcd;

struct bcd_t2 {
  bcd_ctx *ctx = nullptr;
  HostFrags frags;
  std::vector<void *> bufs;         // uploaded fragments
  FragDev fd{};
  T2Plan plan;
  bool profile = false;
  hipEvent_t ev[kT2Kernels + 1] = {};
  double kt_ms[kT2Kernels] = {};
  long kt_n[kT2Kernels] = {};
  ~bcd_t2() {
    for (void *p : bufs) (void)bcd::sync_free(p);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

#define HTRY(x)                                  \
  do {                                           \
    hipError_t e_ = (x);                         \
    if (e_ != hipSuccess) return BCD_ERR_DEVICE; \
  } while (0)

#define BTRY(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != BCD_OK) return r_; \
  } while (0)

bool prop_ok(const HostSystem &S, const bcd_prop *g) {
  if (!g || !g->np || !g->nh || g->nspb < S.nspb) return false;
  for (int s = 0; s < S.nsubsh; ++s)
    if (g->np[s] < 0 || g->nh[s] < 0) return false;
  return true;
}

// Fragment subshells and energies of g (n_clj_p / n_clj_h: fragments are
// stored subshell after subshell).
HostFrags host_frags(const HostSystem &S, const bcd_prop *g) {
  HostFrags F;
  for (int s = 0; s < S.nsubsh; ++s) {
    for (int k = 0; k < g->np[s]; ++k) F.sub[0].push_back(s);
    for (int k = 0; k < g->nh[s]; ++k) F.sub[1].push_back(s);
  }
  F.n[0] = (int)F.sub[0].size();
  F.n[1] = (int)F.sub[1].size();
  F.e[0].assign(g->ep, g->ep + F.n[0]);
  F.e[1].assign(g->eh, g->eh + F.n[1]);
  return F;
}

template <class T>
int upload(std::vector<void *> &bufs, const T *src, size_t n, const T **dst) {
  void *p = nullptr;
  if (bcd::sync_malloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return BCD_ERR_NOMEM;
  bufs.push_back(p);
  if (n && bcd::sync_memcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return BCD_ERR_DEVICE;
  *dst = (const T *)p;
  return BCD_OK;
}

int upload_frags(const HostFrags &F, const bcd_prop *g, std::vector<void *> &bufs, FragDev &fd) {
  fd.nspb = g->nspb;
  BTRY(upload(bufs, g->fqp, (size_t)F.n[0] * g->nspb, &fd.fq[0]));
  BTRY(upload(bufs, g->fqh, (size_t)F.n[1] * g->nspb, &fd.fq[1]));
  BTRY(upload(bufs, F.sub[0].data(), F.sub[0].size(), &fd.sub[0]));
  BTRY(upload(bufs, F.sub[1].data(), F.sub[1].size(), &fd.sub[1]));
  return BCD_OK;
}

int ensure_t2(bcd_ctx *ctx) {
  if (ctx->t2) return BCD_OK;
  std::unique_ptr<T2System> t(new (std::nothrow) T2System);
  if (!t) return BCD_ERR_NOMEM;
  HTRY(hipSetDevice(ctx->device));
  BTRY(t->upload(ctx->sys));
  ctx->t2 = std::move(t);
  return BCD_OK;
}

int basis_common(bcd_ctx *ctx, const bcd_prop *g, int kind, int J, int dT, int pif, int icut, int ntcut,
                 const int *n1p, const int *n1h, T2Job &job) {
  if (!ctx || !prop_ok(ctx->sys, g) || J < 0 || pif < 0 || pif > 1) return BCD_ERR_ARG;
  const HostFrags F = host_frags(ctx->sys, g);
  return kind == 0 ? t2_ladd_basis(ctx->sys, F, J, dT, pif, icut, ntcut, n1p, n1h, job)
                   : t2_ph_basis(ctx->sys, F, J, dT, pif, icut, ntcut, n1p, n1h, job);
}

int count_common(bcd_ctx *ctx, const bcd_prop *g, int kind, int J, int dT, int pif, int icut, int ntcut,
                 const int *n1p, const int *n1h, int *na, int *nb) {
  if (!na || !nb) return BCD_ERR_ARG;
  T2Job job;
  BTRY(basis_common(ctx, g, kind, J, dT, pif, icut, ntcut, n1p, n1h, job));
  *na = (int)job.ea.size();
  *nb = (int)job.eb.size();
  return BCD_OK;
}

int basis_out(bcd_ctx *ctx, const bcd_prop *g, int kind, int J, int dT, int pif, int icut, int ntcut,
              const int *n1p, const int *n1h, int cap_a, int cap_b, int *a1, int *a2, double *ea, int *b1,
              int *b2, double *eb, int *na, int *nb) {
  if (!na || !nb) return BCD_ERR_ARG;
  T2Job job;
  BTRY(basis_common(ctx, g, kind, J, dT, pif, icut, ntcut, n1p, n1h, job));
  *na = (int)job.ea.size();
  *nb = (int)job.eb.size();
  if (*na > cap_a || *nb > cap_b) return BCD_ERR_CAPACITY;
  if ((*na && (!a1 || !a2 || !ea)) || (*nb && (!b1 || !b2 || !eb))) return BCD_ERR_ARG;
  std::copy(job.a1.begin(), job.a1.end(), a1);
  std::copy(job.a2.begin(), job.a2.end(), a2);
  std::copy(job.ea.begin(), job.ea.end(), ea);
  std::copy(job.b1.begin(), job.b1.end(), b1);
  std::copy(job.b2.begin(), job.b2.end(), b2);
  std::copy(job.eb.begin(), job.eb.end(), eb);
  return BCD_OK;
}

// Checks a caller-supplied basis against the channel and the propagator:
// ladder rows are (qp, qp) / (qh, qh) pairs of total charge dT, ring rows
// (qp, qh) pairs of charge difference +dT (ph) / -dT (hp).
int basis_check(const HostSystem &S, const HostFrags &F, const T2Job &j) {
  auto ok = [&](int side_u, int u, int side_w, int w, int want_dch) {
    if (u < 0 || u >= F.n[side_u] || w < 0 || w >= F.n[side_w]) return false;
    const int su = F.sub[side_u][u], sw = F.sub[side_w][w];
    if ((j.pi + S.ip[su] + S.ip[sw]) % 2) return false;
    const int s1 = S.j2[su] + S.j2[sw] - 2 * j.J, s2 = S.j2[sw] + 2 * j.J - S.j2[su];
    const int s3 = 2 * j.J + S.j2[su] - S.j2[sw];
    if (s1 < 0 || s2 < 0 || s3 < 0 || ((s1 | s2 | s3) & 1)) return false;
    return j.kind == 0 ? S.ch[su] + S.ch[sw] == j.dT : S.ch[su] - S.ch[sw] == want_dch;
  };
  for (size_t k = 0; k < j.ea.size(); ++k)
    if (!(j.kind == 0 ? ok(0, j.a1[k], 0, j.a2[k], 0) : ok(0, j.a1[k], 1, j.a2[k], j.dT))) return BCD_ERR_ARG;
  for (size_t k = 0; k < j.eb.size(); ++k)
    if (!(j.kind == 0 ? ok(1, j.b1[k], 1, j.b2[k], 0) : ok(0, j.b1[k], 1, j.b2[k], -j.dT))) return BCD_ERR_ARG;
  return BCD_OK;
}

// One matrix through the batched machinery; copies back exactly the extent
// Fill_*_mtx writes.
int fill_common(bcd_ctx *ctx, const bcd_prop *g, T2Job job, double *Sab, int ld, double *free_poles) {
  if (!ctx || !prop_ok(ctx->sys, g) || !Sab) return BCD_ERR_ARG;
  const int ndim = t2_job_dims(job);
  if (ndim < 0) return ndim;          // -100: invalid mode
  // the reference checks ld < 1 first for the ladder (gII_b.cpp:332) and last
  // for the ring (PolProp_b.cpp:291-301); ndim > ld stops its program (-102 here)
  if (job.kind == 0 && ld < 1) return -101;
  if (ndim > ld) return -102;
  if (ld < 1) return -101;
  const HostFrags F = host_frags(ctx->sys, g);
  BTRY(basis_check(ctx->sys, F, job));
  BTRY(ensure_t2(ctx));
  HTRY(hipSetDevice(ctx->device));
  bcd_t2 t;
  t.ctx = ctx;
  BTRY(upload_frags(F, g, t.bufs, t.fd));
  job.ld = ld;
  t.plan.jobs.push_back(job);
  BTRY(t.plan.prepare(*ctx->t2, ctx->sys, F, ctx->stream));
  BTRY(t.plan.launch(*ctx->t2, t.fd, ctx->stream));
  HTRY(hipStreamSynchronize(ctx->stream));
  if (ndim > 0)
    HTRY(bcd::sync_memcpy(Sab, t.plan.d_out + t.plan.jobs[0].out_off, sizeof(double) * ((size_t)(ndim - 1) * ld + ndim),
                   hipMemcpyDeviceToHost));
  if (free_poles)
    for (int k = 0; k < ndim; ++k)
      if (!std::isnan(t.plan.jobs[0].free_poles[k])) free_poles[k] = t.plan.jobs[0].free_poles[k];
  return BCD_OK;
}

}  // namespace

extern "C" {

int bcd_ladd_count(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut, const int *n1p,
                   const int *n1h, int *npp, int *nhh) {
  return count_common(ctx, g, 0, J, dT, pif, icut, ntcut, n1p, n1h, npp, nhh);
}

int bcd_ladd_basis(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut, const int *n1p,
                   const int *n1h, int cap_pp, int cap_hh, int *n1_pp, int *n2_pp, double *e_pp, int *k1_hh,
                   int *k2_hh, double *e_hh, int *npp, int *nhh) {
  return basis_out(ctx, g, 0, J, dT, pif, icut, ntcut, n1p, n1h, cap_pp, cap_hh, n1_pp, n2_pp, e_pp, k1_hh, k2_hh,
                   e_hh, npp, nhh);
}

int bcd_ladd_fill(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int npp, const int *n1_pp,
                  const int *n2_pp, const double *e_pp, int nhh, const int *k1_hh, const int *k2_hh,
                  const double *e_hh, int i_RPAcalc, double *Sab, int ndim_vec_alloc, double *free_poles,
                  int add_free_poles) {
  if (npp < 0 || nhh < 0 || (npp && (!n1_pp || !n2_pp || !e_pp)) || (nhh && (!k1_hh || !k2_hh || !e_hh)))
    return BCD_ERR_ARG;
  T2Job j;
  j.kind = 0; j.J = J; j.dT = dT; j.pi = pif; j.mode = i_RPAcalc; j.add = add_free_poles != 0;
  j.a1.assign(n1_pp, n1_pp + npp); j.a2.assign(n2_pp, n2_pp + npp); j.ea.assign(e_pp, e_pp + npp);
  j.b1.assign(k1_hh, k1_hh + nhh); j.b2.assign(k2_hh, k2_hh + nhh); j.eb.assign(e_hh, e_hh + nhh);
  return fill_common(ctx, g, std::move(j), Sab, ndim_vec_alloc, free_poles);
}

int bcd_ph_count(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut, const int *n1p,
                 const int *n1h, int *nph, int *nhp) {
  return count_common(ctx, g, 1, J, dT, pif, icut, ntcut, n1p, n1h, nph, nhp);
}

int bcd_ph_basis(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut, const int *n1p,
                 const int *n1h, int cap_ph, int cap_hp, int *n_ph, int *k_ph, double *e_ph, int *n_hp, int *k_hp,
                 double *e_hp, int *nph, int *nhp) {
  return basis_out(ctx, g, 1, J, dT, pif, icut, ntcut, n1p, n1h, cap_ph, cap_hp, n_ph, k_ph, e_ph, n_hp, k_hp,
                   e_hp, nph, nhp);
}

int bcd_ph_fill(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int nph, const int *n_ph, const int *k_ph,
                const double *e_ph, int nhp, const int *n_hp, const int *k_hp, const double *e_hp, int i_RPAcalc,
                double *Sab, int ndim_vec_alloc, double *free_poles, int add_free_poles) {
  if (nph < 0 || nhp < 0 || (nph && (!n_ph || !k_ph || !e_ph)) || (nhp && (!n_hp || !k_hp || !e_hp)))
    return BCD_ERR_ARG;
  T2Job j;
  j.kind = 1; j.J = J; j.dT = dT; j.pi = pif; j.mode = i_RPAcalc; j.add = add_free_poles != 0;
  j.a1.assign(n_ph, n_ph + nph); j.a2.assign(k_ph, k_ph + nph); j.ea.assign(e_ph, e_ph + nph);
  j.b1.assign(n_hp, n_hp + nhp); j.b2.assign(k_hp, k_hp + nhp); j.eb.assign(e_hp, e_hp + nhp);
  return fill_common(ctx, g, std::move(j), Sab, ndim_vec_alloc, free_poles);
}

int bcd_t2_create(bcd_ctx *ctx, const bcd_prop *g, int ladd_mode, int ring_mode, int add_free_poles,
                  bcd_t2 **out) {
  if (!out) return BCD_ERR_ARG;
  *out = nullptr;
  if (!ctx || !prop_ok(ctx->sys, g) || ladd_mode < -1 || ladd_mode > 4 || ring_mode < -1 || ring_mode > 4 ||
      ring_mode == 2)
    return BCD_ERR_ARG;
  const HostSystem &S = ctx->sys;
  std::unique_ptr<bcd_t2> t(new (std::nothrow) bcd_t2);
  if (!t) return BCD_ERR_NOMEM;
  t->ctx = ctx;
  t->frags = host_frags(S, g);
  // charge bounds of Calculate_Jch_bounds (Mod_Space.cpp:589-620)
  int pp_mn = 1 << 20, pp_mx = -(1 << 20), ph_mn = 0, ph_mx = 0;
  for (int a = 0; a < S.nsubsh; ++a)
    for (int b = a; b < S.nsubsh; ++b) {
      pp_mn = std::min(pp_mn, S.ch[a] + S.ch[b]);
      pp_mx = std::max(pp_mx, S.ch[a] + S.ch[b]);
      ph_mn = std::min(ph_mn, std::min(S.ch[a] - S.ch[b], S.ch[b] - S.ch[a]));
      ph_mx = std::max(ph_mx, std::max(S.ch[a] - S.ch[b], S.ch[b] - S.ch[a]));
    }
  for (int J = 0; J <= S.jmax; ++J)
    for (int pi = 0; pi < 2; ++pi) {
      if (ladd_mode >= 0)
        for (int dT = pp_mn; dT <= pp_mx; ++dT) {
          T2Job j;
          BTRY(t2_ladd_basis(S, t->frags, J, dT, pi, 0, 0, nullptr, nullptr, j));
          if (j.ea.empty() && j.eb.empty()) continue;
          j.mode = ladd_mode; j.add = add_free_poles != 0;
          t->plan.jobs.push_back(std::move(j));
        }
      if (ring_mode >= 0)
        for (int dT = ph_mn; dT <= ph_mx; ++dT) {
          T2Job j;
          BTRY(t2_ph_basis(S, t->frags, J, dT, pi, 0, 0, nullptr, nullptr, j));
          if (j.ea.empty() && j.eb.empty()) continue;
          j.mode = ring_mode; j.add = add_free_poles != 0;
          t->plan.jobs.push_back(std::move(j));
        }
    }
  BTRY(ensure_t2(ctx));
  HTRY(hipSetDevice(ctx->device));
  BTRY(upload_frags(t->frags, g, t->bufs, t->fd));
  BTRY(t->plan.prepare(*ctx->t2, S, t->frags, ctx->stream));
  for (hipEvent_t &e : t->ev) HTRY(hipEventCreate(&e));
  *out = t.release();
  return BCD_OK;
}

void bcd_t2_destroy(bcd_t2 *t) { delete t; }

int bcd_t2_count(const bcd_t2 *t) { return t ? (int)t->plan.jobs.size() : BCD_ERR_ARG; }

int bcd_t2_info(const bcd_t2 *t, int k, int *info) {
  if (!t || !info || k < 0 || k >= (int)t->plan.jobs.size()) return BCD_ERR_ARG;
  const T2Job &j = t->plan.jobs[k];
  const int v[10] = {j.kind, j.J, j.pi, j.dT, j.mode, j.nf, j.nb, j.ld, (int)j.ea.size(), (int)j.eb.size()};
  std::copy(v, v + 10, info);
  return BCD_OK;
}

int bcd_t2_basis(const bcd_t2 *t, int k, int *a1, int *a2, double *ea, int *b1, int *b2, double *eb) {
  if (!t || k < 0 || k >= (int)t->plan.jobs.size()) return BCD_ERR_ARG;
  const T2Job &j = t->plan.jobs[k];
  if (a1) std::copy(j.a1.begin(), j.a1.end(), a1);
  if (a2) std::copy(j.a2.begin(), j.a2.end(), a2);
  if (ea) std::copy(j.ea.begin(), j.ea.end(), ea);
  if (b1) std::copy(j.b1.begin(), j.b1.end(), b1);
  if (b2) std::copy(j.b2.begin(), j.b2.end(), b2);
  if (eb) std::copy(j.eb.begin(), j.eb.end(), eb);
  return BCD_OK;
}

int bcd_t2_build(bcd_t2 *t) {
  if (!t) return BCD_ERR_ARG;
  bcd_ctx *ctx = t->ctx;
  HTRY(hipSetDevice(ctx->device));
  BTRY(t->plan.launch(*ctx->t2, t->fd, ctx->stream, t->profile ? t->ev : nullptr));
  HTRY(hipStreamSynchronize(ctx->stream));
  if (t->profile)
    for (int k = 0; k < kT2Kernels; ++k) {
      float ms = 0.f;
      HTRY(hipEventElapsedTime(&ms, t->ev[k], t->ev[k + 1]));
      t->kt_ms[k] += ms;
      ++t->kt_n[k];
    }
  return BCD_OK;
}

int bcd_t2_matrix(const bcd_t2 *t, int k, double *Sab, double *free_poles) {
  if (!t || k < 0 || k >= (int)t->plan.jobs.size()) return BCD_ERR_ARG;
  const T2Job &j = t->plan.jobs[k];
  const int ndim = j.nf + j.nb;
  if (Sab && ndim > 0)
    HTRY(bcd::sync_memcpy(Sab, t->plan.d_out + j.out_off, sizeof(double) * ((size_t)(ndim - 1) * j.ld + ndim),
                   hipMemcpyDeviceToHost));
  if (free_poles)
    for (int i = 0; i < ndim; ++i)
      if (!std::isnan(j.free_poles[i])) free_poles[i] = j.free_poles[i];
  return BCD_OK;
}

int bcd_t2_set_profile(bcd_t2 *t, int on) {
  if (!t) return BCD_ERR_ARG;
  t->profile = on != 0;
  for (int k = 0; k < kT2Kernels; ++k) { t->kt_ms[k] = 0.0; t->kt_n[k] = 0; }
  return BCD_OK;
}

int bcd_t2_kernel_times(const bcd_t2 *t, int k, double *avg_ms, long *launches) {
  if (!t || k < 0 || k >= kT2Kernels || !avg_ms || !launches) return BCD_ERR_ARG;
  *launches = t->kt_n[k];
  *avg_ms = t->kt_n[k] ? t->kt_ms[k] / t->kt_n[k] : 0.0;
  return BCD_OK;
}

int bcd_t2_work(const bcd_t2 *t, double *work, int n) {
  if (!t || !work || n < 3) return BCD_ERR_ARG;
  work[0] = t->plan.flops_g1;
  work[1] = t->plan.flops_g2;
  work[2] = t->plan.out_doubles * 8.0;
  return BCD_OK;
}

}  // extern "C"
