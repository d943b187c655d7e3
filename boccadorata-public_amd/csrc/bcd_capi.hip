// bcd_capi.hip -- the extern "C" boundary (include/bcdor_amd.h).
//
// Every entry point validates its arguments on the host, then runs the work
// on the GPU through the kernels of bcd_kernels.hip.  Without a device the
// entry points return BCD_ERR_DEVICE; nothing falls back to the CPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "bcd_hipsync.hpp"
#include "bcd_batch.hpp"
#include "bcd_mlp.hpp"
#include "bcd_ctx.hpp"
#include "bcdor_amd.h"

using namespace bcd;

struct bcd_batch {
  // instances added one at a time (bcd_batch_begin / _add / _end): each one's
  // interaction goes to HBM as it is added and its host copy is freed, so
  // the host never holds more than the instances being added
  std::vector<DevArena> vbufs;
  std::vector<std::unique_ptr<HostSystem>> systems;
  std::vector<HostProp> g0;
  std::vector<int> lanc_n;
  bcd_solve_params par{};
  int device = 0;
  bool ready = false;
  Batch b;
  ~bcd_batch() {
    for (DevArena &a : vbufs) a.release();
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

int device_ok(int dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || dev < 0 || dev >= n) return BCD_ERR_DEVICE;
  return BCD_OK;
}

HostProp to_host_prop(const HostSystem &S, const bcd_prop *g) {
  HostProp H;
  H.nspb = g->nspb;
  H.np.assign(g->np, g->np + S.nsubsh);
  H.nh.assign(g->nh, g->nh + S.nsubsh);
  long tp = 0, th = 0;
  for (int s = 0; s < S.nsubsh; ++s) { tp += g->np[s]; th += g->nh[s]; }
  H.ep.assign(g->ep, g->ep + tp);
  H.eh.assign(g->eh, g->eh + th);
  H.fqp.assign(g->fqp, g->fqp + tp * g->nspb);
  H.fqh.assign(g->fqh, g->fqh + th * g->nspb);
  return H;
}

// SpProp_t(ModSpace_t*, i_xe=1) (Sp_Propagator.cpp:292-378): holes at spe-10,
// particles at spe, unit overlaps, orbit order.
HostProp bare_prop(const HostSystem &S) {
  HostProp H;
  H.nspb = S.nspb;
  H.np.assign(S.nsubsh, 0);
  H.nh.assign(S.nsubsh, 0);
  for (int s = 0; s < S.nsubsh; ++s)
    for (int k = 0; k < S.no[s]; ++k) {
      const int o = S.orb0[s] + k;
      std::vector<double> x(S.nspb, 0.0);
      if (S.orb_n[o] < S.no[s]) x[S.orb_n[o]] = 1.0;
      if (S.orb_F[o] == 1) {
        H.eh.push_back(S.orb_e[o] - 10.0);
        H.fqh.insert(H.fqh.end(), x.begin(), x.end());
        ++H.nh[s];
      } else {
        H.ep.push_back(S.orb_e[o]);
        H.fqp.insert(H.fqp.end(), x.begin(), x.end());
        ++H.np[s];
      }
    }
  return H;
}

bool prop_ok(const HostSystem &S, const bcd_prop *g) {
  if (!g || !g->np || !g->nh || g->nspb < S.nspb) return false;
  for (int s = 0; s < S.nsubsh; ++s)
    if (g->np[s] < 0 || g->nh[s] < 0) return false;
  return true;
}

// A one-instance batch over the context's system, started from g.
int stage_batch(bcd_ctx *ctx, const bcd_prop *g, int mode, Batch &b) {
  std::vector<const HostSystem *> one{&ctx->sys};
  bcd_solve_params p;
  std::memset(&p, 0, sizeof p);
  p.mode = mode;
  p.hf_itr_max = 0;
  p.sc_itr_max = 0;
  p.conv_kind = 0;
  p.hf_conv = p.sc_conv = 1e-7;
  std::vector<HostProp> g0{to_host_prop(ctx->sys, g)};
  BTRY(b.setup(ctx->device, one, p, g0, true));
  return b.prepare();
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)bcd::sync_free(p); }
  int alloc(size_t bytes) { return bcd::sync_malloc(&p, bytes > 0 ? bytes : 8) == hipSuccess ? BCD_OK : BCD_ERR_NOMEM; }
};

// Argument checks of Reduce_Mtx_Lncz (Lanczos.cpp:50-82), returning its
// error codes instead of exiting; *nLanc is the number of vectors to build.
int lanczos_checks(int NA_out, int NB_out, int NPLS_out, int NB_in, int NPLS_in, int nt_pvts, const int *n_itr,
                   int *nLanc_out) {
  if (NPLS_in < 1 || NPLS_out < 1) return -1;
  if (NB_out != NB_in) return -100;
  if (NB_in > kMaxNB || NB_in < 1 || nt_pvts > kMaxPiv || nt_pvts < 1) return BCD_ERR_ARG;
  int nLanc = 0, nact = 0;
  for (int i = 0; i < nt_pvts; ++i)
    if (n_itr[i] > 0) { nLanc += n_itr[i]; ++nact; }
  if (nLanc < 1) return -2;
  if (nLanc > NPLS_in) nLanc = NPLS_in;
  if (nLanc > NPLS_out) return -101;
  if (nact + 1 > NA_out && nLanc < NPLS_in) return -102;
  if (lanczos_bucket(NPLS_in) < 0) return BCD_ERR_CAPACITY;
  *nLanc_out = nLanc;
  return BCD_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device layout of one Reduce_Mtx_Lncz call inside one region: the input rows
// (when they come from the host), pivots, iteration counts, output rows,
// i_piv, the row count, the job and the Krylov workspace.
struct LancLayout {
  size_t min, piv, itr, out, ipv, n, job, V, total;
};

LancLayout lanczos_layout(bool with_min, int NPLS_in, int Tin, int nt_pvts, int NB_in, int Tout, int nLanc) {
  LancLayout L;
  size_t o = 0;
  L.min = o; o += with_min ? align256(sizeof(double) * (size_t)NPLS_in * Tin) : 0;
  L.piv = o; o += align256(sizeof(double) * (size_t)nt_pvts * NB_in);
  L.itr = o; o += align256(sizeof(int) * (size_t)nt_pvts);
  L.out = o; o += align256(sizeof(double) * (size_t)Tout * nLanc);
  L.ipv = o; o += align256(sizeof(int) * (size_t)(nt_pvts + 2));
  L.n = o; o += align256(sizeof(int));
  L.job = o; o += align256(sizeof(LancJob));
  L.V = o; o += align256(sizeof(double) * (size_t)lanczos_ws_vectors(nLanc, NB_in) * lanczos_stride(NPLS_in));
  L.total = o;
  return L;
}

// One Lanczos job with every device buffer in the region at base (laid out
// by lanczos_layout).  The input rows are d_Min on the device, or h_Min
// copied into the region.  Allocates nothing.
int lanczos_in_region(bcd_ctx *ctx, char *base, const LancLayout &L, int nLanc, int NA_out, int NB_out, double *Mout,
                      int *i_piv_out, int NA_in, int NB_in, int NPLS_in, const double *h_Min, const double *d_Min,
                      int nt_pvts, double *pivots, int *n_itr) {
  const int Tout = NA_out + NB_out, Tin = NA_in + NB_in;
  if (h_Min) {
    HTRY(hipMemcpyAsync(base + L.min, h_Min, sizeof(double) * (size_t)NPLS_in * Tin, hipMemcpyHostToDevice,
                        ctx->stream));
    d_Min = (const double *)(base + L.min);
  }
  HTRY(hipMemcpyAsync(base + L.piv, pivots, sizeof(double) * nt_pvts * NB_in, hipMemcpyHostToDevice, ctx->stream));
  HTRY(hipMemcpyAsync(base + L.itr, n_itr, sizeof(int) * nt_pvts, hipMemcpyHostToDevice, ctx->stream));
  LancJob J;
  std::memset(&J, 0, sizeof J);
  J.Min = d_Min; J.Tin = Tin; J.NA_in = NA_in; J.NB = NB_in; J.M = NPLS_in;
  J.piv = (const double *)(base + L.piv); J.nt_pvts = nt_pvts; J.nLanc = nLanc; J.itr = (const int *)(base + L.itr);
  J.Mout = (double *)(base + L.out); J.Tout = Tout; J.NA_out = NA_out; J.i_piv_out = (int *)(base + L.ipv);
  J.n_out = (int *)(base + L.n); J.V = (double *)(base + L.V); J.inst = -1;
  HTRY(hipMemcpyAsync(base + L.job, &J, sizeof J, hipMemcpyHostToDevice, ctx->stream));
  launch_lanczos(kLancDefault[lanczos_bucket(NPLS_in)], 1, (const LancJob *)(base + L.job), nullptr, ctx->stream);
  HTRY(hipGetLastError());
  int nout = 0;
  HTRY(hipMemcpyAsync(&nout, base + L.n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HTRY(hipStreamSynchronize(ctx->stream));
  if (nout < 0) return nout;
  HTRY(bcd::sync_memcpy(Mout, base + L.out, sizeof(double) * (size_t)Tout * nLanc, hipMemcpyDeviceToHost));
  int npv = 0;
  HTRY(bcd::sync_memcpy(&npv, base + L.ipv, sizeof(int), hipMemcpyDeviceToHost));
  HTRY(bcd::sync_memcpy(i_piv_out, base + L.ipv, sizeof(int) * (npv + 2), hipMemcpyDeviceToHost));
  // the reference normalises every pivot it starts from in place
  int ip = -1;
  for (int k = 0; k < npv; ++k) {
    ++ip;
    while (ip < nt_pvts && n_itr[ip] < 1) ++ip;
    double x1 = 0.0;
    for (int l = 0; l < NB_in; ++l) x1 += pivots[ip * NB_in + l] * pivots[ip * NB_in + l];
    x1 = std::sqrt(x1);
    for (int l = 0; l < NB_in; ++l) pivots[ip * NB_in + l] /= x1;
  }
  return nout;
}

// One Lanczos job on device data (the rows in d_Min), buffers allocated for
// the call.
int run_lanczos(bcd_ctx *ctx, int NA_out, int NB_out, int NPLS_out, double *Mout, int *i_piv_out, int NA_in,
                int NB_in, int NPLS_in, const double *d_Min, int nt_pvts, double *pivots, int *n_itr) {
  int nLanc = 0;
  const int rc = lanczos_checks(NA_out, NB_out, NPLS_out, NB_in, NPLS_in, nt_pvts, n_itr, &nLanc);
  if (rc != BCD_OK) return rc;
  const LancLayout L = lanczos_layout(false, NPLS_in, NA_in + NB_in, nt_pvts, NB_in, NA_out + NB_out, nLanc);
  DevBuf region;
  BTRY(region.alloc(L.total));
  return lanczos_in_region(ctx, (char *)region.p, L, nLanc, NA_out, NB_out, Mout, i_piv_out, NA_in, NB_in, NPLS_in,
                           nullptr, d_Min, nt_pvts, pivots, n_itr);
}

// Scratch of bcd_reduce_mtx_lncz_noalloc: every device buffer of one call
// lives here (input rows, pivots, counts, output rows, i_piv, the job and the
// Krylov workspace), so that entry point allocates nothing.  Each device the
// library runs on holds one copy until the process ends; calls on one device
// share it.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
__device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of this device's scratch, looked up once per device
// with the device current (the array's name is only a host-side stand-in).
int noalloc_scratch(int device, char **out) {
  if (device < 0 || device >= kNoAllocMaxDevices) return BCD_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BCD_ERR_DEVICE;
    g_scratch_ptr[device] = (char *)p;
  }
  *out = g_scratch_ptr[device];
  return BCD_OK;
}

}  // namespace

extern "C" {

int bcd_version(void) { return 1; }

const char *bcd_error_string(int code) {
  switch (code) {
    case BCD_OK: return "ok";
    case BCD_ERR_ARG: return "invalid argument";
    case BCD_ERR_DEVICE: return "no usable GPU or a HIP call failed";
    case BCD_ERR_CAPACITY: return "output capacity or a kernel limit exceeded";
    case BCD_ERR_LANCZOS: return "Lanczos reduction failed";
    case BCD_ERR_EIGEN: return "eigensolver did not converge";
    case BCD_ERR_NOMEM: return "device allocation failed";
    default: return code < 0 ? "Reduce_Mtx_Lncz error code" : "unknown";
  }
}

int bcd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int bcd_ctx_create(int device, const bcd_modspace *ms, const bcd_vpp *vpp, bcd_ctx **out) {
  if (!out || !ms) return BCD_ERR_ARG;
  *out = nullptr;
  BTRY(device_ok(device));
  std::unique_ptr<bcd_ctx> c(new (std::nothrow) bcd_ctx);
  if (!c) return BCD_ERR_NOMEM;
  c->device = device;
  BTRY(c->sys.build(ms, vpp));
  HTRY(hipSetDevice(device));
  HTRY(hipStreamCreate(&c->stream));
  *out = c.release();
  return BCD_OK;
}

void bcd_ctx_destroy(bcd_ctx *ctx) {
  if (!ctx) return;
  ctx->t2.reset();
  ctx->scratch.release();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int bcd_count_poles(bcd_ctx *ctx, int ish, const bcd_prop *g, int *nfw, int *nbk) {
  if (!ctx || !nfw || !nbk || ish < 0 || ish >= ctx->sys.nsubsh || !prop_ok(ctx->sys, g)) return BCD_ERR_ARG;
  Batch b;
  BTRY(stage_batch(ctx, g, 1, b));
  PropView P = b.view(true);
  HTRY(hipMemsetAsync(b.d_work_count, 0, sizeof(int), b.stream));
  hipLaunchKernelGGL(k_sigma2_plan, dim3(b.n_gc), dim3(256), 0, b.stream, b.d_chans, b.d_bpack, 0, b.n_gc, P,
                     b.d_insts, b.d_cs, b.d_blk_fw_off, b.d_blk_bk_off, b.d_work_list, b.d_work_count, ish, nullptr);
  HTRY(hipGetLastError());
  ChanState cs;
  HTRY(hipMemcpyAsync(&cs, b.d_cs + ish, sizeof cs, hipMemcpyDeviceToHost, b.stream));
  HTRY(hipStreamSynchronize(b.stream));
  *nfw = cs.mfw;
  *nbk = cs.mbk;
  return BCD_OK;
}

int bcd_sigma2_build(bcd_ctx *ctx, int ish, const bcd_prop *g, double *rows_fw, int cap_fw, double *rows_bk,
                     int cap_bk, int *nfw, int *nbk) {
  if (!ctx || !nfw || !nbk || ish < 0 || ish >= ctx->sys.nsubsh || !prop_ok(ctx->sys, g)) return BCD_ERR_ARG;
  Batch b;
  BTRY(stage_batch(ctx, g, 1, b));
  BTRY(b.build_sigma2(ish));
  HTRY(hipStreamSynchronize(b.stream));
  const ChanState &cs = b.h_cs[ish];
  *nfw = cs.mfw;
  *nbk = cs.mbk;
  if (cs.mfw > cap_fw || cs.mbk > cap_bk) return BCD_ERR_CAPACITY;
  const int w = 2 + ctx->sys.no[ish];
  long bf = 0, bb = 0;
  HTRY(bcd::sync_memcpy(&bf, b.d_row_fw_base + ish, sizeof(long), hipMemcpyDeviceToHost));
  HTRY(bcd::sync_memcpy(&bb, b.d_row_bk_base + ish, sizeof(long), hipMemcpyDeviceToHost));
  if (cs.mfw && rows_fw) HTRY(bcd::sync_memcpy(rows_fw, b.d_rows_fw + bf, sizeof(double) * cs.mfw * w, hipMemcpyDeviceToHost));
  if (cs.mbk && rows_bk) HTRY(bcd::sync_memcpy(rows_bk, b.d_rows_bk + bb, sizeof(double) * cs.mbk * w, hipMemcpyDeviceToHost));
  return BCD_OK;
}

int bcd_hf_pot(bcd_ctx *ctx, int ish, const bcd_prop *g, double *u, int ndim) {
  if (!ctx || !u || ish < 0 || ish >= ctx->sys.nsubsh || ndim < 1 || !prop_ok(ctx->sys, g)) return BCD_ERR_ARG;
  Batch b;
  BTRY(stage_batch(ctx, g, 0, b));
  PropView P = b.view(true);
  hipLaunchKernelGGL(k_wbuild, dim3(1), dim3(256), 0, b.stream, b.d_insts, b.d_chans, ish, ish + 1, b.d_w);
  hipLaunchKernelGGL(k_density, dim3(b.n_gc), dim3(64), 0, b.stream, b.d_insts, b.d_chans, 0, b.n_gc, P, b.d_rho);
  hipLaunchKernelGGL(k_sigma_inf, dim3(1), dim3(256), 0, b.stream, b.d_insts, b.d_chans, ish, ish + 1, b.d_is,
                     b.d_w, b.d_rho, b.d_h);
  HTRY(hipGetLastError());
  const int n = ctx->sys.no[ish];
  std::vector<double> h(n * n);
  HTRY(hipMemcpyAsync(h.data(), b.d_h + b.h_chans[ish].h_off, sizeof(double) * n * n, hipMemcpyDeviceToHost,
                      b.stream));
  HTRY(hipStreamSynchronize(b.stream));
  for (int r = 0; r < n && r < ndim; ++r)
    for (int c = 0; c < n && c < ndim; ++c) u[r * ndim + c] = h[r * n + c];
  return n > ndim ? n - ndim : BCD_OK;
}

int bcd_reduce_mtx_lncz(bcd_ctx *ctx, int NA_out, int NB_out, int NPLS_out, double *Mout, int *i_piv_out,
                        int NA_in, int NB_in, int NPLS_in, const double *Min, int nt_pvts, double *pivots,
                        int *n_itr) {
  if (!ctx || !Mout || !i_piv_out || !Min || !pivots || !n_itr) return BCD_ERR_ARG;
  if (NPLS_in < 1 || NPLS_out < 1) return -1;
  HTRY(hipSetDevice(ctx->device));
  DevBuf dmin;
  BTRY(dmin.alloc(sizeof(double) * (size_t)NPLS_in * (NA_in + NB_in)));
  HTRY(bcd::sync_memcpy(dmin.p, Min, sizeof(double) * (size_t)NPLS_in * (NA_in + NB_in), hipMemcpyHostToDevice));
  return run_lanczos(ctx, NA_out, NB_out, NPLS_out, Mout, i_piv_out, NA_in, NB_in, NPLS_in,
                     (const double *)dmin.p, nt_pvts, pivots, n_itr);
}

int bcd_reduce_mtx_lncz_noalloc(bcd_ctx *ctx, int NA_out, int NB_out, int NPLS_out, double *Mout, int *i_piv_out,
                                int NA_in, int NB_in, int NPLS_in, const double *Min, int nt_pvts, double *pivots,
                                int *n_itr) {
  if (!ctx || !Mout || !i_piv_out || !Min || !pivots || !n_itr) return BCD_ERR_ARG;
  int nLanc = 0;
  const int rc = lanczos_checks(NA_out, NB_out, NPLS_out, NB_in, NPLS_in, nt_pvts, n_itr, &nLanc);
  if (rc != BCD_OK) return rc;
  const LancLayout L = lanczos_layout(true, NPLS_in, NA_in + NB_in, nt_pvts, NB_in, NA_out + NB_out, nLanc);
  if (L.total > kNoAllocScratchBytes) return BCD_ERR_CAPACITY;
  HTRY(hipSetDevice(ctx->device));
  char *base = nullptr;
  BTRY(noalloc_scratch(ctx->device, &base));
  return lanczos_in_region(ctx, base, L, nLanc, NA_out, NB_out, Mout, i_piv_out, NA_in, NB_in, NPLS_in, Min, nullptr,
                           nt_pvts, pivots, n_itr);
}

long bcd_noalloc_scratch_bytes(void) { return (long)kNoAllocScratchBytes; }

int bcd_reduce_selfen_lncz(bcd_ctx *ctx, int n_orb, int nfw, double *rows_fw, int nbk, double *rows_bk,
                           int nt_pvts, double *pivots, int *n_itr_fw, int *n_itr_bk, double *out_fw,
                           int *i_piv_fw, int *nfw_out, double *out_bk, int *i_piv_bk, int *nbk_out,
                           int *ntot_den) {
  if (!ctx || n_orb < 1 || nfw < 0 || nbk < 0 || !pivots || !n_itr_fw || !n_itr_bk || !nfw_out || !nbk_out ||
      !ntot_den || !i_piv_fw || !i_piv_bk)
    return BCD_ERR_ARG;
  HTRY(hipSetDevice(ctx->device));
  const int T = 2 + n_orb, nden = std::max(2, nt_pvts + 1), dim = nden + n_orb;
  for (int k = 0; k < dim; ++k) { i_piv_fw[k] = -100; i_piv_bk[k] = -100; }
  *ntot_den = nden;
  *nfw_out = *nbk_out = 0;
  // sort each side on the device (fw ascending, bk descending) with the
  // reference's heapsort, so rows of equal energy keep the reference's order
  // (Self_energy_ext_Op_e.cpp:119-130, Utilities.cpp:25-94)
  DevBuf dfw_in, dfw, dbk_in, dbk, dkey, didx;
  BTRY(dfw_in.alloc(sizeof(double) * (size_t)std::max(nfw, 1) * T));
  BTRY(dfw.alloc(sizeof(double) * (size_t)std::max(nfw, 1) * T));
  BTRY(dbk_in.alloc(sizeof(double) * (size_t)std::max(nbk, 1) * T));
  BTRY(dbk.alloc(sizeof(double) * (size_t)std::max(nbk, 1) * T));
  const int mx = std::max(nfw, nbk);
  if (mx > kHeapLdsRows) {
    BTRY(dkey.alloc(sizeof(double) * (size_t)mx));
    BTRY(didx.alloc(sizeof(int) * (size_t)mx));
  }
  auto heap = [&](const DevBuf &in, int M, int desc, const DevBuf &out) {
    const bool lds = M <= kHeapLdsRows;
    hipLaunchKernelGGL(k_heap_rows, dim3(1), dim3(256), lds ? (size_t)M * 12 + 16 : 0, ctx->stream,
                       (const double *)in.p, M, T, 0, desc, (double *)dkey.p, (int *)didx.p, lds ? 1 : 0,
                       (double *)out.p);
  };
  if (nfw) {
    HTRY(bcd::sync_memcpy(dfw_in.p, rows_fw, sizeof(double) * (size_t)nfw * T, hipMemcpyHostToDevice));
    heap(dfw_in, nfw, 0, dfw);
  }
  if (nbk) {
    HTRY(bcd::sync_memcpy(dbk_in.p, rows_bk, sizeof(double) * (size_t)nbk * T, hipMemcpyHostToDevice));
    heap(dbk_in, nbk, 1, dbk);
  }
  HTRY(hipGetLastError());
  HTRY(hipStreamSynchronize(ctx->stream));
  if (nfw) HTRY(bcd::sync_memcpy(rows_fw, dfw.p, sizeof(double) * (size_t)nfw * T, hipMemcpyDeviceToHost));
  if (nbk) HTRY(bcd::sync_memcpy(rows_bk, dbk.p, sizeof(double) * (size_t)nbk * T, hipMemcpyDeviceToHost));
  int tot[2] = {0, 0};
  for (int k = 0; k < nt_pvts; ++k) {
    if (n_itr_fw[k] > 0) tot[0] += n_itr_fw[k];
    if (n_itr_bk[k] > 0) tot[1] += n_itr_bk[k];
  }
  for (int side = 0; side < 2; ++side) {
    const int M = side ? nbk : nfw;
    int *itr = side ? n_itr_bk : n_itr_fw;
    const int nl = std::min(tot[side], M);
    if (M < 1 || nl < 1) continue;
    if (tot[side] > M) {
      const double x1 = (double)M / (double)tot[side];
      for (int k = 0; k < nt_pvts; ++k) itr[k] = (int)(x1 * itr[k] + 1.0);
    }
    const int r = run_lanczos(ctx, nden, n_orb, std::max(1, nl), side ? out_bk : out_fw, side ? i_piv_bk : i_piv_fw,
                              2, n_orb, M, (const double *)(side ? dbk.p : dfw.p), nt_pvts, pivots, itr);
    if (r < 0) return r;
    (side ? *nbk_out : *nfw_out) = r;
  }
  return BCD_OK;
}

int bcd_dyson(bcd_ctx *ctx, int n_orb, int j2, const double *h, const double *u1, int ntot_den_fw, int nfw,
              const double *rows_fw, const int *i_piv_fw, int npiv_fw, int ntot_den_bk, int nbk,
              const double *rows_bk, const int *i_piv_bk, int npiv_bk, double efermi, double *sol_qp, int *nqp,
              double *sol_qh, int *nqh, double *kolt3) {
  if (!ctx || n_orb < 1 || n_orb > kMaxNI || !h || !u1 || !sol_qp || !sol_qh || !nqp || !nqh) return BCD_ERR_ARG;
  nfw = std::max(nfw, 0);
  nbk = std::max(nbk, 0);
  if ((nfw && !rows_fw) || (nbk && !rows_bk)) return BCD_ERR_ARG;
  if ((npiv_fw > 0 && !i_piv_fw) || (npiv_bk > 0 && !i_piv_bk) || npiv_fw > kMaxPiv || npiv_bk > kMaxPiv)
    return BCD_ERR_ARG;
  HTRY(hipSetDevice(ctx->device));
  const int n = n_orb + nfw + nbk, Tf = ntot_den_fw + n_orb, Tb = ntot_den_bk + n_orb;
  const int ns[2] = {nfw, nbk}, Ts[2] = {Tf, Tb}, dens[2] = {ntot_den_fw, ntot_den_bk};
  const int npv[2] = {npiv_fw > 0 ? npiv_fw : 0, npiv_bk > 0 ? npiv_bk : 0};
  // a reduced side (pivots given) is a fishbone block and is diagonalised
  // first; an unreduced side is diagonal already
  const int lside = std::max(npv[0] ? nfw : 0, npv[1] ? nbk : 0);
  const size_t t_aux = ((size_t)(2 + n_orb) * lside + n_orb + 16) * sizeof(double);
  if (lside && t_aux > (size_t)kLdsBytes) return BCD_ERR_CAPACITY;
  const bool t_lds = t_aux + (size_t)lside * lside * sizeof(double) <= (size_t)kLdsBytes;
  const size_t aws = arrow_ws_bytes(n, n_orb);
  const bool a_lds = aws <= (size_t)kLdsBytes;
  DevBuf dh, du, dfw, dbk, dpf, dpb, dtri, dmat, dpole, dws, de, dx, dc, dk, dj, dja, dst;
  BTRY(dh.alloc(sizeof(double) * n_orb * n_orb));
  BTRY(du.alloc(sizeof(double) * n_orb * n_orb));
  BTRY(dfw.alloc(sizeof(double) * (size_t)std::max(nfw, 1) * Tf));
  BTRY(dbk.alloc(sizeof(double) * (size_t)std::max(nbk, 1) * Tb));
  BTRY(dpf.alloc(sizeof(int) * (kMaxPiv + 2)));
  BTRY(dpb.alloc(sizeof(int) * (kMaxPiv + 2)));
  BTRY(dtri.alloc(sizeof(double) * (size_t)std::max(lside, 1) * (2 + n_orb) * 2));
  BTRY(dpole.alloc(sizeof(double) * (size_t)(1 + n_orb) * std::max(nfw + nbk, 1)));
  if (lside && !t_lds) BTRY(dmat.alloc(sizeof(double) * (size_t)lside * lside * 2));
  if (!a_lds) BTRY(dws.alloc(aws));
  BTRY(de.alloc(sizeof(double) * n));
  BTRY(dx.alloc(sizeof(double) * (size_t)n * n_orb));
  BTRY(dc.alloc(sizeof(int) * 2));
  BTRY(dk.alloc(sizeof(double) * 3));
  BTRY(dj.alloc(sizeof(DysonJob) * 2));
  BTRY(dja.alloc(sizeof(ArrowJob)));
  BTRY(dst.alloc(sizeof(int)));
  HTRY(bcd::sync_memcpy(dh.p, h, sizeof(double) * n_orb * n_orb, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(du.p, u1, sizeof(double) * n_orb * n_orb, hipMemcpyHostToDevice));
  if (nfw) HTRY(bcd::sync_memcpy(dfw.p, rows_fw, sizeof(double) * (size_t)nfw * Tf, hipMemcpyHostToDevice));
  if (nbk) HTRY(bcd::sync_memcpy(dbk.p, rows_bk, sizeof(double) * (size_t)nbk * Tb, hipMemcpyHostToDevice));
  int pf[kMaxPiv + 2], pb[kMaxPiv + 2];
  for (int k = 0; k < kMaxPiv + 2; ++k) { pf[k] = -100; pb[k] = -100; }
  if (npiv_fw > 0) std::copy(i_piv_fw, i_piv_fw + npiv_fw + 2, pf);
  if (npiv_bk > 0) std::copy(i_piv_bk, i_piv_bk + npiv_bk + 2, pb);
  HTRY(bcd::sync_memcpy(dpf.p, pf, sizeof pf, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memcpy(dpb.p, pb, sizeof pb, hipMemcpyHostToDevice));
  HTRY(bcd::sync_memset(dst.p, 0, sizeof(int)));
  const double *rows[2] = {(const double *)dfw.p, (const double *)dbk.p};
  const int *ipv[2] = {(const int *)dpf.p, (const int *)dpb.p};
  double *pole = (double *)dpole.p;
  ArrowJob A;
  std::memset(&A, 0, sizeof A);
  DysonJob P[2];
  int njob = 0;
  long poff = 0;
  for (int s = 0; s < 2; ++s) {
    A.np[s] = ns[s];
    if (npv[s] && ns[s] > 0) {
      DysonJob &J = P[njob];
      std::memset(&J, 0, sizeof J);
      J.nI = n_orb; J.j2 = j2; J.n = ns[s]; J.inst = -1; J.poles = 1; J.path = 1;   // the general kernels below
      J.nfw = ns[s]; J.den_fw = J.den_bk = dens[s]; J.npiv_fw = npv[s];
      J.rfw = rows[s]; J.ipiv_fw = ipv[s];
      J.tri = (double *)dtri.p + (size_t)njob * lside * (2 + n_orb);
      J.mat = t_lds ? nullptr : (double *)dmat.p + (size_t)njob * lside * lside;
      J.out_lam = pole + poff; J.out_ct = pole + poff + ns[s]; J.ct_ld = ns[s];
      A.E[s] = J.out_lam; A.es[s] = 1; A.C[s] = J.out_ct; A.ci[s] = 1; A.cr[s] = ns[s];
      poff += (long)(1 + n_orb) * ns[s];
      ++njob;
    } else {
      A.E[s] = rows[s]; A.es[s] = Ts[s]; A.C[s] = rows[s] + dens[s]; A.ci[s] = Ts[s]; A.cr[s] = 1;
    }
  }
  A.nI = n_orb; A.j2 = j2; A.n = n; A.inst = -1;
  A.h = (const double *)dh.p; A.u1 = (const double *)du.p; A.efermi = efermi;
  A.ws = a_lds ? nullptr : (double *)dws.p;
  A.out_e = (double *)de.p; A.out_x = (double *)dx.p;
  A.out_nqp = (int *)dc.p; A.out_nqh = (int *)dc.p + 1; A.kolt = (double *)dk.p;
  HTRY(hipFuncSetAttribute((const void *)k_tridiag<true>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  HTRY(hipFuncSetAttribute((const void *)k_tridiag<false>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  HTRY(hipFuncSetAttribute((const void *)k_ql, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
  if (arrow_set_lds_limit(kLdsBytes)) return BCD_ERR_DEVICE;
  if (njob) {
    HTRY(bcd::sync_memcpy(dj.p, P, sizeof(DysonJob) * njob, hipMemcpyHostToDevice));
    const DysonJob *jp = (const DysonJob *)dj.p;
    if (t_lds)
      hipLaunchKernelGGL(k_tridiag<true>, dim3(njob), dim3(256), t_aux + (size_t)lside * lside * sizeof(double),
                         ctx->stream, jp, njob);
    else
      hipLaunchKernelGGL(k_tridiag<false>, dim3(njob), dim3(256), t_aux, ctx->stream, jp, njob);
    hipLaunchKernelGGL(k_ql, dim3(njob), dim3(64), (size_t)(2 + n_orb) * lside * sizeof(double), ctx->stream, jp,
                       njob, (int *)dst.p);
  }
  HTRY(hipMemcpyAsync(dja.p, &A, sizeof A, hipMemcpyHostToDevice, ctx->stream));
  const int nt = n <= 64 ? 64 : (n <= 128 ? 128 : kArrowMaxThreads);
  launch_arrow((const ArrowJob *)dja.p, 1, (int *)dst.p, n_orb, a_lds, nt, a_lds ? aws : 0, ctx->stream);
  HTRY(hipGetLastError());
  HTRY(hipStreamSynchronize(ctx->stream));
  int st = 0, cnt[2] = {0, 0};
  HTRY(bcd::sync_memcpy(&st, dst.p, sizeof(int), hipMemcpyDeviceToHost));
  if (st) return BCD_ERR_EIGEN;
  HTRY(bcd::sync_memcpy(cnt, dc.p, sizeof cnt, hipMemcpyDeviceToHost));
  std::vector<double> e(n), x((size_t)n * n_orb);
  HTRY(bcd::sync_memcpy(e.data(), de.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  HTRY(bcd::sync_memcpy(x.data(), dx.p, sizeof(double) * (size_t)n * n_orb, hipMemcpyDeviceToHost));
  if (kolt3) HTRY(bcd::sync_memcpy(kolt3, dk.p, sizeof(double) * 3, hipMemcpyDeviceToHost));
  const int w = n_orb + 2;
  for (int k = 0; k < n; ++k) {
    double *rec = k < cnt[0] ? sol_qp + (size_t)k * w : sol_qh + (size_t)(k - cnt[0]) * w;
    double z = 0.0;
    for (int r = 0; r < n_orb; ++r) { rec[2 + r] = x[(size_t)k * n_orb + r]; z += rec[2 + r] * rec[2 + r]; }
    rec[0] = e[k];
    rec[1] = z;
  }
  *nqp = cnt[0];
  *nqh = cnt[1];
  return BCD_OK;
}

int bcd_syevd_batched(bcd_ctx *ctx, int nmat, const int *n, const long *off, double *a, const long *woff,
                      double *w) {
  if (!ctx || nmat < 0 || (nmat && (!n || !off || !a || !woff || !w))) return BCD_ERR_ARG;
  HTRY(hipSetDevice(ctx->device));
  for (int k = 0; k < nmat; ++k) {
    const int m = n[k];
    if (m < 1) continue;
    // every row of the orthogonal factor is tracked (nI = m): the tracked rows
    // and work vectors sit in LDS, the matrix in LDS or HBM
    const size_t aux = ((size_t)(2 + m) * m + m + 16) * sizeof(double);
    if (aux > (size_t)kLdsBytes) return BCD_ERR_CAPACITY;
    const bool in_lds = aux + (size_t)m * m * sizeof(double) <= (size_t)kLdsBytes;
    DevBuf dh, du, dtri, dmat, de, dx, dc, dk, dj, dst;
    BTRY(dh.alloc(sizeof(double) * m * m));
    BTRY(du.alloc(sizeof(double) * m * m));
    BTRY(dtri.alloc(sizeof(double) * (size_t)m * (2 + m)));
    if (!in_lds) BTRY(dmat.alloc(sizeof(double) * (size_t)m * m));
    BTRY(de.alloc(sizeof(double) * m));
    BTRY(dx.alloc(sizeof(double) * (size_t)m * m));
    BTRY(dc.alloc(sizeof(int) * 2));
    BTRY(dk.alloc(sizeof(double) * 3));
    BTRY(dj.alloc(sizeof(DysonJob)));
    BTRY(dst.alloc(sizeof(int)));
    HTRY(bcd::sync_memcpy(dh.p, a + off[k], sizeof(double) * m * m, hipMemcpyHostToDevice));
    HTRY(bcd::sync_memset(du.p, 0, sizeof(double) * m * m));
    HTRY(bcd::sync_memset(dst.p, 0, sizeof(int)));
    DysonJob J;
    std::memset(&J, 0, sizeof J);
    J.nI = m; J.n = m; J.inst = -1; J.den_fw = J.den_bk = 2; J.efermi = HUGE_VAL;
    J.h = (const double *)dh.p; J.u1 = (const double *)du.p;
    J.tri = (double *)dtri.p; J.mat = (double *)dmat.p;
    J.out_e = (double *)de.p; J.out_x = (double *)dx.p;
    J.out_nqp = (int *)dc.p; J.out_nqh = (int *)dc.p + 1; J.kolt = (double *)dk.p;
    HTRY(bcd::sync_memcpy(dj.p, &J, sizeof J, hipMemcpyHostToDevice));
    const DysonJob *jp = (const DysonJob *)dj.p;
    HTRY(hipFuncSetAttribute((const void *)k_tridiag<true>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
    HTRY(hipFuncSetAttribute((const void *)k_tridiag<false>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
    HTRY(hipFuncSetAttribute((const void *)k_ql, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
    if (in_lds)
      hipLaunchKernelGGL(k_tridiag<true>, dim3(1), dim3(256), aux + (size_t)m * m * sizeof(double), ctx->stream, jp, 1);
    else
      hipLaunchKernelGGL(k_tridiag<false>, dim3(1), dim3(256), aux, ctx->stream, jp, 1);
    hipLaunchKernelGGL(k_ql, dim3(1), dim3(64), (size_t)(2 + m) * m * sizeof(double), ctx->stream, jp, 1,
                       (int *)dst.p);
    HTRY(hipGetLastError());
    HTRY(hipStreamSynchronize(ctx->stream));
    int st = 0;
    HTRY(bcd::sync_memcpy(&st, dst.p, sizeof(int), hipMemcpyDeviceToHost));
    if (st) return BCD_ERR_EIGEN;
    std::vector<double> e(m), x((size_t)m * m);
    HTRY(bcd::sync_memcpy(e.data(), de.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    HTRY(bcd::sync_memcpy(x.data(), dx.p, sizeof(double) * (size_t)m * m, hipMemcpyDeviceToHost));
    // all values sit below the infinite split point, so the records come out
    // by descending value: reverse to LAPACK's ascending order
    for (int i = 0; i < m; ++i) {
      w[woff[k] + i] = e[m - 1 - i];
      for (int r = 0; r < m; ++r) a[off[k] + (size_t)r * m + i] = x[(size_t)(m - 1 - i) * m + r];
    }
  }
  return BCD_OK;
}

// ------------------------------------------------------------- batched solve

int bcd_batch_create(int device, int n_inst, const bcd_modspace *ms, const bcd_vpp *vpp, const bcd_prop *g0,
                     const bcd_solve_params *par, bcd_batch **out) {
  return bcd_batch_create_ex(device, n_inst, ms, vpp, g0, par, nullptr, out);
}

int bcd_batch_create_ex(int device, int n_inst, const bcd_modspace *ms, const bcd_vpp *vpp, const bcd_prop *g0,
                        const bcd_solve_params *par, const int *lanc_n, bcd_batch **out) {
  if (!out || n_inst < 1 || !ms || !vpp || !par) return BCD_ERR_ARG;
  *out = nullptr;
  bcd_batch *B = nullptr;
  BTRY(bcd_batch_begin(device, n_inst, par, &B));
  std::unique_ptr<bcd_batch> hold(B);
  for (int i = 0; i < n_inst; ++i) {
    // the instances of one call keep their interaction in the batch's own
    // arena (one upload in bcd_batch_end), as small batches prefer
    std::unique_ptr<HostSystem> S(new (std::nothrow) HostSystem);
    if (!S) return BCD_ERR_NOMEM;
    BTRY(S->build(ms + i, vpp + i));
    if (g0) {
      if (!prop_ok(*S, g0 + i)) return BCD_ERR_ARG;
      B->g0[i] = to_host_prop(*S, g0 + i);
    } else {
      B->g0[i] = bare_prop(*S);
    }
    B->lanc_n[i] = lanc_n ? lanc_n[i] : 0;
    B->systems[i] = std::move(S);
  }
  BTRY(bcd_batch_end(B));
  *out = hold.release();
  return BCD_OK;
}

int bcd_batch_begin(int device, int n_inst, const bcd_solve_params *par, bcd_batch **out) {
  if (!out || n_inst < 1 || !par) return BCD_ERR_ARG;
  *out = nullptr;
  if (par->mode < 0 || par->mode > 2 || par->dyson_solver < 0 || par->dyson_solver > 1 || par->precision < 0 ||
      par->precision > 3 || par->accel < 0 || par->accel > kMixMax || par->hw < 0.0 ||
      (par->t_fac != 0.0 && par->hw <= 0.0))
    return BCD_ERR_ARG;
  BTRY(device_ok(device));
  std::unique_ptr<bcd_batch> B(new (std::nothrow) bcd_batch);
  if (!B) return BCD_ERR_NOMEM;
  B->device = device;
  B->par = *par;
  B->systems.resize(n_inst);
  B->g0.resize(n_inst);
  B->lanc_n.assign(n_inst, 0);
  B->vbufs.resize(n_inst);
  *out = B.release();
  return BCD_OK;
}

int bcd_batch_add(bcd_batch *b, int inst, const bcd_modspace *ms, const bcd_vpp *vpp, const bcd_prop *g0,
                  int lanc_n) {
  if (!b || b->ready || inst < 0 || inst >= (int)b->systems.size() || b->systems[inst] || !ms || !vpp)
    return BCD_ERR_ARG;
  std::unique_ptr<HostSystem> S(new (std::nothrow) HostSystem);
  if (!S) return BCD_ERR_NOMEM;
  BTRY(S->build(ms, vpp));
  if (g0 && !prop_ok(*S, g0)) return BCD_ERR_ARG;
  HostProp G = g0 ? to_host_prop(*S, g0) : bare_prop(*S);
  HTRY(hipSetDevice(b->device));
  DevArena &A = b->vbufs[inst];
  BTRY(A.grow(sizeof(double) * (size_t)std::max<long>(S->v_len, 1)));
  if (S->v_len) HTRY(bcd::sync_memcpy(A.ptr, S->V.data(), sizeof(double) * S->v_len, hipMemcpyHostToDevice));
  S->dV = (const double *)A.ptr;
  std::vector<double>().swap(S->V);
  b->g0[inst] = std::move(G);
  b->lanc_n[inst] = lanc_n;
  b->systems[inst] = std::move(S);
  return BCD_OK;
}

int bcd_batch_end(bcd_batch *b) {
  if (!b || b->ready) return BCD_ERR_ARG;
  std::vector<const HostSystem *> sp;
  for (auto &S : b->systems) {
    if (!S) return BCD_ERR_ARG;   // an instance was never added
    sp.push_back(S.get());
  }
  b->b.inst_lanc_n = b->lanc_n;
  BTRY(b->b.setup(b->device, sp, b->par, b->g0));
  std::vector<HostProp>().swap(b->g0);
  b->ready = true;
  return BCD_OK;
}

void bcd_batch_destroy(bcd_batch *b) { delete b; }

int bcd_batch_error_detail(const bcd_batch *b, char *buf, int n) {
  if (!b || !buf || n < 1) return BCD_ERR_ARG;
  const std::string &s = b->b.err_detail;
  const int k = std::min((int)s.size(), n - 1);
  std::memcpy(buf, s.data(), k);
  buf[k] = 0;
  return (int)s.size();
}

int bcd_batch_set_params(bcd_batch *b, const bcd_solve_params *par) {
  if (!b || !b->ready || !par) return BCD_ERR_ARG;
  return b->b.set_params(*par);
}

int bcd_batch_set_efermi(bcd_batch *b, int inst, int ish, double ef) {
  if (!b) return BCD_ERR_ARG;
  return b->b.set_efermi(inst, ish, ef);
}

int bcd_pivot_round(int device, const double *x, double *out, long n) {
  if (!x || !out || n < 0) return BCD_ERR_ARG;
  BTRY(device_ok(device));
  if (n == 0) return BCD_OK;
  HTRY(hipSetDevice(device));
  DevBuf d;
  BTRY(d.alloc(sizeof(double) * (size_t)n));
  HTRY(bcd::sync_memcpy(d.p, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_pivot_round, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, nullptr,
                     (const double *)d.p, (double *)d.p, n);
  HTRY(hipGetLastError());
  HTRY(bcd::sync_memcpy(out, d.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
  return BCD_OK;
}

int bcd_batch_solve(bcd_batch *b) {
  if (!b || !b->ready) return BCD_ERR_ARG;
  return b->b.solve();
}

int bcd_batch_set_profile(bcd_batch *b, int on) {
  if (!b) return BCD_ERR_ARG;
  b->b.profile = on != 0;
  return BCD_OK;
}

int bcd_batch_set_graphs(bcd_batch *b, int on) {
  if (!b) return BCD_ERR_ARG;
  b->b.use_graphs = on != 0;
  return BCD_OK;
}

int bcd_batch_timings(bcd_batch *b, double *ms, int n) {
  if (!b || !ms || n < 6) return BCD_ERR_ARG;
  for (int k = 0; k < 6; ++k) ms[k] = b->b.stage_ms[k];
  return BCD_OK;
}

int bcd_batch_stats(bcd_batch *b, int inst, int *hf_its, int *sc_its, double *last_diff, double *e_kolt,
                    long *n_fw_poles, long *n_bk_poles) {
  if (!b || inst < 0 || inst >= b->b.n_inst) return BCD_ERR_ARG;
  Batch &B = b->b;
  if (hf_its) *hf_its = B.hf_its.empty() ? 0 : B.hf_its[inst];
  if (sc_its) *sc_its = B.sc_its.empty() ? 0 : B.sc_its[inst];
  if (last_diff) *last_diff = B.h_is[inst].diff;
  if (e_kolt) *e_kolt = B.h_is[inst].e_kolt;
  if (n_fw_poles) *n_fw_poles = B.n_fw_tot.empty() ? 0 : B.n_fw_tot[inst];
  if (n_bk_poles) *n_bk_poles = B.n_bk_tot.empty() ? 0 : B.n_bk_tot[inst];
  return BCD_OK;
}

int bcd_batch_converged(bcd_batch *b, int inst) {
  if (!b || inst < 0 || inst >= b->b.n_inst || b->b.hf_ok.empty()) return BCD_ERR_ARG;
  return (b->b.hf_ok[inst] && b->b.sc_ok[inst]) ? 1 : 0;
}

int bcd_batch_result(bcd_batch *b, int inst, int nspb, int *np, int *nh, int *n_qp, int *n_qh, double *ep,
                     double *zp, double *fqp, double *eh, double *zh, double *fqh) {
  if (!b) return BCD_ERR_ARG;
  return b->b.result(inst, nspb, np, nh, n_qp, n_qh, ep, zp, fqp, eh, zh, fqh);
}

int bcd_batch_set_learned_start(bcd_batch *b, bcd_mlp *m) {
  if (!b || !b->ready) return BCD_ERR_ARG;
  if (m && (m->net.n_in != BCD_LEARN_IN || m->net.n_out != BCD_LEARN_OUT || m->net.device != b->device))
    return BCD_ERR_ARG;
  b->b.mlp = m ? &m->net : nullptr;
  return BCD_OK;
}

int bcd_batch_set_fallback_tol(bcd_batch *b, double tol) {
  if (!b || !(tol == tol)) return BCD_ERR_ARG;
  b->b.fallback_tol = tol;
  return BCD_OK;
}

int bcd_batch_record_learning(bcd_batch *b, int on) {
  if (!b) return BCD_ERR_ARG;
  b->b.learn_record = on != 0;
  return BCD_OK;
}

int bcd_batch_num_channels(const bcd_batch *b) { return (b && b->ready) ? b->b.n_gc : BCD_ERR_ARG; }

int bcd_batch_learning_data(bcd_batch *b, float *feat, float *target, int *chan_inst, int n_chan) {
  if (!b || !b->ready || n_chan != b->b.n_gc || !b->b.learn_valid) return BCD_ERR_ARG;
  Batch &B = b->b;
  HTRY(hipSetDevice(b->device));
  HTRY(hipStreamSynchronize(B.stream));
  if (feat) HTRY(bcd::sync_memcpy(feat, B.d_feat, sizeof(float) * (size_t)n_chan * BCD_LEARN_IN, hipMemcpyDeviceToHost));
  if (target)
    HTRY(bcd::sync_memcpy(target, B.d_target, sizeof(float) * (size_t)n_chan * BCD_LEARN_OUT, hipMemcpyDeviceToHost));
  if (chan_inst)
    for (int gc = 0; gc < n_chan; ++gc) chan_inst[gc] = B.h_chans[gc].inst;
  return BCD_OK;
}

int bcd_batch_start_info(bcd_batch *b, int inst, int *seeded, int *fell_back, int *first_its, double *residual) {
  if (!b || inst < 0 || inst >= b->b.n_inst || b->b.seeded.empty()) return BCD_ERR_ARG;
  Batch &B = b->b;
  if (seeded) *seeded = B.seeded[inst];
  if (fell_back) *fell_back = B.fell_back[inst];
  if (first_its) *first_its = B.first_its[inst];
  if (residual) *residual = B.resid[inst];
  return BCD_OK;
}

int bcd_batch_kernel_times(bcd_batch *b, int k, double *avg_ms, long *launches) {
  if (!b || k < 0 || k >= KT_N || !avg_ms || !launches) return BCD_ERR_ARG;
  *launches = b->b.kcount[k];
  *avg_ms = b->b.kcount[k] ? b->b.ktime[k] / b->b.kcount[k] : 0.0;
  return BCD_OK;
}

int bcd_batch_autotune(bcd_batch *b) {
  if (!b) return BCD_ERR_ARG;
  b->b.tune_lanczos = true;
  return BCD_OK;
}

int bcd_batch_lanczos_shapes(bcd_batch *b, int *shapes, float *tune_ms, int n) {
  if (!b || !shapes || n < kNLancBuckets) return BCD_ERR_ARG;
  for (int k = 0; k < kNLancBuckets; ++k) {
    shapes[k] = b->b.lanc_shape[k];
    if (tune_ms)
      for (int sh = 0; sh < kNLancShapes; ++sh) tune_ms[k * kNLancShapes + sh] = b->b.lanc_tune_ms[k][sh];
  }
  return BCD_OK;
}

int bcd_batch_set_lanczos_shapes(bcd_batch *b, const int *shapes, int n) {
  if (!b || !shapes || n < kNLancBuckets) return BCD_ERR_ARG;
  for (int k = 0; k < kNLancBuckets; ++k)
    if (!lanczos_shape_fits(shapes[k], kLancBucket[k])) return BCD_ERR_ARG;
  for (int k = 0; k < kNLancBuckets; ++k) b->b.lanc_shape[k] = shapes[k];
  return BCD_OK;
}

int bcd_batch_work(bcd_batch *b, double *work, int n) {
  if (!b || !work || n < 7) return BCD_ERR_ARG;
  return b->b.work(work, n >= 17 ? 17 : n >= 15 ? 15 : 7);
}

}  // extern "C"
