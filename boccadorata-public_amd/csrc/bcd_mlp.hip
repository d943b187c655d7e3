// The learned start's network on the GPU: BF16 MFMA GEMMs with fused bias /
// ReLU / ReLU-mask epilogues, the mean-squared loss, column sums for the
// bias gradients, Adam, and the C ABI (bcd_mlp_*).  Nothing here has a
// counterpart in the reference: the network predicts a starting density for
// its sc0 iteration (Main.cpp:216-451).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "bcd_hipsync.hpp"
#include "bcd_mlp.hpp"
#include "bcdor_amd.h"

namespace bcd {

namespace {

#define MTRY(x)                                  \
  do {                                           \
    hipError_t e_ = (x);                         \
    if (e_ != hipSuccess) return BCD_ERR_DEVICE; \
  } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __host__ inline unsigned short to_bf16(float f) {
  unsigned int u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

constexpr int kTile = 64;      // block tile (rows and columns)
constexpr int kKs = 16;        // K step: one 16x16x16 MFMA
constexpr int kApad = 20;      // LDS row pitch of the A tile (bf16 elements; 8-byte aligned rows)
constexpr int kBpad = 68;      // LDS row pitch of the B tile

// C[m][n] = epi( sum_k A(m,k) B(k,n) ), A(m,k) = A[m*sam + k*sak] (FP32,
// rounded to BF16 on the way into LDS), B(k,n) = Bf[k*sbk + n*sbn] (FP32) or
// Bh[...] (BF16).  Epilogue: + bias[n] (if bias), * (mask[m*ldm+n] > 0) (if
// mask), max(., 0) (if RELU).  Four waves per block, each a 32x32 quarter of
// the 64x64 tile as 2x2 MFMA tiles.  MFMA operand layout (16x16x16 bf16):
// lane l holds A(l%16, 4*(l/16)+0..3) and B(4*(l/16)+0..3, l%16); the result
// element r of lane l is C(4*(l/16)+r, l%16).
template <bool RELU>
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const float *A, long sam, long sak,
                                                  const float *Bf, const unsigned short *Bh, long sbk, long sbn,
                                                  const float *bias, const float *mask, long ldm, float *C,
                                                  long ldc) {
  __shared__ unsigned short As[kTile * kApad];
  __shared__ unsigned short Bs[kKs * kBpad];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, wm = wv >> 1, wn = wv & 1;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kKs) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + 256 * r;
      const int am = idx / kKs, ak = idx % kKs;
      const int gm = m0 + am, gk = k0 + ak;
      As[am * kApad + ak] = (gm < M && gk < K) ? to_bf16(A[(long)gm * sam + (long)gk * sak]) : (unsigned short)0;
      const int bk = idx / kTile, bn = idx % kTile;
      const int hk = k0 + bk, hn = n0 + bn;
      unsigned short bv = 0;
      if (hk < K && hn < N) bv = Bh ? Bh[(long)hk * sbk + (long)hn * sbn] : to_bf16(Bf[(long)hk * sbk + (long)hn * sbn]);
      Bs[bk * kBpad + bn] = bv;
    }
    __syncthreads();
    bf16x4 a[2], b[2];
    const int kq = 4 * (lane >> 4), l16 = lane & 15;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const unsigned short *src = As + (wm * 32 + i * 16 + l16) * kApad + kq;
      a[i] = bf16x4{(short)src[0], (short)src[1], (short)src[2], (short)src[3]};
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = wn * 32 + j * 16 + l16;
      b[j] = bf16x4{(short)Bs[(kq + 0) * kBpad + col], (short)Bs[(kq + 1) * kBpad + col],
                    (short)Bs[(kq + 2) * kBpad + col], (short)Bs[(kq + 3) * kBpad + col]};
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  const int rq = 4 * (lane >> 4), l16 = lane & 15;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + l16;
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + rq + r;
        if (row >= M) continue;
        float v = acc[i][j][r];
        if (bias) v += bias[col];
        if (mask && !(mask[(long)row * ldm + col] > 0.f)) v = 0.f;
        if (RELU) v = v > 0.f ? v : 0.f;
        C[(long)row * ldc + col] = v;
      }
    }
}

// dy = 2 (y - t) / (n * nout); part[block] = sum of (y - t)^2 over the block's elements.
__global__ __launch_bounds__(256) void k_mlp_loss(long total, const float *y, const float *t, float scale, float *dy,
                                                  float *part) {
  __shared__ float red[256];
  float s = 0.f;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += 256L * gridDim.x) {
    const float d = y[i] - t[i];
    dy[i] = 2.f * d * scale;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// loss = scale * sum of the nb partials (one thread, fixed order).
__global__ void k_mlp_loss_total(const float *part, int nb, float scale, float *loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[i];
    *loss = (float)(s * scale);
  }
}

// out[c] = sum over the n rows of a[r*ld + c], rows in order.
__global__ __launch_bounds__(256) void k_mlp_colsum(int n, int cols, const float *a, long ld, float *out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < n; ++r) s += a[(long)r * ld + c];
  out[c] = s;
}

__global__ __launch_bounds__(256) void k_mlp_adam(long n, float *p, const float *g, float *m1, float *m2, float lr,
                                                  float b1, float b2, float c1, float c2, float eps) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_to_bf16(long n, const float *p, unsigned short *w) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) w[i] = to_bf16(p[i]);
}

// Rows perm[0..n) of X (width wx) and T (width wt) into xb, tb.
__global__ __launch_bounds__(256) void k_mlp_gather(int n, const int *perm, const float *X, int wx, const float *T,
                                                    int wt, float *xb, float *tb) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  const int w = wx + wt;
  if (i >= (long)n * w) return;
  const int r = (int)(i / w), c = (int)(i % w);
  const long src = perm[r];
  if (c < wx) xb[(long)r * wx + c] = X[src * wx + c];
  else tb[(long)r * wt + c - wx] = T[src * wt + c - wx];
}

template <bool RELU>
void gemm(hipStream_t s, int M, int N, int K, const float *A, long sam, long sak, const float *Bf,
          const unsigned short *Bh, long sbk, long sbn, const float *bias, const float *mask, long ldm, float *C,
          long ldc) {
  const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_mlp_gemm<RELU>, grid, dim3(256), 0, s, M, N, K, A, sam, sak, Bf, Bh, sbk, sbn, bias, mask,
                     ldm, C, ldc);
}

}  // namespace

Mlp::~Mlp() {
  if (device >= 0) (void)hipSetDevice(device);
  for (void *q : {(void *)p, (void *)g, (void *)m1, (void *)m2, (void *)wb, (void *)x, (void *)h1, (void *)h2,
                  (void *)y, (void *)t, (void *)dy, (void *)dh2, (void *)dh1, (void *)part, (void *)perm})
    if (q) (void)bcd::sync_free(q);
  if (stream) (void)hipStreamDestroy(stream);
}

int Mlp::init(int dev, int nin, int nhid, int nout, unsigned long long seed) {
  device = dev;
  n_in = nin; n_hid = nhid; n_out = nout;
  const long sz[6] = {(long)nhid * nin, nhid, (long)nhid * nhid, nhid, (long)nout * nhid, nout};
  long o = 0;
  for (int l = 0; l < 3; ++l) {
    off_w[l] = o; o += sz[2 * l];
    off_b[l] = o; o += sz[2 * l + 1];
  }
  n_par = o;
  MTRY(hipSetDevice(dev));
  MTRY(hipStreamCreate(&stream));
  MTRY(bcd::sync_malloc(&p, sizeof(float) * n_par));
  MTRY(bcd::sync_malloc(&g, sizeof(float) * n_par));
  MTRY(bcd::sync_malloc(&m1, sizeof(float) * n_par));
  MTRY(bcd::sync_malloc(&m2, sizeof(float) * n_par));
  MTRY(bcd::sync_malloc(&wb, sizeof(unsigned short) * n_par));
  // He initialisation of the weights, zero biases
  std::vector<float> h(n_par, 0.f);
  std::mt19937_64 rng(seed);
  const int fan[3] = {nin, nhid, nhid};
  for (int l = 0; l < 3; ++l) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (float)fan[l]) * (l == 2 ? 0.1f : 1.f));
    for (long i = 0; i < sz[2 * l]; ++i) h[off_w[l] + i] = nd(rng);
  }
  MTRY(bcd::sync_memcpy(p, h.data(), sizeof(float) * n_par, hipMemcpyHostToDevice));
  MTRY(bcd::sync_memset(m1, 0, sizeof(float) * n_par));
  MTRY(bcd::sync_memset(m2, 0, sizeof(float) * n_par));
  MTRY(bcd::sync_memset(g, 0, sizeof(float) * n_par));
  adam_t = 0;
  int r = refresh_bf16(stream);
  if (r) return r;
  MTRY(hipStreamSynchronize(stream));
  return BCD_OK;
}

int Mlp::reserve(int rows) {
  if (rows <= cap) return BCD_OK;
  MTRY(hipSetDevice(device));
  MTRY(hipStreamSynchronize(stream));
  for (float **q : {&x, &h1, &h2, &y, &t, &dy, &dh2, &dh1, &part})
    if (*q) { MTRY(bcd::sync_free(*q)); *q = nullptr; }
  if (perm) { MTRY(bcd::sync_free(perm)); perm = nullptr; }
  const long r = rows;
  MTRY(bcd::sync_malloc(&x, sizeof(float) * r * n_in));
  MTRY(bcd::sync_malloc(&h1, sizeof(float) * r * n_hid));
  MTRY(bcd::sync_malloc(&h2, sizeof(float) * r * n_hid));
  MTRY(bcd::sync_malloc(&y, sizeof(float) * r * n_out));
  MTRY(bcd::sync_malloc(&t, sizeof(float) * r * n_out));
  MTRY(bcd::sync_malloc(&dy, sizeof(float) * r * n_out));
  MTRY(bcd::sync_malloc(&dh2, sizeof(float) * r * n_hid));
  MTRY(bcd::sync_malloc(&dh1, sizeof(float) * r * n_hid));
  MTRY(bcd::sync_malloc(&part, sizeof(float) * 1056));   // loss partials [0, 1024), the loss at 1024
  MTRY(bcd::sync_malloc(&perm, sizeof(int) * r));
  cap = rows;
  return BCD_OK;
}

int Mlp::refresh_bf16(hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3((unsigned)((n_par + 255) / 256)), dim3(256), 0, s, n_par, p, wb);
  MTRY(hipGetLastError());
  return BCD_OK;
}

int Mlp::forward_dev(int n, const float *xd, float *yd, hipStream_t s) {
  if (n > cap) return BCD_ERR_CAPACITY;
  // h1 = relu(x W1^T + b1), h2 = relu(h1 W2^T + b2), y = h2 W3^T + b3
  gemm<true>(s, n, n_hid, n_in, xd, n_in, 1, nullptr, wb + off_w[0], 1, n_in, p + off_b[0], nullptr, 0, h1, n_hid);
  gemm<true>(s, n, n_hid, n_hid, h1, n_hid, 1, nullptr, wb + off_w[1], 1, n_hid, p + off_b[1], nullptr, 0, h2, n_hid);
  gemm<false>(s, n, n_out, n_hid, h2, n_hid, 1, nullptr, wb + off_w[2], 1, n_hid, p + off_b[2], nullptr, 0, yd,
              n_out);
  MTRY(hipGetLastError());
  return BCD_OK;
}

int Mlp::loss_grad_dev(int n, const float *xd, const float *td, float *loss_dev, hipStream_t s) {
  int r = forward_dev(n, xd, y, s);
  if (r) return r;
  const long tot = (long)n * n_out;
  const int nb = (int)std::min<long>(1024, (tot + 255) / 256);
  const float scale = 1.f / (float)tot;
  hipLaunchKernelGGL(k_mlp_loss, dim3(nb), dim3(256), 0, s, tot, y, td, scale, dy, part);
  if (loss_dev) hipLaunchKernelGGL(k_mlp_loss_total, dim3(1), dim3(64), 0, s, part, nb, scale, loss_dev);
  float *gw3 = g + off_w[2], *gb3 = g + off_b[2], *gw2 = g + off_w[1], *gb2 = g + off_b[1];
  float *gw1 = g + off_w[0], *gb1 = g + off_b[0];
  // layer 3: dW3 = dy^T h2, db3 = colsum dy, dh2 = (dy W3) * [h2 > 0]
  gemm<false>(s, n_out, n_hid, n, dy, 1, n_out, h2, nullptr, n_hid, 1, nullptr, nullptr, 0, gw3, n_hid);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((n_out + 255) / 256), dim3(256), 0, s, n, n_out, dy, (long)n_out, gb3);
  gemm<false>(s, n, n_hid, n_out, dy, n_out, 1, nullptr, wb + off_w[2], n_hid, 1, nullptr, h2, n_hid, dh2, n_hid);
  // layer 2
  gemm<false>(s, n_hid, n_hid, n, dh2, 1, n_hid, h1, nullptr, n_hid, 1, nullptr, nullptr, 0, gw2, n_hid);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((n_hid + 255) / 256), dim3(256), 0, s, n, n_hid, dh2, (long)n_hid, gb2);
  gemm<false>(s, n, n_hid, n_hid, dh2, n_hid, 1, nullptr, wb + off_w[1], n_hid, 1, nullptr, h1, n_hid, dh1, n_hid);
  // layer 1
  gemm<false>(s, n_hid, n_in, n, dh1, 1, n_hid, xd, nullptr, n_in, 1, nullptr, nullptr, 0, gw1, n_in);
  hipLaunchKernelGGL(k_mlp_colsum, dim3((n_hid + 255) / 256), dim3(256), 0, s, n, n_hid, dh1, (long)n_hid, gb1);
  MTRY(hipGetLastError());
  return BCD_OK;
}

int Mlp::adam(float lr, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  ++adam_t;
  const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((n_par + 255) / 256)), dim3(256), 0, s, n_par, p, g, m1, m2, lr, b1,
                     b2, c1, c2, 1e-8f);
  MTRY(hipGetLastError());
  return refresh_bf16(s);
}

}  // namespace bcd

namespace {

#define CTRY(x)                                  \
  do {                                           \
    hipError_t e_ = (x);                         \
    if (e_ != hipSuccess) return BCD_ERR_DEVICE; \
  } while (0)

#define RTRY(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != BCD_OK) return r_; \
  } while (0)

constexpr int kMlpChunk = 8192;   // rows per forward chunk of the host entry points

}  // namespace

extern "C" {

int bcd_mlp_create(int device, int n_in, int n_hidden, int n_out, unsigned long long seed, bcd_mlp **out) {
  if (!out || n_in < 1 || n_hidden < 1 || n_out < 1 || n_in > 4096 || n_hidden > 4096 || n_out > 4096)
    return BCD_ERR_ARG;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return BCD_ERR_DEVICE;
  bcd_mlp *m = new (std::nothrow) bcd_mlp;
  if (!m) return BCD_ERR_NOMEM;
  const int r = m->net.init(device, n_in, n_hidden, n_out, seed);
  if (r) { delete m; return r; }
  *out = m;
  return BCD_OK;
}

void bcd_mlp_destroy(bcd_mlp *m) { delete m; }

long bcd_mlp_num_params(const bcd_mlp *m) { return m ? m->net.n_par : 0; }

int bcd_mlp_get_params(bcd_mlp *m, float *p) {
  if (!m || !p) return BCD_ERR_ARG;
  CTRY(hipSetDevice(m->net.device));
  CTRY(hipStreamSynchronize(m->net.stream));
  CTRY(bcd::sync_memcpy(p, m->net.p, sizeof(float) * m->net.n_par, hipMemcpyDeviceToHost));
  return BCD_OK;
}

int bcd_mlp_set_params(bcd_mlp *m, const float *p) {
  if (!m || !p) return BCD_ERR_ARG;
  CTRY(hipSetDevice(m->net.device));
  CTRY(hipStreamSynchronize(m->net.stream));
  CTRY(bcd::sync_memcpy(m->net.p, p, sizeof(float) * m->net.n_par, hipMemcpyHostToDevice));
  RTRY(m->net.refresh_bf16(m->net.stream));
  CTRY(hipStreamSynchronize(m->net.stream));
  return BCD_OK;
}

int bcd_mlp_forward(bcd_mlp *m, int n, const float *x, float *y) {
  if (!m || n < 0 || (n && (!x || !y))) return BCD_ERR_ARG;
  bcd::Mlp &N = m->net;
  CTRY(hipSetDevice(N.device));
  RTRY(N.reserve(std::min(n, kMlpChunk)));
  for (int r0 = 0; r0 < n; r0 += kMlpChunk) {
    const int k = std::min(kMlpChunk, n - r0);
    CTRY(hipMemcpyAsync(N.x, x + (long)r0 * N.n_in, sizeof(float) * k * N.n_in, hipMemcpyHostToDevice, N.stream));
    RTRY(N.forward_dev(k, N.x, N.y, N.stream));
    CTRY(hipMemcpyAsync(y + (long)r0 * N.n_out, N.y, sizeof(float) * k * N.n_out, hipMemcpyDeviceToHost, N.stream));
  }
  CTRY(hipStreamSynchronize(N.stream));
  return BCD_OK;
}

int bcd_mlp_loss_grad(bcd_mlp *m, int n, const float *x, const float *t, float *loss, float *grad) {
  if (!m || n < 1 || !x || !t) return BCD_ERR_ARG;
  bcd::Mlp &N = m->net;
  CTRY(hipSetDevice(N.device));
  RTRY(N.reserve(n));
  CTRY(hipMemcpyAsync(N.x, x, sizeof(float) * n * N.n_in, hipMemcpyHostToDevice, N.stream));
  CTRY(hipMemcpyAsync(N.t, t, sizeof(float) * n * N.n_out, hipMemcpyHostToDevice, N.stream));
  RTRY(N.loss_grad_dev(n, N.x, N.t, N.part + 1024, N.stream));
  if (loss) CTRY(hipMemcpyAsync(loss, N.part + 1024, sizeof(float), hipMemcpyDeviceToHost, N.stream));
  if (grad) CTRY(hipMemcpyAsync(grad, N.g, sizeof(float) * N.n_par, hipMemcpyDeviceToHost, N.stream));
  CTRY(hipStreamSynchronize(N.stream));
  return BCD_OK;
}

int bcd_mlp_fit(bcd_mlp *m, int n, const float *x, const float *t, int epochs, int batch, float lr,
                unsigned long long seed, float *epoch_loss) {
  if (!m || n < 1 || !x || !t || epochs < 0 || batch < 1 || !(lr > 0.f)) return BCD_ERR_ARG;
  bcd::Mlp &N = m->net;
  CTRY(hipSetDevice(N.device));
  batch = std::min(batch, n);
  RTRY(N.reserve(batch));
  // the training set resident on the device for the whole fit
  float *X = nullptr, *T = nullptr, *L = nullptr;
  int *P = nullptr;
  const int nbat = (n + batch - 1) / batch;
  auto cleanup = [&]() {
    if (X) (void)bcd::sync_free(X);
    if (T) (void)bcd::sync_free(T);
    if (L) (void)bcd::sync_free(L);
    if (P) (void)bcd::sync_free(P);
  };
  if (bcd::sync_malloc(&X, sizeof(float) * (long)n * N.n_in) != hipSuccess ||
      bcd::sync_malloc(&T, sizeof(float) * (long)n * N.n_out) != hipSuccess ||
      bcd::sync_malloc(&L, sizeof(float) * std::max(1, epochs) * (long)nbat) != hipSuccess ||
      bcd::sync_malloc(&P, sizeof(int) * (long)n) != hipSuccess) {
    cleanup();
    return BCD_ERR_NOMEM;
  }
  int rc = BCD_OK;
  std::vector<int> perm(n);
  std::mt19937_64 rng(seed);
  do {
    if (bcd::sync_memcpy(X, x, sizeof(float) * (long)n * N.n_in, hipMemcpyHostToDevice) != hipSuccess ||
        bcd::sync_memcpy(T, t, sizeof(float) * (long)n * N.n_out, hipMemcpyHostToDevice) != hipSuccess) {
      rc = BCD_ERR_DEVICE;
      break;
    }
    for (int e = 0; e < epochs && rc == BCD_OK; ++e) {
      for (int i = 0; i < n; ++i) perm[i] = i;
      std::shuffle(perm.begin(), perm.end(), rng);
      // the previous epoch's gathers have finished reading P
      if (hipStreamSynchronize(N.stream) != hipSuccess ||
          bcd::sync_memcpy(P, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = BCD_ERR_DEVICE;
        break;
      }
      for (int bi = 0; bi < nbat && rc == BCD_OK; ++bi) {
        const int r0 = bi * batch, k = std::min(batch, n - r0);
        const long tot = (long)k * (N.n_in + N.n_out);
        hipLaunchKernelGGL(bcd::k_mlp_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, N.stream, k,
                           P + r0, X, N.n_in, T, N.n_out, N.x, N.t);
        rc = N.loss_grad_dev(k, N.x, N.t, L + (long)e * nbat + bi, N.stream);
        if (rc == BCD_OK) rc = N.adam(lr, N.stream);
      }
    }
    if (rc != BCD_OK) break;
    if (hipStreamSynchronize(N.stream) != hipSuccess) { rc = BCD_ERR_DEVICE; break; }
    if (epoch_loss && epochs > 0) {
      std::vector<float> h((size_t)epochs * nbat);
      if (bcd::sync_memcpy(h.data(), L, sizeof(float) * h.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = BCD_ERR_DEVICE;
        break;
      }
      for (int e = 0; e < epochs; ++e) {
        double s = 0.0;
        for (int bi = 0; bi < nbat; ++bi) s += h[(size_t)e * nbat + bi];
        epoch_loss[e] = (float)(s / nbat);
      }
    }
  } while (false);
  (void)hipStreamSynchronize(N.stream);
  cleanup();
  return rc;
}

}  // extern "C"
