// The learned start's network: a 3-layer perceptron (n_in -> n_hid -> n_hid
// -> n_out, ReLU between layers) whose GEMMs run on the bf16 matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation).  Master
// weights, biases, gradients and the Adam moments are FP32 in HBM; the
// weights the GEMMs read are BF16 copies refreshed after every update, and
// activations are rounded to BF16 where they enter a GEMM.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace bcd {

struct Mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_in = 0, n_hid = 0, n_out = 0;
  long n_par = 0;
  // parameters in one FP32 array: W1 [hid x in], b1, W2 [hid x hid], b2, W3 [out x hid], b3
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  unsigned short *wb = nullptr;   // BF16 copies of W1, W2, W3 at the same offsets as in p
  long off_w[3] = {0, 0, 0}, off_b[3] = {0, 0, 0};
  long adam_t = 0;
  // activations of up to cap rows: x, h1, h2, y, t, dy, dh2, dh1, plus the
  // loss partials and a row permutation
  int cap = 0;
  float *x = nullptr, *h1 = nullptr, *h2 = nullptr, *y = nullptr, *t = nullptr;
  float *dy = nullptr, *dh2 = nullptr, *dh1 = nullptr, *part = nullptr;
  int *perm = nullptr;

  ~Mlp();
  int init(int dev, int nin, int nhid, int nout, unsigned long long seed);
  int reserve(int rows);
  // y = net(x) for n rows of device arrays (n <= cap after reserve(n))
  int forward_dev(int n, const float *xd, float *yd, hipStream_t s);
  // loss = mean (y - t)^2 over n rows and n_out outputs; gradient into g
  int loss_grad_dev(int n, const float *xd, const float *td, float *loss_dev, hipStream_t s);
  int adam(float lr, hipStream_t s);
  int refresh_bf16(hipStream_t s);
};

}  // namespace bcd

// The C ABI's handle (bcd_mlp_*, include/bcdor_amd.h).
struct bcd_mlp {
  bcd::Mlp net;
};
