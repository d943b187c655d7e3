// bcd_order.hpp -- two small sequential routines the reference's results
// depend on bit for bit, written once for the kernels and the host tests.
//
//  * heap_perm: the row permutation of the reference's heapsorts
//    Sort_u_double2dim / Sort_d_double2dim (Utilities.cpp:25-94).  They move
//    whole rows, but which row goes where depends only on the key column, so
//    sorting (key, row index) pairs with the same sift order gives the same
//    permutation -- including the order of equal keys, which a stable sort
//    would not reproduce.  Poles of one (a, b, g) triple in different J
//    blocks have equal energies, so ties are common.
//  * dys_round: a pivot component as SpProp_t::DysPivots prints it ("%18.6le",
//    Sp_Propagator.cpp:1252) and Make_LancPivots reads it back with %lf
//    (Miscellanea.cpp:366-370): x rounded to 7 significant decimal digits,
//    half to even on exact ties, as the correctly rounded conversions do.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace bcd {

// Sort (key[i], idx[i]), i < n, ascending (desc = 0, Sort_u) or descending
// (desc = 1, Sort_d) with the reference's heapsort.  K and I are indexable
// (plain pointers into LDS, HBM or host memory).
template <class K, class I>
__host__ __device__ inline void heap_perm(K key, I idx, int n, int desc) {
  if (n < 2) return;
  int l = n / 2, ir = n - 1;
  while (true) {
    double xk;
    int xi;
    if (l > 0) {
      --l;
      xk = key[l]; xi = idx[l];
    } else {
      xk = key[ir]; xi = idx[ir];
      key[ir] = key[0]; idx[ir] = idx[0];
      if (--ir == 0) { key[0] = xk; idx[0] = xi; break; }
    }
    int i = l, j = l + l + 1;
    while (j <= ir) {
      if (j < ir && (desc ? key[j] > key[j + 1] : key[j] < key[j + 1])) ++j;
      if (desc ? xk > key[j] : xk < key[j]) {
        key[i] = key[j]; idx[i] = idx[j];
        i = j;
        j = j + j + 1;
      } else {
        break;
      }
    }
    key[i] = xk; idx[i] = xi;
  }
}

// 10^k, k = 0..22: every one of them is exact in binary64.
__host__ __device__ inline double pow10_exact(int k) {
  double p = 1.0;
  for (int i = 0; i < k; ++i) p *= 10.0;
  return p;
}

// x printed with 7 significant digits and read back.  With |x| in
// [1e-16, 1e29) the scale 10^k (or 10^-k) is exact, the integer mantissa N
// is chosen from the exact sign of x*10^k - N - 1/2 (one fused multiply-add
// each, no double rounding) and N / 10^k (N * 10^-k) is one correctly rounded
// operation, like the reader's.  Outside that range (pivot components are
// at most 1 in magnitude, so only components below 1e-16) the same steps
// run with an inexact scale.
__host__ __device__ inline double dys_round(double x) {
  if (x == 0.0 || !(x == x) || std::isinf(x)) return x;
  const double ax = std::fabs(x);
  // decimal exponent: 10^6 <= ax * 10^(6-e) < 10^7, correcting log10's
  // rounding.  Within an ulp of a power of ten either choice reads back as
  // that power, so the approximate test is enough.
  int e = (int)std::floor(std::log10(ax));
  const double t = ax * std::pow(10.0, 6 - e);
  if (t >= 1.0e7) ++e;
  else if (t < 1.0e6) --e;
  const int k = 6 - e;
  double N;
  if (k >= 0 && k <= 22) {
    const double P = pow10_exact(k);
    N = std::floor(ax * P);
    if (std::fma(ax, P, -N) < 0.0) N -= 1.0;          // the product was rounded up to N
    const double r = std::fma(ax, P, -(N + 0.5));
    if (r > 0.0 || (r == 0.0 && std::fmod(N, 2.0) != 0.0)) N += 1.0;
    return std::copysign(N / P, x);
  }
  if (k < 0 && -k <= 22) {
    const double Q = pow10_exact(-k);
    N = std::floor(ax / Q);
    if (std::fma(N, Q, -ax) > 0.0) N -= 1.0;
    const double r = std::fma(N + 0.5, Q, -ax);         // > 0: ax below the midpoint
    if (r < 0.0 || (r == 0.0 && std::fmod(N, 2.0) != 0.0)) N += 1.0;
    return std::copysign(N * Q, x);
  }
  const double P = std::pow(10.0, k);
  N = std::nearbyint(ax * P);
  return std::copysign(N / P, x);
}

}  // namespace bcd
