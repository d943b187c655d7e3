// Runs the sequential routines of bcd_order.hpp (the code the kernels run)
// on the host, built with AddressSanitizer and UndefinedBehaviorSanitizer.
//
// stdin : int32 mode, int32 n, then
//         mode 0: float64 x[n]              -> stdout float64 dys_round(x[i])
//         mode 1: int32 desc, float64 key[n] -> stdout int32 permutation[n]
#include <cstdint>
#include <cstdio>
#include <vector>

#include "bcd_order.hpp"

int main() {
  int32_t mode = 0, n = 0;
  if (fread(&mode, 4, 1, stdin) != 1 || fread(&n, 4, 1, stdin) != 1 || n < 0) return 2;
  if (mode == 0) {
    std::vector<double> x(n);
    if (fread(x.data(), 8, n, stdin) != (size_t)n) return 2;
    for (double &v : x) v = bcd::dys_round(v);
    fwrite(x.data(), 8, n, stdout);
    return 0;
  }
  int32_t desc = 0;
  if (fread(&desc, 4, 1, stdin) != 1) return 2;
  std::vector<double> key(n);
  std::vector<int> idx(n);
  if (fread(key.data(), 8, n, stdin) != (size_t)n) return 2;
  for (int i = 0; i < n; ++i) idx[i] = i;
  bcd::heap_perm(key.data(), idx.data(), n, desc);
  fwrite(idx.data(), 4, n, stdout);
  return 0;
}
