// Runs bcd::static_dyson -- the body of the k_dyson_static kernel -- on the
// host, so the tests can check it against a dense eigensolver on the CPU
// (built with AddressSanitizer and UndefinedBehaviorSanitizer).
//
// stdin : int32 ncase; per case int32 nI, j2; float64 efermi;
//         float64 h[nI*nI], u1[nI*nI]
// stdout: per case int32 rc, nqp, nqh; float64 e[nI], x[nI*nI], kolt[3]
#include <cstdint>
#include <cstdio>
#include <vector>

#include "bcd_arrow.hpp"

namespace {

template <class T>
bool rd(T *p, size_t n) { return std::fread(p, sizeof(T), n, stdin) == n; }

template <class T>
void wr(const T *p, size_t n) { std::fwrite(p, sizeof(T), n, stdout); }

}  // namespace

int main() {
  int32_t ncase = 0;
  if (!rd(&ncase, 1)) return 2;
  for (int c = 0; c < ncase; ++c) {
    int32_t hdr[2];
    double ef = 0.0;
    if (!rd(hdr, 2) || !rd(&ef, 1)) return 2;
    const int nI = hdr[0], j2 = hdr[1];
    const size_t nn = (size_t)(nI > 0 ? nI : 0) * (nI > 0 ? nI : 0);
    std::vector<double> h(nn + 1), u1(nn + 1), e(nI > 0 ? nI : 1), x(nn + 1), kolt(3);
    if (!rd(h.data(), nn) || !rd(u1.data(), nn)) return 2;
    int32_t out[3] = {0, 0, 0};
    int nqp = 0, nqh = 0;
    out[0] = bcd::static_dyson<bcd::kStaticNI>(nI, j2, h.data(), u1.data(), ef, e.data(), x.data(), &nqp, &nqh,
                                               kolt.data());
    out[1] = nqp;
    out[2] = nqh;
    wr(out, 3);
    wr(e.data(), nI > 0 ? nI : 0);
    wr(x.data(), nn);
    wr(kolt.data(), 3);
  }
  return 0;
}
