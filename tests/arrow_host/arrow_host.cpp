// Runs bcd::arrow_body -- the code of the k_arrow kernel -- on host threads:
// one std::thread per GPU thread of the block and a barrier for each
// __syncthreads, so the tests can check the algorithm and every index on the
// CPU (built with AddressSanitizer).
//
// stdin : int32 ncase; per case int32 nI, np0, np1, nt, j2, warm; float64 efermi;
//         float64 h[nI*nI], u1[nI*nI], E0[np0], C0[np0*nI], E1[np1], C1[np1*nI]
//         (C rows = poles, columns = orbits); with warm = 1 also h2[nI*nI]
// stdout: per case int32 nqp, nqh, failed; float64 e[n], x[n*nI], kolt[3], flops;
//         with warm = 1 a second such record: the same poles with static part
//         h2, started from the first run's roots (as the next sweep is)
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "bcd_arrow.hpp"

namespace {

struct Barrier {
  std::mutex mu;
  std::condition_variable cv;
  int n, waiting = 0;
  long gen = 0;
  explicit Barrier(int n_) : n(n_) {}
  void wait() {
    std::unique_lock<std::mutex> lk(mu);
    const long g = gen;
    if (++waiting == n) { waiting = 0; ++gen; cv.notify_all(); return; }
    cv.wait(lk, [&] { return gen != g; });
  }
};

struct HostBlock {
  int t, n;
  Barrier *bar;
  int *failed;
  double *acc;
  int tid() const { return t; }
  int nt() const { return n; }
  void sync() const { bar->wait(); }
  void fail() const { *failed = 1; }
  double *work;
  void add_work(double f) const { work[t] += f; }
  double rcp(double x) const { return 1.0 / x; }
  unsigned long long lap(int, unsigned long long) const { return 0; }
  int scan01(int *, int v, int &tot) const {
    acc[t] = v;
    sync();
    int pre = 0;
    tot = 0;
    for (int i = 0; i < n; ++i) { pre += i < t ? (int)acc[i] : 0; tot += (int)acc[i]; }
    sync();
    return pre;
  }
  double sum(double *, double v) const { return reduce(v, false); }
  template <int N>
  void sums(double *, double *v) const {
    for (int k = 0; k < N; ++k) v[k] = reduce(v[k], false);
  }
  double max(double *, double v) const { return reduce(v, true); }
  double reduce(double v, bool mx) const {
    acc[t] = v;
    sync();
    double s = mx ? acc[0] : 0.0;
    for (int i = 0; i < n; ++i) s = mx ? (acc[i] > s ? acc[i] : s) : s + acc[i];
    sync();
    return s;
  }
};

template <class T>
bool get(std::vector<T> &v, size_t n) {
  v.resize(n);
  return n == 0 || fread(v.data(), sizeof(T), n, stdin) == n;
}

}  // namespace

int main() {
  int32_t ncase = 0;
  if (fread(&ncase, 4, 1, stdin) != 1) return 2;
  for (int c = 0; c < ncase; ++c) {
    int32_t hd[6];
    double ef;
    if (fread(hd, 4, 6, stdin) != 6 || fread(&ef, 8, 1, stdin) != 1) return 2;
    const int nI = hd[0], np0 = hd[1], np1 = hd[2], nt = hd[3], j2 = hd[4], warm = hd[5], n = nI + np0 + np1;
    if (nI < 1 || nI > bcd::kArrowNI || nt < 1 || nt > bcd::kArrowMaxThreads) return 3;
    std::vector<double> h, u1, e0, c0, e1, c1, h2;
    if (!get(h, (size_t)nI * nI) || !get(u1, (size_t)nI * nI) || !get(e0, np0) || !get(c0, (size_t)np0 * nI) ||
        !get(e1, np1) || !get(c1, (size_t)np1 * nI) || (warm && !get(h2, (size_t)nI * nI)))
      return 2;
    std::vector<double> roots((size_t)nI * n, 0.0);
    std::vector<int> counts(nI, -1);
    for (int run = 0; run < (warm ? 2 : 1); ++run) {
      std::vector<double> oe(n), ox((size_t)n * nI), kolt(3), work(nt, 0.0);
      std::vector<char> ws(bcd::arrow_ws_bytes(n, nI));
      int nqp = -1, nqh = -1, failed = 0;
      bcd::ArrowJob J{};
      J.nI = nI; J.j2 = j2; J.n = n; J.inst = -1;
      J.np[0] = np0; J.np[1] = np1;
      J.E[0] = e0.data(); J.E[1] = e1.data(); J.C[0] = c0.data(); J.C[1] = c1.data();
      J.es[0] = J.es[1] = 1; J.ci[0] = J.ci[1] = nI; J.cr[0] = J.cr[1] = 1;
      J.h = run ? h2.data() : h.data(); J.u1 = u1.data(); J.efermi = ef;
      J.out_e = oe.data(); J.out_x = ox.data(); J.out_nqp = &nqp; J.out_nqh = &nqh; J.kolt = kolt.data();
      if (warm) { J.warm = roots.data(); J.warm_p = counts.data(); J.wld = n; }
      Barrier bar(nt);
      std::vector<double> acc(nt);
      std::vector<std::thread> th;
      for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
          bcd::arrow_body<bcd::kArrowNI>(HostBlock{t, nt, &bar, &failed, acc.data(), work.data()}, J,
                                         (double *)ws.data());
        });
      for (auto &x : th) x.join();
      double flops = 0.0;
      for (double f : work) flops += f;
      int32_t o[3] = {nqp, nqh, failed};
      fwrite(o, 4, 3, stdout);
      fwrite(oe.data(), 8, n, stdout);
      fwrite(ox.data(), 8, (size_t)n * nI, stdout);
      fwrite(kolt.data(), 8, 3, stdout);
      fwrite(&flops, 8, 1, stdout);
    }
  }
  return 0;
}
