// ref_harness_2b.cpp -- golden-vector writer for the two-body propagator
// matrices of the compiled reference.
//
// TEST INFRASTRUCTURE.  Linked against the reference's own objects (built from
// /root/reference/source by oracle/Makefile, target `ref`); it only drives
// reference functions and prints their outputs with %.17g.
//
//   ref_harness_2b <model_space> <Vpp.bcd> <sp_prop> <out.txt> <jmodes>
//   ref_harness_2b <model_space> <Vpp.bcd> <sp_prop> <out.txt> time <seconds>
//
// The second form is the CPU baseline of the two-body benchmark: it builds
// the mode-3 matrices of every channel with a non-empty basis (ladder, then
// ring, in the CoupledCluster_t loop order) until <seconds> have passed and
// prints one line per finished channel, "kind J pi dT ndim seconds".
//
// For every two-body channel (J <= Jmax, parity, charge) it prints the ladder
// basis (gII_prop_t::Build_LaddRPA_basis, gII_b.cpp:115-200) and the ring basis
// (Pol_prop_t::Build_phRPA_basis, PolProp_b.cpp:96-175), the mode-3 matrices
// without free poles (the call CoupledCluster_t makes, CCD_a.cpp:176-230) and,
// for J <= jmodes, every other mode with free poles added
// (Fill_LaddRPA_mtx, gII_b.cpp:288-563; Fill_phRPA_mtx, PolProp_b.cpp:255-563).
// Matrices are written into a buffer of leading dimension ndim+2 prefilled
// with -777.25, so the extent the reference clears is recorded too.  The
// free-propagator mode 4 runs last: the reference keeps its switch in a
// file-scope static that stays set for every later call of the process.
// Finally the bases are rebuilt with the n1 labels set to (index % 3) and
// the cuts icut = +1 / -1 with ntcut = 2.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <vector>

using namespace std;
#include "BcDor-Run_vars.hh"
#include "BcDor-gII_classes.hh"
#include "BcDor-PolProp_classes.hh"

static const double kFill = -777.25;

static void put_ints(FILE *f, const int *x, int n) {
  for (int i = 0; i < n; ++i) fprintf(f, "%s%d", i ? " " : "", x[i]);
  fprintf(f, "\n");
}

static void put_dbls(FILE *f, const double *x, int n) {
  for (int i = 0; i < n; ++i) fprintf(f, "%s%.17g", i ? " " : "", x[i]);
  fprintf(f, "\n");
}

static void put_ladd_basis(FILE *f, const char *tag, gII_prop_t &L, int J, int pi, int dT) {
  int npp = L.n_pp_basis > 0 ? L.n_pp_basis : 0, nhh = L.n_hh_basis > 0 ? L.n_hh_basis : 0;
  fprintf(f, "%s %d %d %d %d %d\n", tag, J, pi, dT, npp, nhh);
  put_ints(f, L.n1_pp, npp); put_ints(f, L.n2_pp, npp); put_dbls(f, L.e_pp, npp);
  put_ints(f, L.k1_hh, nhh); put_ints(f, L.k2_hh, nhh); put_dbls(f, L.e_hh, nhh);
}

static void put_ring_basis(FILE *f, const char *tag, Pol_prop_t &R, int J, int pi, int dT) {
  int nph = R.n_ph_basis > 0 ? R.n_ph_basis : 0, nhp = R.n_hp_basis > 0 ? R.n_hp_basis : 0;
  fprintf(f, "%s %d %d %d %d %d\n", tag, J, pi, dT, nph, nhp);
  put_ints(f, R.n_ph, nph); put_ints(f, R.k_ph, nph); put_dbls(f, R.e_ph, nph);
  put_ints(f, R.n_hp, nhp); put_ints(f, R.k_hp, nhp); put_dbls(f, R.e_hp, nhp);
}

static int fill(gII_prop_t &X, int mode, double *buf, int ld, double *fp, int add) {
  return X.Fill_LaddRPA_mtx(mode, buf, ld, fp, add != 0);
}

static int fill(Pol_prop_t &X, int mode, double *buf, int ld, double *fp, int add) {
  return X.Fill_phRPA_mtx(mode, buf, ld, fp, add != 0);
}

// One Fill call into a prefilled buffer; prints "<tag> J pi dT mode add ndim ld rc",
// ndim rows of ld values, then the ndim free poles (prefilled too).
template <class T>
static void put_fill(FILE *f, const char *tag, T &X, int J, int pi, int dT, int mode, int add, int ndim) {
  const int ld = (ndim > 0 ? ndim : 1) + 2;
  const int rows = ndim > 0 ? ndim : 1;
  vector<double> buf((size_t)rows * ld, kFill), fp(rows, kFill);
  const int rc = fill(X, mode, &buf[0], ld, &fp[0], add);
  fprintf(f, "%s %d %d %d %d %d %d %d %d\n", tag, J, pi, dT, mode, add, ndim, ld, rc);
  for (int r = 0; r < ndim; ++r) put_dbls(f, &buf[(size_t)r * ld], ld);
  put_dbls(f, &fp[0], ndim);
}

static double now_s() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

// Timed builds of every channel's mode-3 matrices (no output of the matrices).
static void time_builds(FILE *out, ModSpace_t &msp, VppInt_t &vpp, SpProp_t &g, double budget) {
  const double t_end = now_s() + budget;
  for (int J = 0; J <= msp.Jmax; ++J)
    for (int pi = 0; pi < 2; ++pi) {
      for (int dT = msp.ch_pp_mn; dT <= msp.ch_pp_mx; ++dT) {
        if (now_s() > t_end) return;
        gII_prop_t L(&msp, &g, &vpp);
        L.Build_LaddRPA_basis(J, dT, pi);
        const int n = (L.n_pp_basis > 0 ? L.n_pp_basis : 0) + (L.n_hh_basis > 0 ? L.n_hh_basis : 0);
        if (n == 0) continue;
        vector<double> buf((size_t)n * n), fp(n);
        const double t0 = now_s();
        L.Fill_LaddRPA_mtx(3, &buf[0], n, &fp[0], false);
        fprintf(out, "0 %d %d %d %d %.6f\n", J, pi, dT, n, now_s() - t0);
        fflush(out);
      }
      for (int dT = msp.ch_ph_mn; dT <= msp.ch_ph_mx; ++dT) {
        if (now_s() > t_end) return;
        Pol_prop_t R(&msp, &g, &vpp);
        R.Build_phRPA_basis(J, dT, pi);
        const int n = (R.n_ph_basis > 0 ? R.n_ph_basis : 0) + (R.n_hp_basis > 0 ? R.n_hp_basis : 0);
        if (n == 0) continue;
        vector<double> buf((size_t)n * n), fp(n);
        const double t0 = now_s();
        R.Fill_phRPA_mtx(3, &buf[0], n, &fp[0], false);
        fprintf(out, "1 %d %d %d %d %.6f\n", J, pi, dT, n, now_s() - t0);
        fflush(out);
      }
    }
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s msp vpp prop out jmodes | time seconds\n", argv[0]);
    return 2;
  }
  BcDor_Init();
  ofstream sink("/dev/null");
  streambuf *keep = cout.rdbuf(sink.rdbuf());

  ModSpace_t msp(argv[1]);
  msp.Calculate_Jch_bounds();
  VppInt_t vpp(argv[2], &msp);
  SpProp_t g(argv[3], &msp);
  FILE *out = fopen(argv[4], "w");
  if (strcmp(argv[5], "time") == 0) {
    time_builds(out, msp, vpp, g, argc > 6 ? atof(argv[6]) : 10.0);
    fclose(out);
    cout.rdbuf(keep);
    return 0;
  }
  const int jmodes = atoi(argv[5]);
  fprintf(out, "bounds %d %d %d %d %d\n", msp.Jmax, msp.ch_pp_mn, msp.ch_pp_mx, msp.ch_ph_mn, msp.ch_ph_mx);

  for (int pass = 0; pass < 2; ++pass) {     // pass 1: the free-propagator mode only
    for (int J = 0; J <= msp.Jmax; ++J)
      for (int pi = 0; pi < 2; ++pi) {
        for (int dT = msp.ch_pp_mn; dT <= msp.ch_pp_mx; ++dT) {
          gII_prop_t L(&msp, &g, &vpp);
          L.Build_LaddRPA_basis(J, dT, pi);
          const int npp = L.n_pp_basis > 0 ? L.n_pp_basis : 0, nhh = L.n_hh_basis > 0 ? L.n_hh_basis : 0;
          if (pass == 1) { put_fill(out, "lmat", L, J, pi, dT, 4, 0, npp + nhh); continue; }
          put_ladd_basis(out, "lbasis", L, J, pi, dT);
          put_fill(out, "lmat", L, J, pi, dT, 3, 0, npp + nhh);
          if (J <= jmodes) {
            put_fill(out, "lmat", L, J, pi, dT, 0, 1, npp + nhh);
            put_fill(out, "lmat", L, J, pi, dT, 1, 1, npp);
            put_fill(out, "lmat", L, J, pi, dT, 2, 1, nhh);
            put_fill(out, "lmat", L, J, pi, dT, 3, 1, npp + nhh);
          }
        }
        for (int dT = msp.ch_ph_mn; dT <= msp.ch_ph_mx; ++dT) {
          Pol_prop_t R(&msp, &g, &vpp);
          R.Build_phRPA_basis(J, dT, pi);
          const int nph = R.n_ph_basis > 0 ? R.n_ph_basis : 0, nhp = R.n_hp_basis > 0 ? R.n_hp_basis : 0;
          if (pass == 1) { put_fill(out, "rmat", R, J, pi, dT, 4, 0, nph + nhp); continue; }
          put_ring_basis(out, "rbasis", R, J, pi, dT);
          put_fill(out, "rmat", R, J, pi, dT, 3, 0, nph + nhp);
          if (J <= jmodes) {
            put_fill(out, "rmat", R, J, pi, dT, 0, 1, nph + nhp);
            put_fill(out, "rmat", R, J, pi, dT, 1, 1, nph);
            put_fill(out, "rmat", R, J, pi, dT, 3, 1, nph + nhp);
          }
        }
      }
  }

  // bases under the n1 cut
  for (int i = 0; i < g.n_qp; ++i) g.n1p[i] = i % 3;
  for (int i = 0; i < g.n_qh; ++i) g.n1h[i] = i % 3;
  for (int icut = -1; icut <= 1; icut += 2)
    for (int J = 0; J <= msp.Jmax; ++J)
      for (int pi = 0; pi < 2; ++pi) {
        for (int dT = msp.ch_pp_mn; dT <= msp.ch_pp_mx; ++dT) {
          gII_prop_t L(&msp, &g, &vpp);
          L.Build_LaddRPA_basis(J, dT, pi, icut, 2);
          fprintf(out, "lcut %d\n", icut);
          put_ladd_basis(out, "lbasis", L, J, pi, dT);
        }
        for (int dT = msp.ch_ph_mn; dT <= msp.ch_ph_mx; ++dT) {
          Pol_prop_t R(&msp, &g, &vpp);
          R.Build_phRPA_basis(J, dT, pi, icut, 2);
          fprintf(out, "rcut %d\n", icut);
          put_ring_basis(out, "rbasis", R, J, pi, dT);
        }
      }
  fclose(out);
  cout.rdbuf(keep);
  return 0;
}
