// ref_harness.cpp -- golden-vector writer around the compiled reference.
//
// TEST INFRASTRUCTURE.  Linked against the reference's own hot-path objects
// (built from /root/reference/source by oracle/Makefile, target `ref`); it
// only drives reference functions and prints their outputs with %.17g.
//
//   ref_harness <model_space> <Vpp.bcd> <sp_prop> <pivots|-> <out.txt> [bcdwk_dir]
//
// Per subshell it records: the pole counts (Count_2p1h_2h1p_poles,
// Self_energy_ext_Op_c.cpp:35-113), Sigma^inf (Build_ExtendedHartreeFock ->
// ExtendedHartreeFock_Pot, Self_energy_ext_Op_b.cpp:71-80 and
// Sp_Propagator_a.cpp:39-130), the Sigma(2) rows in build order
// (Build_Dynamic_SlfEn_2ndOrder, Self_energy_ext_Op_c.cpp:323-561) and, with a
// pivots file, the Lanczos-reduced rows (Reduce_SelfEn_Lncz,
// Self_energy_ext_Op_e.cpp:55-232 -> Reduce_Mtx_Lncz, Lanczos.cpp:44-286).
// With bcdwk_dir it also stores each channel's Sigma(2) with the reference's
// own save_SelfEn_bin (Self_energy_ext_Op_d.cpp:55-133): the unreduced one as
// "Unred" (both sides), "UnredFw" (i_FwBk=+1) and "UnredBk" (i_FwBk=-1), and
// the Lanczos-reduced one likewise as "Lanc", "LancFw" and "LancBk".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

// the reference headers expect std names in scope
using namespace std;
#include "BcDor-Run_vars.hh"

static void save_three(Wide_self_energy_t &s, const char *dir, const char *label) {
  char lab[64];
  strcpy(s.STOREDIR, dir);
  const int side[3] = {0, 1, -1};
  const char *suffix[3] = {"", "Fw", "Bk"};
  for (int k = 0; k < 3; ++k) {
    sprintf(lab, "%s%s", label, suffix[k]);
    s.save_SelfEn_bin(side[k], lab);
  }
}

static void put_rows(FILE *f, const char *tag, int ish, int n, int w, const double *r) {
  fprintf(f, "%s %d %d %d\n", tag, ish, n, w);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < w; ++k) fprintf(f, "%s%.17g", k ? " " : "", r[(long)i * w + k]);
    fprintf(f, "\n");
  }
}

// Pivots file blocks "shell,npvts: ish n" followed by n lines
// "n_itr_fw n_itr_bk u(0..no-1)" (the layout Make_LancPivots reads).
static int read_pivots(const char *fn, int ish, int no, vector<double> &vec, vector<int> &ifw,
                       vector<int> &ibk) {
  FILE *in = fopen(fn, "r");
  if (!in) return -1;
  char line[4096];
  int n = 0;
  while (fgets(line, sizeof line, in)) {
    int s, np;
    if (sscanf(line, "shell,npvts:%d%d", &s, &np) != 2 || s != ish) continue;
    for (int i = 0; i < np; ++i) {
      if (!fgets(line, sizeof line, in)) break;
      int a, b, used;
      char *p = line;
      if (sscanf(p, "%d %d%n", &a, &b, &used) < 2) continue;
      p += used;
      ifw.push_back(a);
      ibk.push_back(b);
      for (int k = 0; k < no; ++k) {
        double x = 0.0;
        if (sscanf(p, "%lf%n", &x, &used) < 1) used = 0;
        p += used;
        vec.push_back(x);
      }
      ++n;
    }
    break;
  }
  fclose(in);
  return n;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s msp vpp prop pivots|- out [bcdwk_dir]\n", argv[0]);
    return 2;
  }
  BcDor_Init();
  ofstream sink("/dev/null");
  streambuf *keep = cout.rdbuf(sink.rdbuf());

  ModSpace_t msp(argv[1]);
  msp.Calculate_Jch_bounds();
  VppInt_t vpp(argv[2], &msp);
  SpProp_t g(argv[3], &msp);
  const bool lanc = strcmp(argv[4], "-") != 0;
  const char *wkdir = argc > 6 ? argv[6] : 0;
  if (wkdir && strlen(wkdir) >= 100) {
    fprintf(stderr, "bcdwk_dir longer than STOREDIR holds\n");
    return 2;
  }
  FILE *out = fopen(argv[5], "w");
  fprintf(out, "nme %d\n", vpp.n_me_stored);

  for (int ish = 0; ish < msp.nsubsh; ++ish) {
    Wide_self_energy_t sf(ish, &msp, &vpp, &g);
    int nfw = 0, nbk = 0;
    sf.Count_2p1h_2h1p_poles(&nfw, &nbk);
    fprintf(out, "count %d %d %d\n", ish, nfw, nbk);

    sf.Build_ExtendedHartreeFock();
    put_rows(out, "hf", ish, sf.NTOT_ORB, sf.NTOT_ORB, sf.Sig_HF);

    int n3 = 0, n4 = 0;
    sf.Build_Dynamic_SlfEn_2ndOrder(&n3, &n4, 1, 0, 0);
    put_rows(out, "fw", ish, sf.N_PLS_fw, sf.NTOT_DIM, sf.Sig_dyn_fw);
    put_rows(out, "bk", ish, sf.N_PLS_bk, sf.NTOT_DIM, sf.Sig_dyn_bk);
    if (wkdir) save_three(sf, wkdir, "Unred");

    if (lanc) {
      vector<double> pv;
      vector<int> ifw, ibk;
      int npv = read_pivots(argv[4], ish, msp.MSp_no[ish], pv, ifw, ibk);
      if (npv <= 0) continue;
      vector<double *> rows(npv);
      for (int k = 0; k < npv; ++k) rows[k] = &pv[(long)k * msp.MSp_no[ish]];
      Wide_self_energy_t sl(ish, &msp, &vpp, &g);
      sl.Reduce_SelfEn_Lncz(&sf, npv, &rows[0], &ifw[0], &ibk[0], 1);
      fprintf(out, "itr %d %d", ish, npv);
      for (int k = 0; k < npv; ++k) fprintf(out, " %d %d", ifw[k], ibk[k]);
      fprintf(out, "\n");
      put_rows(out, "lfw", ish, sl.N_PLS_fw, sl.NTOT_DIM, sl.Sig_dyn_fw);
      put_rows(out, "lbk", ish, sl.N_PLS_bk, sl.NTOT_DIM, sl.Sig_dyn_bk);
      fprintf(out, "ipiv %d %d %d", ish, sl.n_piv_fw, sl.n_piv_bk);
      for (int k = 0; k < sl.n_piv_fw + 2 && sl.N_PLS_fw > 0; ++k) fprintf(out, " %d", sl.i_piv_fw[k]);
      fprintf(out, " |");
      for (int k = 0; k < sl.n_piv_bk + 2 && sl.N_PLS_bk > 0; ++k) fprintf(out, " %d", sl.i_piv_bk[k]);
      fprintf(out, "\n");
      if (wkdir) save_three(sl, wkdir, "Lanc");
    }
  }
  fclose(out);
  cout.rdbuf(keep);
  return 0;
}
