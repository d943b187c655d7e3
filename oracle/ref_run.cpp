// ref_run.cpp -- full-precision recorder around the compiled reference program.
//
// TEST INFRASTRUCTURE.  Linked against every object of the reference
// (compiled by oracle/Makefile, target `refrun`, from /root/reference/source;
// BcDor-Main.cpp with its main() renamed) and LAPACK dsyevd_ from MKL.  It runs
// the reference's own main() -- command line, Run_itrs (Main.cpp:105-479),
// MakeSelfEn, MakePivots -- unchanged, and records what the reference's text
// outputs print with only ~7 digits:
//
//   * every DysonMatrix call (DysonMatrix.cpp:123-532), through the linker's
//     --wrap of the call in Main.o: the subshell, the Fermi energy it splits
//     at (Seek_EFermi, DysonMatrix.cpp:619-635, evaluated on the same input
//     propagator just before the call), the number of lost poles, the Calc_En
//     partial sums x1/x2/x3 (DysonMatrix.cpp:542-594) and the fragment counts
//     it appended;
//   * every convergence check (SpProp_t::quick_n_dirty_diff,
//     Sp_Propagator.cpp:1189-1213, called at Main.cpp:438-439): the summed
//     difference and the whole new propagator at %.17g.  At each check the
//     previous propagator file moves to <file>.prev, so after the run
//     <file>.prop holds the last iteration's output and <file>.prev its input.
//
//   BCD_REF_DUMP=<file> ref_run <reference command line ...>
//
// <file> receives the records.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

using namespace std;
#include "BcDor-Run_vars.hh"

int bcdor_ref_main(int argc, char **argv);
double Seek_EFermi(SpProp_t *gin, int ish, int iplot);

static FILE *g_dump = 0;
static const char *g_dump_name = 0;
static int g_ncheck = 0, g_ndyson = 0;

extern "C" {
int __real__Z11DysonMatrixP18Wide_self_energy_tP8SpProp_tPdS3_S3_(Wide_self_energy_t *, SpProp_t *, double *,
                                                                  double *, double *);
double __real__ZN8SpProp_t18quick_n_dirty_diffEPS_(SpProp_t *self, SpProp_t *other);

int __wrap__Z11DysonMatrixP18Wide_self_energy_tP8SpProp_tPdS3_S3_(Wide_self_energy_t *se, SpProp_t *dout,
                                                                  double *x1, double *x2, double *x3) {
  const int ish = se->I_SH;
  const double ef = Seek_EFermi(se->gin, ish, 0);
  const int np0 = dout ? dout->Sp_clj_np[ish] : 0, nh0 = dout ? dout->Sp_clj_nh[ish] : 0;
  const int lost = __real__Z11DysonMatrixP18Wide_self_energy_tP8SpProp_tPdS3_S3_(se, dout, x1, x2, x3);
  if (g_dump) {
    fprintf(g_dump, "dyson %d %d %.17g %d %.17g %.17g %.17g %d %d %d %d\n", g_ndyson, ish, ef, lost, *x1, *x2, *x3,
            dout ? dout->Sp_clj_np[ish] - np0 : 0, dout ? dout->Sp_clj_nh[ish] - nh0 : 0, se->N_PLS_fw,
            se->N_PLS_bk);
    fflush(g_dump);
  }
  ++g_ndyson;
  return lost;
}

double __wrap__ZN8SpProp_t18quick_n_dirty_diffEPS_(SpProp_t *self, SpProp_t *other) {
  const double d = __real__ZN8SpProp_t18quick_n_dirty_diffEPS_(self, other);
  if (g_dump) {
    fprintf(g_dump, "check %d %.17g %d %d\n", g_ncheck, d, other->n_qp, other->n_qh);
    fflush(g_dump);
    char fn[4096], fprev[4096];
    snprintf(fn, sizeof fn, "%s.prop", g_dump_name);
    snprintf(fprev, sizeof fprev, "%s.prev", g_dump_name);
    (void)rename(fn, fprev);
    FILE *f = fopen(fn, "w");
    if (f) {
      const ModSpace_t *ms = other->MdSp;
      fprintf(f, "prop %d %d %d %d\n", ms->nsubsh, other->N_SPB_MX, other->n_qp, other->n_qh);
      for (int s = 0; s < ms->nsubsh; ++s) {
        fprintf(f, "sub %d %d %d\n", s, other->Sp_clj_np[s], other->Sp_clj_nh[s]);
        for (int side = 0; side < 2; ++side) {
          const int n = side ? other->Sp_clj_nh[s] : other->Sp_clj_np[s];
          const double *e = side ? other->Sp_clj_eh[s] : other->Sp_clj_ep[s];
          const double *z = side ? other->Sp_clj_zh[s] : other->Sp_clj_zp[s];
          const double *x = side ? other->Sp_clj_fqh[s] : other->Sp_clj_fqp[s];
          for (int k = 0; k < n; ++k) {
            fprintf(f, "%c %.17g %.17g", side ? 'h' : 'p', e[k], z[k]);
            for (int r = 0; r < ms->MSp_no[s]; ++r) fprintf(f, " %.17g", x[(long)k * other->N_SPB_MX + r]);
            fprintf(f, "\n");
          }
        }
      }
      fclose(f);
    }
  }
  ++g_ncheck;
  return d;
}
}

int main(int argc, char **argv) {
  g_dump_name = getenv("BCD_REF_DUMP");
  if (g_dump_name) g_dump = fopen(g_dump_name, "w");
  // the reference's chatter goes to stdout; callers redirect it
  const int rc = bcdor_ref_main(argc, argv);
  if (g_dump) fclose(g_dump);
  return rc;
}
