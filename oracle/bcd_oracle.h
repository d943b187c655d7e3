/*
 * bcd_oracle.h -- CPU FP64 restatement of BoccaDorata's second-order SCGF hot path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This library is the checker: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product path (boccadorata-public_amd/) never calls it and has no CPU
 * fallback.
 *
 * Every function restates one reference routine (file:line under source/ of
 * craolus/BoccaDorata-public) in plain C99.  Loop orders, pole enumeration
 * orders and floating-point operation orders follow the reference so that the
 * outputs agree with the compiled reference to the last few ulps.  Parity of
 * each function is pinned in tests/test_oracle_*.py:
 *   orc_vpp_get ............ golden vectors of the compiled reference (V lookup via Sigma rows)
 *   orc_count_poles ........ reference pole counts (SpPropStat totals) + golden vectors
 *   orc_build_sigma2 ....... golden vectors of the compiled reference (Op_c.cpp)
 *   orc_hf_pot ............. golden vectors of the compiled reference (Sp_Propagator_a.cpp)
 *   orc_sort_u / orc_sort_d  ordering invariants + golden Lanczos rows (sorted input)
 *   orc_reduce_mtx_lncz .... golden vectors of the compiled reference (Lanczos.cpp),
 *                            saturation invariant (L = M reproduces unreduced Dyson)
 *   orc_eigh ............... numpy.linalg.eigh (LAPACK) on random and structured matrices
 *   orc_dyson .............. reference-held fixture bcdwk + sp_prop_O16_HF -> sp_prop_O16_2nd_Lanc50,
 *                            numpy eigh on the assembled matrix, one-orbit secular equation
 *   orc_run ................ Koltun energies of the reference fixtures, HF fixed point,
 *                            golden first-iteration Sigma + Dyson chains
 *   orc_s6j ................ exact rational 6j values (tests/test_oracle_2b.py)
 *   orc_pandya, orc_ladd_*,
 *   orc_ph_* ............... golden vectors of the compiled reference (gII_b.cpp, PolProp_b.cpp)
 * The Dyson stage of the reference (DysonMatrix.cpp) links LAPACK dsyevd_, an
 * external library, so it is not built from the reference sources; its parity
 * rests on the reference-held fixture and on the closed-form checks above.
 */
#ifndef BCD_ORACLE_H
#define BCD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- model space + interaction (ModSpace_t, VppInt_t) ------------------ */
typedef struct orc_sys {
  int nsubsh, norb, nspb;      /* nspb = N_SPB_MX = max orbits per subshell      */
  int Jmax;                    /* max 2j, as Calculate_Jch_bounds (Mod_Space.cpp:589-594) */
  int *l, *j2, *ip, *ch, *no, *orb0, *nqp, *nqh;   /* [nsubsh]                */
  int *orb_sh, *orb_n, *orb_F;                     /* [norb]                  */
  double *orb_e;                                   /* [norb]  (Mor_xe)        */
  long n_me;                                       /* VppInt_t storage        */
  int *iabJ, *icdJ;
  double *v;
  int *chipJ_str, *chipJ_end;                      /* [5*128*2]               */
} orc_sys;

/* Lehmann propagator (SpProp_t, Sp_classes.hh:103-170): fragments stored per
 * subshell, subshells contiguous in model-space order. */
typedef struct orc_prop {
  int nsubsh, nspb;
  int cap_qp, cap_qh, n_qp, n_qh;
  int *np, *nh, *op, *oh;      /* [nsubsh] counts and first index              */
  double *ep, *zp, *fqp;       /* [cap_qp], [cap_qp], [cap_qp*nspb]            */
  double *eh, *zh, *fqh;
  int *clj_p, *clj_h;          /* subshell of each fragment (n_clj_p/h)        */
} orc_prop;

orc_sys *orc_sys_create(int nsubsh, const int *l, const int *j2, const int *ip,
                        const int *ch, const int *no, const int *orb_n,
                        const int *orb_F, const double *orb_e,
                        long n_raw, const int *ra, const int *rb, const int *rc,
                        const int *rd, const int *rJ, const double *rv);
void orc_sys_free(orc_sys *s);
long orc_sys_n_me(const orc_sys *s);

double orc_vpp_get(const orc_sys *s, int ia, int ib, int ic, int id, int J);

orc_prop *orc_prop_create(const orc_sys *s, int cap_qp, int cap_qh);
orc_prop *orc_prop_bare(const orc_sys *s);
void orc_prop_free(orc_prop *p);
void orc_prop_clean(orc_prop *p);
int orc_prop_add_qp(orc_prop *p, int ish, int nin, int ndim, const double *sol);
int orc_prop_add_qh(orc_prop *p, int ish, int nin, int ndim, const double *sol);
double orc_prop_diff(const orc_prop *a, const orc_prop *b);
double orc_prop_maxdiff(const orc_prop *a, const orc_prop *b);
orc_prop *orc_prop_copy(const orc_prop *p);

/* heapsorts of pole rows by column isort (Utilities.cpp:25-94) */
void orc_sort_u(int ntot, int ndim, int isort, double *dbl);
void orc_sort_d(int ntot, int ndim, int isort, double *dbl);

void orc_count_poles(const orc_sys *s, int ish, const orc_prop *g, int *nfw, int *nbk);
int orc_build_sigma2(const orc_sys *s, int ish, const orc_prop *g,
                     double *rows_fw, int cap_fw, double *rows_bk, int cap_bk,
                     int *nfw, int *nbk);
int orc_hf_pot(const orc_sys *s, int ish, const orc_prop *g, double *u, int ndim);
void orc_kin_ho(const orc_sys *s, int ish, double htom, double *u, int ndim);

int orc_reduce_mtx_lncz(int NA_out, int NB_out, int NPLS_out, double *Mout, int *i_piv_out,
                        int NA_in, int NB_in, int NPLS_in, const double *Min,
                        int nt_pvts, double *pivots, int *n_itr);
int orc_reduce_selfen_lncz(int n_orb, int nfw, double *rows_fw, int nbk, double *rows_bk,
                           int nt_pvts, double *pivots, int *n_itr_fw, int *n_itr_bk,
                           double *out_fw, int *i_piv_fw, int *nfw_out,
                           double *out_bk, int *i_piv_bk, int *nbk_out, int *ntot_den);

int orc_eigh(int n, double *a, double *w);

double orc_seek_efermi(const orc_sys *s, const orc_prop *g, int ish);
int orc_dyson(int n_orb, int j2, const double *h, const double *u1,
              int ntot_den_fw, int nfw, double *rows_fw, const int *i_piv_fw, int n_piv_fw,
              int ntot_den_bk, int nbk, double *rows_bk, const int *i_piv_bk, int n_piv_bk,
              double efermi, double *sol_qp, int *nqp, double *sol_qh, int *nqh,
              double *kolt3);

/* SCGF driver (Run_itrs, Main.cpp:105-479) */
typedef struct orc_params {
  int mode;        /* 0 = HF (HF=), 1 = 2nd order one shot (2nd), 2 = sc0 (ExtSE=)    */
  int itr_max;     /* ItrMax; < 0 iterates until converged                            */
  int lanczos;     /* 0 = full Sigma(2), 1 = Lanczos-reduced Sigma(2)                 */
  int lanc_kind;   /* 0: every pivot gets lanc_n its per side (MakePivots=lanc_n);
                      1: lanc_n Krylov vectors per side split over the pivots       */
  int lanc_n;
  int conv_kind;   /* 0: sum|de| < conv (quick_n_dirty_diff); 1: max|de| < conv       */
  double conv;
  double t_fac;    /* U1body_fac multiplying htom*T_HO (0 for Trel=7)                 */
  int ext_u1;      /* 1: add diag(model-space spe) (ExtU1=1)                          */
  int n_pivot_sets;/* >0: explicit pivots supplied per subshell (see orc_run)          */
  double hw;       /* htom of the kinetic table (Get_Tab_kin_ho, Utilities.cpp:106-124) */
  int piv_round;   /* 1: derived pivots rounded to the %18.6le text DysPivots writes   */
  int n_ef;        /* forced Fermi energies (Set_external_EFermi, DysonMatrix.cpp:601-617) */
  const int *ef_ish;
  const double *ef_val;
} orc_params;

/* Pivots of subshell ish as DysPivots (Sp_Propagator.cpp:1225-1289) writes
 * them from g: quasiholes then quasiparticles, each by ascending energy with
 * 100*Z >= z_cut; round_print: each component rounded as "%18.6le" prints it
 * and Make_LancPivots (Miscellanea.cpp:366-370) reads it back.  Returns the
 * number of pivots; vec receives them row-major [npv x no(ish)]. */
int orc_make_pivots(const orc_sys *s, const orc_prop *g, int ish, double z_cut, int round_print, double *vec);

/* Seek_EFermi with the forced values of par (DysonMatrix.cpp:619-635): the
 * first forced entry for ish wins, else the midpoint of g. */
double orc_fermi(const orc_sys *s, const orc_prop *g, int ish, const orc_params *par);

typedef struct orc_stats {
  int iterations;           /* number of Dyson sweeps performed                    */
  double last_diff;         /* convergence measure of the last sweep               */
  double e_kolt, u1_tot, a_tot;
  long n_fw_poles, n_bk_poles;
  int lost_poles;
} orc_stats;

/* Explicit pivots: piv_n[ish] pivots for subshell ish, vectors stored row-major
 * [npiv x no(ish)] at piv_vec + piv_off[ish], iteration counts at itr_fw/itr_bk
 * + piv_ioff[ish].  Pass NULLs to derive them from the input propagator as
 * DysPivots (Sp_Propagator.cpp:1225-1289) does with z_cut = 60%. */
int orc_run(const orc_sys *s, const orc_prop *gin, const orc_params *par,
            const int *piv_n, const int *piv_off, const int *piv_ioff,
            const double *piv_vec, const int *itr_fw, const int *itr_bk,
            orc_prop *gout, orc_stats *st);

/* Two-body propagator matrices (bcd_oracle_2b.c): the pp/hh ladder of
 * gII_prop_t (gII_b.cpp:62-563) and the ph ring of Pol_prop_t
 * (PolProp_b.cpp:43-614).  Matrices are stored with entry (c, r) at
 * Sab[c*ld + r]. */
double orc_s6j(int ja, int jb, int jc, int la, int lb, int lc);
double orc_pandya(const orc_sys *s, int ia, int ib, int ig, int id, int Jf);
int orc_ladd_basis(const orc_sys *s, const orc_prop *g, int Jf, int dT, int pif, int icut, int ntcut,
                   const int *n1p, const int *n1h, int cap_pp, int cap_hh, int *n1_pp, int *n2_pp,
                   double *e_pp, int *k1_hh, int *k2_hh, double *e_hh, int *npp, int *nhh);
int orc_ladd_fill(const orc_sys *s, const orc_prop *g, int Jf, int npp, const int *n1_pp, const int *n2_pp,
                  const double *e_pp, int nhh, const int *k1_hh, const int *k2_hh, const double *e_hh,
                  int mode, double *Sab, int ld, double *free_poles, int add_free_poles);
int orc_ph_basis(const orc_sys *s, const orc_prop *g, int Jf, int dT, int pif, int icut, int ntcut,
                 const int *n1p, const int *n1h, int cap_ph, int cap_hp, int *n_ph, int *k_ph, double *e_ph,
                 int *n_hp, int *k_hp, double *e_hp, int *nph, int *nhp);
int orc_ph_fill(const orc_sys *s, const orc_prop *g, int Jf, int dT, int nph, const int *n_ph, const int *k_ph,
                const double *e_ph, int nhp, const int *n_hp, const int *k_hp, const double *e_hp, int mode,
                double *Sab, int ld, double *free_poles, int add_free_poles);

/* accessors for ctypes */
int orc_prop_counts(const orc_prop *p, int *n_qp, int *n_qh, int *nspb);
int orc_prop_export(const orc_prop *p, int *np, int *nh, double *ep, double *zp, double *fqp,
                    double *eh, double *zh, double *fqh);
orc_prop *orc_prop_import(const orc_sys *s, const int *np, const int *nh,
                          const double *ep, const double *fqp,
                          const double *eh, const double *fqh);

#ifdef __cplusplus
}
#endif
#endif
