/*
 * bcdor_amd.h -- C ABI of the MI355X (gfx950) second-order SCGF hot path.
 *
 * Plain pointers and sizes only.  Host arrays use the reference's own
 * layouts (BoccaDorata v1.0, craolus/BoccaDorata-public, files under source/):
 * the model space as ModSpace_t holds it, the interaction as VppInt_t stores
 * it (packed keys + values), propagator fragments as SpProp_t stores them and
 * self-energy pole rows as Wide_self_energy_t stores them, so a reference
 * caller passes its existing arrays unchanged.  All arithmetic runs in FP64
 * on the GPU; there is no CPU fallback: without a usable device every entry
 * point returns BCD_ERR_DEVICE.
 *
 * Return codes: 0 on success, < 0 on error (bcd_error_string).  No entry
 * point calls exit().  Calls on one context are not reentrant; different
 * contexts may be used from different host threads.
 */
#ifndef BCDOR_AMD_H
#define BCDOR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Library error codes sit below -1000 so that they never collide with the
 * reference's own Reduce_Mtx_Lncz codes (-1, -2, -100..-201), which
 * bcd_reduce_mtx_lncz returns unchanged. */
#define BCD_OK 0
#define BCD_ERR_ARG (-1001)
#define BCD_ERR_DEVICE (-1002)
#define BCD_ERR_CAPACITY (-1003)
#define BCD_ERR_LANCZOS (-1004)
#define BCD_ERR_EIGEN (-1005)
#define BCD_ERR_NOMEM (-1006)

/* ModSpace_t arrays (Sp_classes.hh:31-71; filled by ModSpace_t(char*),
 * Mod_Space.cpp:415-504).  Orbits are numbered subshell by subshell. */
typedef struct bcd_modspace {
  int nsubsh;
  const int *l, *j2, *ip, *ch, *no;      /* [nsubsh]: l, 2j, parity, charge, orbits   */
  const int *orb_n, *orb_F;               /* [norb]: radial n (Mor_n), hole flag (Mor_F) */
  const double *orb_e;                    /* [norb]: spe column (Mor_xe)                */
} bcd_modspace;

/* VppInt_t storage (Int_classes.hh:20-33): n_me_stored unique elements with
 * keys ((i1*1024+i2)*128+J), i1 <= i2, iabJ <= icdJ (Vpp_potential.cpp:82)
 * and values Vpp_wrk, phases already applied as read_ascii does. */
typedef struct bcd_vpp {
  long n_me;
  const int *iabJ, *icdJ;
  const double *v;
} bcd_vpp;

/* SpProp_t fragments (Sp_classes.hh:78-102): per subshell np[i] quasiparticles
 * then, in a separate array, nh[i] quasiholes, subshells in model-space
 * order; overlaps fq*[k*nspb + n] (N_SPB_MX stride). */
typedef struct bcd_prop {
  int nspb;
  const int *np, *nh;                     /* [nsubsh]                                  */
  const double *ep, *fqp;                 /* [sum np], [sum np * nspb]                 */
  const double *eh, *fqh;                 /* [sum nh], [sum nh * nspb]                 */
} bcd_prop;

typedef struct bcd_ctx bcd_ctx;
typedef struct bcd_batch bcd_batch;

int bcd_version(void);
const char *bcd_error_string(int code);
int bcd_device_count(void);

/* ---------------------------------------------------------------- context
 * A context holds one system on one device: the model-space tables and the
 * interaction packed into dense J-blocks in HBM.  Replaces the set-up of
 * ModSpace_t + VppInt_t (Load_ModSp / Load_interaction, Initialize.cpp:93-157). */
int bcd_ctx_create(int device, const bcd_modspace *ms, const bcd_vpp *vpp, bcd_ctx **out);
void bcd_ctx_destroy(bcd_ctx *ctx);

/* Count_2p1h_2h1p_poles (Self_energy_ext_Op_c.cpp:35-113), icut == 0. */
int bcd_count_poles(bcd_ctx *ctx, int ish, const bcd_prop *g, int *nfw, int *nbk);

/* Build_Dynamic_SlfEn_2ndOrder (Self_energy_ext_Op_c.cpp:323-561), icut == 0:
 * the 2p1h (fw) and 2h1p (bk) pole rows [E, sum_d m_d^2, m_0..m_{n-1}] of
 * subshell ish in the reference's build order; row width 2 + no[ish]. */
int bcd_sigma2_build(bcd_ctx *ctx, int ish, const bcd_prop *g,
                     double *rows_fw, int cap_fw, double *rows_bk, int cap_bk,
                     int *nfw, int *nbk);

/* SpProp_t::ExtendedHartreeFock_Pot (Sp_Propagator_a.cpp:39-130): Sigma^inf of
 * subshell ish from the quasihole densities of g; u is ndim x ndim row-major. */
int bcd_hf_pot(bcd_ctx *ctx, int ish, const bcd_prop *g, double *u, int ndim);

/* Reduce_Mtx_Lncz (Lanczos.cpp:44-286; declared in Run_vars.hh:193-195), same
 * arguments and return value, with the pivot rows flattened to
 * pivots[nt_pvts * NB_in] (normalised in place, as the reference does). */
int bcd_reduce_mtx_lncz(bcd_ctx *ctx, int NA_out, int NB_out, int NPLS_out, double *Mout,
                        int *i_piv_out, int NA_in, int NB_in, int NPLS_in, const double *Min,
                        int nt_pvts, double *pivots, int *n_itr);

/* bcd_reduce_mtx_lncz without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and return codes as
 * bcd_reduce_mtx_lncz; every device buffer of the call lives in a fixed
 * scratch region compiled into the library (bcd_noalloc_scratch_bytes(),
 * 64 MiB), one per device, held until the process ends.  A call whose
 * buffers exceed it returns BCD_ERR_CAPACITY before any GPU work (the
 * Krylov workspace takes 8*(n_itr total + NB_in + 8)*NPLS_in bytes, rounded
 * up to an even NPLS_in; the input rows 8*NPLS_in*(NA_in+NB_in)).  Calls that
 * overlap on one device share that region and race: issue them from one
 * host thread per device, one after another. */
int bcd_reduce_mtx_lncz_noalloc(bcd_ctx *ctx, int NA_out, int NB_out, int NPLS_out, double *Mout,
                                int *i_piv_out, int NA_in, int NB_in, int NPLS_in, const double *Min,
                                int nt_pvts, double *pivots, int *n_itr);
/* Size in bytes of the per-device scratch of bcd_reduce_mtx_lncz_noalloc. */
long bcd_noalloc_scratch_bytes(void);

/* Wide_self_energy_t::Reduce_SelfEn_Lncz (Self_energy_ext_Op_e.cpp:55-232):
 * sorts the unreduced rows in place (fw ascending, bk descending), rescales
 * n_itr in place when it exceeds the pole count, reduces both sides.  Output
 * rows have *ntot_den = max(2, nt_pvts+1) leading columns. */
int bcd_reduce_selfen_lncz(bcd_ctx *ctx, int n_orb, int nfw, double *rows_fw, int nbk,
                           double *rows_bk, int nt_pvts, double *pivots, int *n_itr_fw,
                           int *n_itr_bk, double *out_fw, int *i_piv_fw, int *nfw_out,
                           double *out_bk, int *i_piv_bk, int *nbk_out, int *ntot_den);

/* DysonMatrix (DysonMatrix.cpp:123-532) without the propagator append: builds
 * [h M_fw^T M_bk^T; M_fw D_fw 0; M_bk 0 D_bk] from the static part u1 + h and the
 * pole rows (diagonal D for npiv == 0, fishbone for a Lanczos-reduced side),
 * diagonalises it, splits at efermi and returns records [e, Z, X(0..n_orb-1)]
 * sorted by descending e.  kolt3 (may be NULL) = Calc_En partial sums
 * {e1b, esum, asum} (DysonMatrix.cpp:542-594). */
int bcd_dyson(bcd_ctx *ctx, int n_orb, int j2, const double *h, const double *u1,
              int ntot_den_fw, int nfw, const double *rows_fw, const int *i_piv_fw, int npiv_fw,
              int ntot_den_bk, int nbk, const double *rows_bk, const int *i_piv_bk, int npiv_bk,
              double efermi, double *sol_qp, int *nqp, double *sol_qh, int *nqh, double *kolt3);

/* Batched symmetric eigensolver standing in for LAPACK dsyevd_('V','U', ...)
 * (DysonMatrix.cpp:35-36, 412): nmat row-major symmetric matrices of orders
 * n[k] at a + off[k]; eigenvalues ascending to w + woff[k]; eigenvectors
 * overwrite the matrices column-wise. */
int bcd_syevd_batched(bcd_ctx *ctx, int nmat, const int *n, const long *off, double *a,
                      const long *woff, double *w);

/* ------------------------------------------------------------ batched solve
 * A batch holds n_inst independent systems (any mix of model spaces) resident
 * on one device and runs the whole SCGF solve of each on the GPU: HF from the
 * given (or bare) propagator, Sigma(2) from the HF result, optional Lanczos
 * reduction, then the sc0 iteration of Sigma^inf + Dyson to convergence
 * (the reference's HF=-1 -> MakeSelfEn 2nd [Lanczos] -> ExtSE=-1 chain,
 * Run_itrs, Main.cpp:105-479). */
typedef struct bcd_solve_params {
  int mode;          /* 0: HF only; 1: HF then one 2nd-order sweep; 2: HF then sc0  */
  int hf_itr_max;    /* ItrMax of the HF stage (< 0: until converged)               */
  int sc_itr_max;    /* ItrMax of the sc0 stage (< 0: until converged)              */
  int lanczos;       /* 1: Lanczos-reduce Sigma(2) before the Dyson sweeps          */
  int lanc_kind;     /* 0: lanc_n iterations per pivot (MakePivots=lanc_n);
                        1: lanc_n Krylov vectors per side split over the pivots     */
  int lanc_n;
  int conv_kind;     /* 0: sum_i |de_i| (quick_n_dirty_diff); 1: max_i |de_i|         */
  double hf_conv;    /* HF stops when the measure drops below this                   */
  double sc_conv;    /* sc0 stops when the measure drops below this                  */
  double t_fac;      /* U1body_fac (Initialize.cpp:65-74; 0: no kinetic term, Trel=7) */
  int ext_u1;        /* 1: add diag(model-space spe) to U1 (ExtU1=1)                */
  int dyson_solver;  /* 0: arrowhead updates over the diagonalised pole blocks
                        (each reduced side diagonalised once per solve);
                        1: dense Householder + QL of the whole Dyson matrix per
                        sweep.  Both give DysonMatrix's eigenpairs.              */
  int precision;     /* 0: FP64 throughout (Sigma(2) rows bit-identical to the
                        reference).  Mixed: the Sigma(2) coupling
                        contraction (Self_energy_ext_Op_c.cpp:454-468) on
                        the matrix cores with FP32 accumulation and 1: FP32
                        operands, 2: BF16 operands; 3: on the vector units
                        with FP32 operands and products, FP64 sums.  The
                        Lanczos reduction, eigensolves and sc loop stay
                        FP64.  The rows' errors propagate to the energies
                        (cancellation in the contraction amplifies them);
                        each mode's error against FP64 is measured, not
                        bounded.                                          */
  int accel;         /* 0: the reference's iteration (the hole density of each
                        sweep's output feeds the next Sigma^inf); m = 1..8:
                        Anderson mixing of that density over the last m
                        iterates in the sc0 stage (the HF stage stays
                        plain).  Same stopping rule, fewer sweeps; where
                        the interaction admits several self-consistent
                        solutions the mixed iteration can settle on another
                        one than the plain iteration does. */
  double hw;         /* htom (hbar*omega) of the kinetic term: U1 = (hw * T_HO) *
                        t_fac, the order of Get_Tab_kin_ho(...) *= U1body_fac
                        (Utilities.cpp:106-124, Main.cpp:266-270)              */
  double mix_start;  /* accel > 0: the density residual norm below which the
                        mixing starts (plain sweeps until then, so that the
                        plain iteration picks the fixed point); <= 0: 1e-3 */
  int piv_round;     /* 1: the Lanczos pivots derived from the HF propagator are
                        rounded as the pivots file carries them (DysPivots
                        prints %18.6le, Make_LancPivots reads it back;
                        Sp_Propagator.cpp:1252, Miscellanea.cpp:366-370), so
                        the chain starts where the reference's MakePivots ->
                        LanDysPiv workflow starts; 0: full precision         */
} bcd_solve_params;

/* g0 may be NULL (bare propagator, SpProp_t(ModSpace_t*, 1)); otherwise g0[i]
 * is the starting propagator of instance i. */
int bcd_batch_create(int device, int n_inst, const bcd_modspace *ms, const bcd_vpp *vpp,
                     const bcd_prop *g0, const bcd_solve_params *par, bcd_batch **out);
/* As bcd_batch_create, with a Lanczos length per instance: lanc_n[i] (> 0)
 * replaces par->lanc_n for instance i, read with par->lanc_kind (iterations
 * per pivot, or Krylov vectors per side split over the pivots -- a pivots
 * file's per-pivot n_itr columns, Miscellanea.cpp:348-380).  lanc_n may be
 * NULL. */
int bcd_batch_create_ex(int device, int n_inst, const bcd_modspace *ms, const bcd_vpp *vpp,
                        const bcd_prop *g0, const bcd_solve_params *par, const int *lanc_n,
                        bcd_batch **out);
/* The same batch built one instance at a time, for batches whose
 * interactions together outgrow host memory: begin sizes it, add packs one
 * instance (model space, interaction, optional start g0, Lanczos length
 * lanc_n or 0 for par->lanc_n), moves its interaction to HBM and frees the
 * host copy; end lays out the device state.  add may be called from several
 * host threads at once for different inst.  Every instance must be added
 * before end; solve and the queries need end. */
int bcd_batch_begin(int device, int n_inst, const bcd_solve_params *par, bcd_batch **out);
int bcd_batch_add(bcd_batch *b, int inst, const bcd_modspace *ms, const bcd_vpp *vpp, const bcd_prop *g0,
                  int lanc_n);
int bcd_batch_end(bcd_batch *b);
/* What the last solve reported beyond its return code: the limit a failing
 * call hit, and instances dropped from the solve (a channel with more poles
 * than the largest Lanczos tile, 40960; bcd_batch_converged then gives 0 for
 * them while the others are solved).  Writes at most n-1 characters and a
 * terminating 0 to buf; returns the full length. */
int bcd_batch_error_detail(const bcd_batch *b, char *buf, int n);
/* Change the iteration settings of a built batch for its later solves:
 * precision, accel, mix_start, hf/sc ItrMax, conv_kind and the tolerances.  The other
 * fields must equal the batch's (they size its device state); otherwise
 * BCD_ERR_ARG. */
int bcd_batch_set_params(bcd_batch *b, const bcd_solve_params *par);
/* Set_external_EFermi (DysonMatrix.cpp:601-617; the SetEf= flag,
 * Command_line.cpp:347-350): every later Dyson sweep of subshell ish of
 * instance inst splits quasiparticles from quasiholes at ef instead of the
 * midpoint Seek_EFermi takes from the input propagator (DysonMatrix.cpp:
 * 619-635).  The first value set for a subshell stays, as in the reference. */
int bcd_batch_set_efermi(bcd_batch *b, int inst, int ish, double ef);
/* The rounding of a pivots file round trip (DysPivots' "%18.6le" read back
 * with %lf) on device: out[i] = x[i] with 7 significant decimal digits, as
 * the batched solve applies it when piv_round = 1.  Exact for |x| >= 1e-16;
 * below, within an ulp or two. */
int bcd_pivot_round(int device, const double *x, double *out, long n);
void bcd_batch_destroy(bcd_batch *b);
/* Run the full solve of every instance; inputs stay resident, so repeated
 * calls redo the same work. */
int bcd_batch_solve(bcd_batch *b);
/* 1: time every kernel family with hipEvents during bcd_batch_solve (adds a
 * synchronisation per launch); 0 (default): stage events only. */
int bcd_batch_set_profile(bcd_batch *b, int on);
/* The sweeps of a stage run in chunks between reads of the convergence
 * flags; a full chunk is captured once per stage into a HIP graph and
 * replayed (on by default; off while profiling, and for the chunk that
 * applies a learned start).  on = 0 launches every sweep directly.  Results
 * are the same either way.  The library makes its own device-synchronising
 * HIP calls (null-stream copies, allocation) under a lock it shares with its
 * captures; a program that makes such calls itself on other host threads
 * while a batch solves should switch the batch's graphs off. */
int bcd_batch_set_graphs(bcd_batch *b, int on);
/* Per-stage device time of the last solve in ms (hipEvents on the batch's
 * stream): [0] W/Sigma^inf set-up, [1] HF loop, [2] Sigma(2) build,
 * [3] Lanczos, [4] sc0 loop, [5] total.  n >= 6. */
int bcd_batch_timings(bcd_batch *b, double *ms, int n);
/* Iterations, last convergence measure and Koltun energy of instance i. */
int bcd_batch_stats(bcd_batch *b, int inst, int *hf_its, int *sc_its, double *last_diff,
                    double *e_kolt, long *n_fw_poles, long *n_bk_poles);
/* 1 when every stage of the last solve of instance i met its convergence
 * criterion (rather than stopping at ItrMax), 0 otherwise, < 0 on error. */
int bcd_batch_converged(bcd_batch *b, int inst);
/* Final propagator of instance i: first call with NULL arrays to get the
 * totals, then with arrays sized sum np / sum nh (overlaps nspb-strided). */
int bcd_batch_result(bcd_batch *b, int inst, int nspb, int *np, int *nh, int *n_qp, int *n_qh,
                     double *ep, double *zp, double *fqp, double *eh, double *zh, double *fqh);
/* Average device duration (ms) and launch count of each kernel family over
 * the last solve, for roofline accounting (hipEvents around every launch,
 * only while profiling is on): k = 0 "sigma2", 1 "lanczos", 2 "tridiag",
 * 3 "ql", 4 "density", 5 "sigma_inf", 6 "finalize", 7 "wbuild",
 * 8 "sigma2_plan", 9 "arrow", 10 "poles", 11 "merge" (merging of the
 * J-degenerate pole groups before the Lanczos reduction). */
int bcd_batch_kernel_times(bcd_batch *b, int k, double *avg_ms, long *launches);
/* Algorithmic work of the last solve: [0] coupling FLOPs (2*sum poles*K*n),
 * [1] Lanczos bytes as the reduction moves them (8*M per vector: the pole rows
 * the jobs run on once, each earlier Krylov vector but the latest once per
 * step, the n coupling columns and E*v_last of finished chains per step, the
 * writes), [2] dense Dyson FLOPs (9 n^3
 * per eigensolve), [3] number of Dyson eigensolves, [4] sum of Dyson n^3,
 * [5] arrowhead-solver FLOPs as counted by the kernels, [6] FLOPs of the
 * once-per-solve diagonalisation of the reduced sides (9 L^3 per side).
 * n >= 7.  With n >= 15 and profiling on, [7..14] are the arrowhead kernel's
 * block time (ms summed over blocks) per phase: load and sort, deflation,
 * roots, Loewner couplings, eigenvectors, reorder, split and output, norms.
 * With n >= 17 and the Lanczos on, [15] is the number of Sigma(2) poles over
 * all channels and sides and [16] the number the Lanczos jobs ran on (fewer
 * where J-degenerate pole groups were merged). */
int bcd_batch_work(bcd_batch *b, double *work, int n);
/* Set the Lanczos tile shape of every bucket (n >= 10 shape indices, each
 * holding its bucket's largest pole count).  Shapes split the block sums
 * differently, so a second solve with other shapes shows how far rounding
 * alone moves each solution. */
int bcd_batch_set_lanczos_shapes(bcd_batch *b, const int *shapes, int n);
/* Tile-shape autotuner of the Lanczos kernel (no counterpart in the
 * reference, which runs one Lanczos at a time on the CPU).  Jobs are grouped
 * by pole count into 10 buckets (M <= 256, 512, 1024, 2048, 4096, 8192, 16384,
 * 24576, 32768, 40960); during the next bcd_batch_solve every shape that fits a bucket's
 * largest job is timed on that bucket's own jobs and the fastest is kept for
 * later solves.  Shapes differ only in how the block sums are split, so
 * results agree to rounding. */
int bcd_batch_autotune(bcd_batch *b);
/* Shape index chosen per bucket (n >= 10) and, if tune_ms is not NULL, the
 * timed ms of every (bucket, shape) pair as tune_ms[bucket * 22 + shape]
 * (0 where not timed). */
int bcd_batch_lanczos_shapes(bcd_batch *b, int *shapes, float *tune_ms, int n);

/* ------------------------------------------------ learned start (network)
 * No counterpart in the reference, whose ExtSE iteration always starts from
 * the propagator file it is given (BcDor-Main.cpp:80, SpProp_t propV(...),
 * Main.cpp:216-451).  A 3-layer perceptron n_in -> n_hidden -> n_hidden ->
 * n_out (ReLU between layers) runs on the bf16 matrix cores: weights held
 * as FP32 masters with BF16 copies, activations rounded to BF16 where they
 * enter a product, FP32 accumulation.  Training (mean-squared loss, Adam,
 * beta = 0.9 / 0.999, eps = 1e-8) runs on the device with the library's own
 * forward and backward kernels.  Parameters are laid out as W1 [n_hidden x
 * n_in], b1, W2 [n_hidden x n_hidden], b2, W3 [n_out x n_hidden], b3.
 * Calls on one network are not reentrant. */
typedef struct bcd_mlp bcd_mlp;
int bcd_mlp_create(int device, int n_in, int n_hidden, int n_out, unsigned long long seed, bcd_mlp **out);
void bcd_mlp_destroy(bcd_mlp *m);
long bcd_mlp_num_params(const bcd_mlp *m);
int bcd_mlp_get_params(bcd_mlp *m, float *p);
int bcd_mlp_set_params(bcd_mlp *m, const float *p);
/* y[n x n_out] = net(x[n x n_in]); host arrays. */
int bcd_mlp_forward(bcd_mlp *m, int n, const float *x, float *y);
/* loss = mean over n rows and n_out outputs of (net(x) - t)^2 and its
 * gradient with respect to every parameter (grad[num_params], may be NULL). */
int bcd_mlp_loss_grad(bcd_mlp *m, int n, const float *x, const float *t, float *loss, float *grad);
/* epochs passes of Adam over the n samples in shuffled minibatches of batch
 * rows (seeded), the whole set resident on the device; epoch_loss[epochs]
 * (may be NULL) receives the mean minibatch loss of every epoch. */
int bcd_mlp_fit(bcd_mlp *m, int n, const float *x, const float *t, int epochs, int batch, float lr,
                unsigned long long seed, float *epoch_loss);

/* The network as the start of a batch's sc0 iteration.  The sc0 sweep reads
 * its input propagator only through the quasihole density rho of every
 * channel (Sigma^inf, Sp_Propagator_a.cpp:39-130), so the start is a
 * predicted converged rho.  The first sc0 sweep runs plainly from the HF
 * propagator; then per channel (subshell of an instance, at most
 * BCD_LEARN_NI orbits) the network maps BCD_LEARN_IN inputs -- HF energies,
 * their distance to the Fermi energy, occupations, the second-order estimate
 * D2 of the density correction from the Sigma(2) rows, the first sweep's
 * correction D1, estimates of hbar*omega, the interaction scale and Nmax,
 * and the channel's l, j, charge -- to the remaining step D - D1 of the
 * converged correction D (rho in the channel's HF orbital basis minus the HF
 * occupations, upper triangle, times 10, in the first 21 of BCD_LEARN_OUT
 * outputs), and the second sweep starts from rho = HF occupations + D1 +
 * that step.  m must have n_in = BCD_LEARN_IN, n_out = BCD_LEARN_OUT and
 * live on the batch's device; NULL switches the learned start off.  The
 * batch keeps the pointer. */
#define BCD_LEARN_IN 80
#define BCD_LEARN_OUT 32
#define BCD_LEARN_NI 6
int bcd_batch_set_learned_start(bcd_batch *b, bcd_mlp *m);
/* Fallback to the plain iteration after a seeded (learned start) or
 * density-mixed (accel > 0) sc0 stage: one plain sweep from every result
 * measures its fixed-point residual (the stage's conv_kind measure); an
 * instance whose residual exceeds tol, or that did not converge, is solved
 * again by the plain iteration from its HF propagator.  Default 1e-6; tol <= 0
 * switches the check off. */
int bcd_batch_set_fallback_tol(bcd_batch *b, double tol);
/* Record the network's per-channel inputs and targets (D of the converged
 * propagator; the network is trained on D - D1) in the following solves. */
int bcd_batch_record_learning(bcd_batch *b, int on);
int bcd_batch_num_channels(const bcd_batch *b);
/* Inputs [n_chan x BCD_LEARN_IN], targets [n_chan x BCD_LEARN_OUT] and the
 * instance of every channel (any may be NULL) of the last recorded solve;
 * n_chan must equal bcd_batch_num_channels.  Channels with more than
 * BCD_LEARN_NI orbits have all-zero rows. */
int bcd_batch_learning_data(bcd_batch *b, float *feat, float *target, int *chan_inst, int n_chan);
/* How the last solve of instance i started and ended: seeded by the network,
 * redone by the plain iteration, sweeps of the seeded or mixed stage before
 * the residual check, and that residual (-1 when not checked). */
int bcd_batch_start_info(bcd_batch *b, int inst, int *seeded, int *fell_back, int *first_its, double *residual);

/* ------------------------------------------ two-body propagator matrices
 * The pp/hh ladder of gII_prop_t (gII_b.cpp:62-563) and the ph ring of
 * Pol_prop_t (PolProp_b.cpp:43-614) built from the fragments of g, as
 * CoupledCluster_t uses them (CCD_a.cpp:158-240).  Fragment indices count
 * quasiparticles (quasiholes) over all subshells in storage order, as
 * SpProp_t's n_qp (n_qh) arrays do.  A channel is (J, dT, pif): total angular
 * momentum, total charge (ladder) or charge difference (ring), parity.
 * n1p / n1h are SpProp_t's per-fragment labels for the selection
 * (icut, ntcut); they may be NULL when icut == 0.
 * Matrices are stored as the reference stores them: entry (c, r) at
 * Sab[c * ndim_vec_alloc + r]; a call writes exactly the extent the
 * reference writes ([0, (ndim-1)*ndim_vec_alloc + ndim)).  free_poles (may be
 * NULL) receives the entries Fill_*_mtx writes.  i_RPAcalc: 0 (D)TDA of both
 * blocks, 1 first block only, 2 second block only (ladder), 3 (D)RPA, 4 free
 * propagator.  Returns 0, -100 (invalid i_RPAcalc), -101 (ndim_vec_alloc < 1),
 * -102 (ndim > ndim_vec_alloc, where the reference stops the program) or a
 * BCD_ERR_ code.  Every call is independent: the reference keeps the
 * free-propagator switch of i_RPAcalc = 4 in a file-scope static that stays
 * set for the rest of the process; this library does not. */

/* gII_prop_t::Count_LaddRPA_basis (gII_b.cpp:62-113) */
int bcd_ladd_count(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut,
                   const int *n1p, const int *n1h, int *npp, int *nhh);
/* gII_prop_t::Build_LaddRPA_basis (gII_b.cpp:115-200): 2qp pairs na <= nb,
 * then 2qh pairs ka <= kb, with e_pp = ep[na] + ep[nb], e_hh = eh[ka] + eh[kb]. */
int bcd_ladd_basis(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut,
                   const int *n1p, const int *n1h, int cap_pp, int cap_hh, int *n1_pp, int *n2_pp,
                   double *e_pp, int *k1_hh, int *k2_hh, double *e_hh, int *npp, int *nhh);
/* gII_prop_t::Fill_LaddRPA_mtx (gII_b.cpp:288-563) */
int bcd_ladd_fill(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int npp, const int *n1_pp,
                  const int *n2_pp, const double *e_pp, int nhh, const int *k1_hh, const int *k2_hh,
                  const double *e_hh, int i_RPAcalc, double *Sab, int ndim_vec_alloc,
                  double *free_poles, int add_free_poles);
/* Pol_prop_t::Count_phRPA_basis (PolProp_b.cpp:43-94) */
int bcd_ph_count(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut,
                 const int *n1p, const int *n1h, int *nph, int *nhp);
/* Pol_prop_t::Build_phRPA_basis (PolProp_b.cpp:96-175): every (qp na, qh kb),
 * ph when ch(a) - ch(b) == dT, hp when ch(b) - ch(a) == dT, e = ep[na] - eh[kb]. */
int bcd_ph_basis(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int icut, int ntcut,
                 const int *n1p, const int *n1h, int cap_ph, int cap_hp, int *n_ph, int *k_ph,
                 double *e_ph, int *n_hp, int *k_hp, double *e_hp, int *nph, int *nhp);
/* Pol_prop_t::Fill_phRPA_mtx (PolProp_b.cpp:255-563) */
int bcd_ph_fill(bcd_ctx *ctx, const bcd_prop *g, int J, int dT, int pif, int nph, const int *n_ph,
                const int *k_ph, const double *e_ph, int nhp, const int *n_hp, const int *k_hp,
                const double *e_hp, int i_RPAcalc, double *Sab, int ndim_vec_alloc,
                double *free_poles, int add_free_poles);

/* Every channel at once (the loops of CoupledCluster_t, CCD_a.cpp:158-240):
 * for J = 0..Jmax and both parities, the ladder of each total charge and the
 * ring of each charge difference whose basis is not empty, with i_RPAcalc
 * ladd_mode / ring_mode (-1: leave that family out).  The fragments and all
 * matrices stay resident on the device; bcd_t2_build recomputes them. */
typedef struct bcd_t2 bcd_t2;
int bcd_t2_create(bcd_ctx *ctx, const bcd_prop *g, int ladd_mode, int ring_mode, int add_free_poles,
                  bcd_t2 **out);
void bcd_t2_destroy(bcd_t2 *t);
int bcd_t2_count(const bcd_t2 *t);
/* info[10] = {kind (0 ladder, 1 ring), J, pif, dT, i_RPAcalc, n_f, n_b, ld, n_a, n_b_basis}:
 * the matrix has dimension n_f + n_b and leading dimension ld; the basis has
 * n_a entries in its first block (2qp / ph) and n_b_basis in its second. */
int bcd_t2_info(const bcd_t2 *t, int k, int *info);
/* basis of matrix k: ladder n1_pp n2_pp e_pp k1_hh k2_hh e_hh, ring n_ph k_ph
 * e_ph n_hp k_hp e_hp (NULL to skip an array) */
int bcd_t2_basis(const bcd_t2 *t, int k, int *a1, int *a2, double *ea, int *b1, int *b2, double *eb);
int bcd_t2_build(bcd_t2 *t);
int bcd_t2_matrix(const bcd_t2 *t, int k, double *Sab, double *free_poles);
/* 1: time each kernel family of bcd_t2_build with hipEvents (resets the sums) */
int bcd_t2_set_profile(bcd_t2 *t, int on);
/* k = 0 "t2_zero", 1 "t2_amp", 2 "t2_pandya", 3 "t2_gemm1", 4 "t2_gemm2", 5 "t2_diag" */
int bcd_t2_kernel_times(const bcd_t2 *t, int k, double *avg_ms, long *launches);
/* work[0] FLOPs of the G = T K products, [1] of the G T^T products (2 M N K
 * each, unpadded), [2] bytes of matrix entries written by the products; n >= 3 */
int bcd_t2_work(const bcd_t2 *t, double *work, int n);

#ifdef __cplusplus
}
#endif
#endif
