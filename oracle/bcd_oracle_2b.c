/*
 * bcd_oracle_2b.c -- CPU FP64 restatement of BoccaDorata's two-body
 * propagator matrices: the pp/hh ladder (gII_prop_t) and the ph ring
 * (Pol_prop_t) bases and (D)RPA/(D)TDA matrices, with the Pandya transform
 * and the 6j symbol they use.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE (see bcd_oracle.h): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Loop orders and floating-point operation orders follow the reference so
 * that results agree with the compiled reference to a few ulps.  Parity is
 * pinned in tests/test_oracle_2b.py against golden vectors of the compiled
 * reference and, for orc_s6j, against exact rational arithmetic.
 */
#include <math.h>
#include <stdlib.h>

#include "bcd_oracle.h"

#define SQR2 1.414213562373095048801688

/* ------------------------------------------------------------------ 6j */

/* Pascal's triangle, C[n][k] for 0 <= k <= n < 60 (Ang_momenta.cpp:74-80). */
#define NBIN 60
static double binom[NBIN][NBIN];
static int binom_ready = 0;

static void binom_init(void) {
  int n, k;
  if (binom_ready) return;
  for (n = 0; n < NBIN; ++n) { binom[n][0] = 1.0; binom[n][n] = 1.0; }
  for (n = 1; n < NBIN - 1; ++n)
    for (k = 1; k <= n; ++k) binom[n + 1][k] = binom[n][k - 1] + binom[n][k];
  binom_ready = 1;
}

/* 0 when 2j1, 2j2, 2j3 close a triangle with integer perimeter
 * (ang_mom_functions::TriIneq, Ang_momenta.hh:40-44). */
static int tri_fail(int a, int b, int c) {
  int s1 = a + b - c, s2 = b + c - a, s3 = c + a - b;
  return !(s1 >= 0 && s2 >= 0 && s3 >= 0 && !(s1 & 1) && !(s2 & 1) && !(s3 & 1));
}

/* {ja/2 jb/2 jc/2; la/2 lb/2 lc/2} in the binomial form of
 * Ang_momenta.cpp:150-182: a normalisation built from eight binomials and an
 * alternating sum of products of four binomials, accumulated as x = -x + t
 * from the lowest summation index up. */
double orc_s6j(int ja, int jb, int jc, int la, int lb, int lc) {
  int mt, ma, mb, mc, na, nb, nc, ka, kb, kc, lo, hi, l;
  double norm, x;
  binom_init();
  if (tri_fail(ja, jb, jc) || tri_fail(ja, lb, lc) || tri_fail(la, jb, lc) || tri_fail(la, lb, jc))
    return 0.0;
  mt = (ja + jb + jc) / 2 + 1;
  ma = (ja + lb + lc) / 2 + 1;
  mb = (la + jb + lc) / 2 + 1;
  mc = (la + lb + jc) / 2 + 1;
  na = mt - ja; nb = mt - jb; nc = mt - jc;
  ka = ma - lc; kb = mb - lc; kc = mc - jc;
  if (mt >= NBIN || ma >= NBIN || mb >= NBIN || mc >= NBIN) return NAN;
  norm = binom[mt][ja + 1] * binom[ja][nc - 1] /
         (binom[ma][ja + 1] * binom[ja][ka - 1] * binom[mb][la + 1] * binom[la][kb - 1] *
          binom[mc][la + 1] * binom[la][kc - 1]);
  norm = sqrt(norm) / (la + 1.0);
  lo = mt;
  if (ma > lo) lo = ma;
  if (mb > lo) lo = mb;
  if (mc > lo) lo = mc;
  hi = ma + na;
  if (mb + nb < hi) hi = mb + nb;
  if (mc + nc < hi) hi = mc + nc;
  if (hi >= NBIN) return NAN;
  x = 0.0;
  for (l = lo; l < hi; ++l)
    x = -x + binom[l][mt] * binom[na - 1][l - ma] * binom[nb - 1][l - mb] * binom[nc - 1][l - mc];
  return (double)(1 - 2 * (hi % 2)) * norm * x;
}

/* ------------------------------------------------------------- Pandya */

/* Particle-hole matrix element from the pp ones (Pandya, PolProp_b.cpp:569-614):
 * -sum_J' (2J'+1) <a d; J'|V|g b; J'> {ja jb Jf; jg jd J'}, times sqrt2 for
 * a == d and for g == b; 0 when the parities do not close. */
double orc_pandya(const orc_sys *s, int ia, int ib, int ig, int id, int Jf) {
  const int sa = s->orb_sh[ia], sb = s->orb_sh[ib], sg = s->orb_sh[ig], sd = s->orb_sh[id];
  int j2a, j2b, j2g, j2d, lo, hi, t1, t2, JJ2;
  double vph = 0.0;
  if ((s->ip[sa] + s->ip[sb] + s->ip[sg] + s->ip[sd]) % 2) return 0.0;
  j2a = s->j2[sa]; j2b = s->j2[sb]; j2g = s->j2[sg]; j2d = s->j2[sd];
  t1 = abs(j2a - j2d); t2 = abs(j2b - j2g);
  lo = t1 > t2 ? t1 : t2;
  t1 = j2a + j2d; t2 = j2b + j2g;
  hi = t1 < t2 ? t1 : t2;
  for (JJ2 = lo; JJ2 <= hi; JJ2 += 2)
    vph -= (double)(JJ2 + 1) * orc_vpp_get(s, ia, id, ig, ib, JJ2 / 2) *
           orc_s6j(j2a, j2b, 2 * Jf, j2g, j2d, JJ2);
  if (ia == id) vph *= SQR2;
  if (ig == ib) vph *= SQR2;
  return vph;
}

/* ------------------------------------------------------- ladder (pp/hh) */

/* Selection of one pair of fragments u <= w of one kind (both qp or both qh)
 * for channel (Jf, dT, pif) (gII_b.cpp:75-97). */
static int ladd_pair_ok(const orc_sys *s, int su, int sw, int same, int Jf, int dT, int pif) {
  if ((pif + s->ip[su] + s->ip[sw]) % 2) return 0;
  if (tri_fail(s->j2[su], s->j2[sw], 2 * Jf)) return 0;
  if (same && (abs((s->j2[su] + s->j2[sw]) / 2 - Jf + 1) % 2)) return 0;
  return s->ch[su] + s->ch[sw] == dT;
}

/* icut == 0 keeps every pair; icut > 0 keeps (n1(u)+n1(w))*icut >= ntcut;
 * icut < 0 keeps (n1(u)+n1(w))*icut >= 1-ntcut (gII_b.cpp:67-70). */
static int cut_drop(int icut, int ntcut, int n1u, int n1w) {
  if (icut == 0) return 0;
  if (icut < 0) ntcut = 1 - ntcut;
  return (n1u + n1w) * icut < ntcut;
}

/* gII_prop_t::Count_LaddRPA_basis / Build_LaddRPA_basis (gII_b.cpp:62-200):
 * 2qp pairs na <= nb (nb inner) then 2qh pairs ka <= kb, with e = e_u + e_w.
 * With NULL outputs only the counts are produced.  n1p / n1h label the
 * fragments for the cut (ignored when icut == 0).  Returns npp + nhh, or -1
 * when a capacity is too small. */
int orc_ladd_basis(const orc_sys *s, const orc_prop *g, int Jf, int dT, int pif, int icut, int ntcut,
                   const int *n1p, const int *n1h, int cap_pp, int cap_hh, int *n1_pp, int *n2_pp,
                   double *e_pp, int *k1_hh, int *k2_hh, double *e_hh, int *npp, int *nhh) {
  int u, w, cp = 0, ch = 0;
  for (u = 0; u < g->n_qp; ++u)
    for (w = u; w < g->n_qp; ++w) {
      if (icut && cut_drop(icut, ntcut, n1p[u], n1p[w])) continue;
      if (!ladd_pair_ok(s, g->clj_p[u], g->clj_p[w], u == w, Jf, dT, pif)) continue;
      if (n1_pp) {
        if (cp >= cap_pp) return -1;
        e_pp[cp] = g->ep[u] + g->ep[w]; n1_pp[cp] = u; n2_pp[cp] = w;
      }
      ++cp;
    }
  for (u = 0; u < g->n_qh; ++u)
    for (w = u; w < g->n_qh; ++w) {
      if (icut && cut_drop(icut, ntcut, n1h[u], n1h[w])) continue;
      if (!ladd_pair_ok(s, g->clj_h[u], g->clj_h[w], u == w, Jf, dT, pif)) continue;
      if (k1_hh) {
        if (ch >= cap_hh) return -1;
        e_hh[ch] = g->eh[u] + g->eh[w]; k1_hh[ch] = u; k2_hh[ch] = w;
      }
      ++ch;
    }
  *npp = cp; *nhh = ch;
  return cp + ch;
}

/* sum over the orbits of fragments (r1, r2) and (c1, c2) of
 * x_r V_J x_c, with the sqrt2 of equal orbits inside and the 1/sqrt2 of equal
 * fragments outside (the inner loops of gII_b.cpp:366-395). */
static double ladd_me(const orc_sys *s, int Jf, const double *fr1, const double *fr2, int sr1, int sr2,
                      int r1, int r2, const double *fc1, const double *fc2, int sc1, int sc2, int c1,
                      int c2) {
  const int a0 = s->orb0[sr1], a1 = a0 + s->no[sr1], b0 = s->orb0[sr2], b1 = b0 + s->no[sr2];
  const int g0 = s->orb0[sc1], g1 = g0 + s->no[sc1], d0 = s->orb0[sc2], d1 = d0 + s->no[sc2];
  int ia, ib, ig, id;
  double x1 = 0.0;
  for (ia = a0; ia < a1; ++ia)
    for (ib = b0; ib < b1; ++ib) {
      double xr = fr1[ia - a0] * fr2[ib - b0];
      if (ia == ib) xr *= SQR2;
      for (ig = g0; ig < g1; ++ig)
        for (id = d0; id < d1; ++id) {
          double xc = fc1[ig - g0] * fc2[id - d0];
          if (ig == id) xc *= SQR2;
          x1 += xr * orc_vpp_get(s, ia, ib, ig, id, Jf) * xc;
        }
    }
  if (r1 == r2) x1 /= SQR2;
  if (c1 == c2) x1 /= SQR2;
  return x1;
}

/* gII_prop_t::Fill_LaddRPA_mtx (gII_b.cpp:288-563).  Sab holds the matrix
 * with entry (c, r) at Sab[c*ld + r]; ndimf = npp (0 for mode 2), ndimb = nhh
 * (0 for mode 1).  Modes: 0 pp and hh TDA blocks, 1 pp only, 2 hh only,
 * 3 full RPA (A, B, D), 4 free poles only (energies on the diagonal).  The
 * A block is +V, the B blocks -V (pp rows) and +V (hh rows), the D block -V;
 * add_free_poles adds e_pp / e_hh on the diagonal of the blocks built.
 * free_poles (may be NULL) receives e_pp then e_hh of the blocks built
 * (nothing for mode 4).  Every call is independent: the reference keeps the
 * "free propagator" switch of mode 4 in a file-scope static that stays set
 * for later calls of the process; this restatement does not.
 * Returns 0, -100 (bad mode), -101 (ld < 1) or -102 (ndim > ld). */
int orc_ladd_fill(const orc_sys *s, const orc_prop *g, int Jf, int npp, const int *n1_pp, const int *n2_pp,
                  const double *e_pp, int nhh, const int *k1_hh, const int *k2_hh, const double *e_hh,
                  int mode, double *Sab, int ld, double *free_poles, int add_free_poles) {
  const int nspb = g->nspb;
  int ndimf, ndimb, ntot, nc, nr, make_pp = 1, make_hh = 1, make_b = 1, free_only = 0;
  long k, kend;
  if (mode < 0 || mode > 4) return -100;
  ndimf = npp > 0 ? npp : 0;
  ndimb = nhh > 0 ? nhh : 0;
  if (ld < 1) return -101;
  if (mode == 0) make_b = 0;
  if (mode == 1) { ndimb = 0; make_hh = 0; make_b = 0; }
  if (mode == 2) { ndimf = 0; make_pp = 0; make_b = 0; }
  if (mode == 4) { free_only = 1; add_free_poles = 1; }
  ntot = ndimf + ndimb;
  if (ntot > ld) return -102;
  kend = (long)(ntot - 1) * ld + ntot;
  for (k = 0; k < kend; ++k) Sab[k] = 0.0;

  if (!free_only && (make_pp || make_b)) {
    for (nc = 0; nc < npp; ++nc) {
      const int ng = n1_pp[nc], nd = n2_pp[nc];
      const int sg = g->clj_p[ng], sd = g->clj_p[nd];
      const double *fg = g->fqp + (size_t)ng * nspb, *fd = g->fqp + (size_t)nd * nspb;
      for (nr = 0; nr < npp; ++nr) {
        const int na = n1_pp[nr], nb = n2_pp[nr];
        Sab[(long)nc * ld + nr] =
            ladd_me(s, Jf, g->fqp + (size_t)na * nspb, g->fqp + (size_t)nb * nspb, g->clj_p[na], g->clj_p[nb],
                    na, nb, fg, fd, sg, sd, ng, nd);
      }
      if (free_poles) free_poles[nc] = e_pp[nc];
      if (!make_b) continue;
      for (nr = 0; nr < nhh; ++nr) {
        const int na = k1_hh[nr], nb = k2_hh[nr];
        const double x1 = ladd_me(s, Jf, g->fqh + (size_t)na * nspb, g->fqh + (size_t)nb * nspb, g->clj_h[na],
                                  g->clj_h[nb], na, nb, fg, fd, sg, sd, ng, nd);
        Sab[(long)nc * ld + ndimf + nr] = -x1;
        Sab[(long)(ndimf + nr) * ld + nc] = x1;
      }
    }
  }
  if (!free_only && make_hh) {
    for (nc = 0; nc < nhh; ++nc) {
      const int ng = k1_hh[nc], nd = k2_hh[nc];
      const int sg = g->clj_h[ng], sd = g->clj_h[nd];
      const double *fg = g->fqh + (size_t)ng * nspb, *fd = g->fqh + (size_t)nd * nspb;
      for (nr = 0; nr < nhh; ++nr) {
        const int na = k1_hh[nr], nb = k2_hh[nr];
        Sab[(long)(ndimf + nc) * ld + ndimf + nr] =
            -ladd_me(s, Jf, g->fqh + (size_t)na * nspb, g->fqh + (size_t)nb * nspb, g->clj_h[na],
                     g->clj_h[nb], na, nb, fg, fd, sg, sd, ng, nd);
      }
      if (free_poles) free_poles[nc + ndimf] = e_hh[nc];
    }
  }
  if (add_free_poles) {
    if (make_pp)
      for (nc = 0; nc < npp; ++nc) Sab[(long)nc * (1 + ld)] += e_pp[nc];
    if (make_hh)
      for (nc = 0; nc < nhh; ++nc) Sab[(long)(nc + ndimf) * (1 + ld)] += e_hh[nc];
  }
  return 0;
}

/* --------------------------------------------------------- ring (ph/hp) */

/* Pol_prop_t::Count_phRPA_basis / Build_phRPA_basis (PolProp_b.cpp:43-175):
 * every (qp na, qh kb) with kb inner; ch(a) - ch(b) == dT gives a ph entry,
 * ch(b) - ch(a) == dT an hp entry (both when dT == 0), each with
 * e = e_p(na) - e_h(kb).  Returns nph + nhp, or -1 when a capacity is too
 * small. */
int orc_ph_basis(const orc_sys *s, const orc_prop *g, int Jf, int dT, int pif, int icut, int ntcut,
                 const int *n1p, const int *n1h, int cap_ph, int cap_hp, int *n_ph, int *k_ph, double *e_ph,
                 int *n_hp, int *k_hp, double *e_hp, int *nph, int *nhp) {
  int u, w, cf = 0, cb = 0;
  for (u = 0; u < g->n_qp; ++u)
    for (w = 0; w < g->n_qh; ++w) {
      const int sa = g->clj_p[u], sb = g->clj_h[w];
      if (icut && cut_drop(icut, ntcut, n1p[u], n1h[w])) continue;
      if ((pif + s->ip[sa] + s->ip[sb]) % 2) continue;
      if (tri_fail(s->j2[sa], s->j2[sb], 2 * Jf)) continue;
      if (s->ch[sa] - s->ch[sb] == dT) {
        if (n_ph) {
          if (cf >= cap_ph) return -1;
          e_ph[cf] = g->ep[u] - g->eh[w]; n_ph[cf] = u; k_ph[cf] = w;
        }
        ++cf;
      }
      if (s->ch[sb] - s->ch[sa] == dT) {
        if (n_hp) {
          if (cb >= cap_hp) return -1;
          e_hp[cb] = g->ep[u] - g->eh[w]; n_hp[cb] = u; k_hp[cb] = w;
        }
        ++cb;
      }
    }
  *nph = cf; *nhp = cb;
  return cf + cb;
}

/* sum over the orbits of (qp r1, qh r2) and (qp c1, qh c2) of
 * x_r P(a b g d) x_c; with swap the column pair enters as (d, g) and the
 * x3 phase of the column (the x2 sums of PolProp_b.cpp:348-368, 437-455). */
static double ring_me(const orc_sys *s, int Jf, const double *fa, const double *fb, int sa, int sb,
                      const double *fg, const double *fd, int sg, int sd, int swap) {
  const int a0 = s->orb0[sa], a1 = a0 + s->no[sa], b0 = s->orb0[sb], b1 = b0 + s->no[sb];
  const int g0 = s->orb0[sg], g1 = g0 + s->no[sg], d0 = s->orb0[sd], d1 = d0 + s->no[sd];
  const double x3 = ((abs(s->j2[sg] - s->j2[sd]) / 2) % 2) ? -1.0 : 1.0;
  int ia, ib, ig, id;
  double x = 0.0;
  for (ia = a0; ia < a1; ++ia)
    for (ib = b0; ib < b1; ++ib) {
      const double xr = fa[ia - a0] * fb[ib - b0];
      for (ig = g0; ig < g1; ++ig)
        for (id = d0; id < d1; ++id) {
          if (swap)
            x += xr * orc_pandya(s, ia, ib, id, ig, Jf) * x3 * fg[ig - g0] * fd[id - d0];
          else
            x += xr * orc_pandya(s, ia, ib, ig, id, Jf) * fg[ig - g0] * fd[id - d0];
        }
    }
  return x;
}

/* Pol_prop_t::Fill_phRPA_mtx (PolProp_b.cpp:255-563), same storage as
 * orc_ladd_fill; n_f = nph, n_b = nhp (0 for mode 1).  Modes: 0 TDA of both
 * blocks, 1 TDA of the ph block only, 3 full RPA, 4 free poles only (mode 2
 * is invalid).  For dT == 0 the hp block is the ph block copied with the
 * opposite sign and the B blocks (mode 3) are the exchange sums x2; for
 * dT != 0 the B and hp blocks come from the hp basis.  Energies: +e_ph on
 * the ph diagonal, -e_ph (dT == 0) or -e_hp (dT != 0) on the hp diagonal.
 * As in orc_ladd_fill, mode 4 does not persist between calls.
 * Returns 0, -100, -101 or -102 as orc_ladd_fill. */
int orc_ph_fill(const orc_sys *s, const orc_prop *g, int Jf, int dT, int nph, const int *n_ph, const int *k_ph,
                const double *e_ph, int nhp, const int *n_hp, const int *k_hp, const double *e_hp, int mode,
                double *Sab, int ld, double *free_poles, int add_free_poles) {
  const int nspb = g->nspb;
  int nf, nb, ntot, nc, nr, make_b = 1, direct = 1, copied = 0, free_only = 0, t0;
  long k, kend;
  if (mode < 0 || mode > 4 || mode == 2) return -100;
  nf = nph > 0 ? nph : 0;
  nb = nhp > 0 ? nhp : 0;
  if (mode == 1) nb = 0;
  ntot = nf + nb;
  if (ntot > ld) return -102;
  if (ld < 1) return -101;
  if (dT == 0) { direct = 0; copied = 1; }
  if (mode == 1) { direct = 0; copied = 0; }
  if (mode != 3) make_b = 0;
  if (mode == 4) { free_only = 1; add_free_poles = 1; }
  t0 = copied && make_b;
  kend = (long)(ntot - 1) * ld + ntot;
  for (k = 0; k < kend; ++k) Sab[k] = 0.0;

  if (!free_only) {
    for (nc = 0; nc < nph; ++nc) {
      const int ng = n_ph[nc], kd = k_ph[nc];
      const int sg = g->clj_p[ng], sd = g->clj_h[kd];
      const double *fg = g->fqp + (size_t)ng * nspb, *fd = g->fqh + (size_t)kd * nspb;
      for (nr = 0; nr < nph; ++nr) {
        const int na = n_ph[nr], kb = k_ph[nr];
        const double *fa = g->fqp + (size_t)na * nspb, *fb = g->fqh + (size_t)kb * nspb;
        const double x1 = ring_me(s, Jf, fa, fb, g->clj_p[na], g->clj_h[kb], fg, fd, sg, sd, 0);
        const double x2 = t0 ? ring_me(s, Jf, fa, fb, g->clj_p[na], g->clj_h[kb], fg, fd, sg, sd, 1) : 0.0;
        if (copied) {
          Sab[(long)(nf + nc) * ld + nf + nr] = -x1;
          Sab[(long)nc * ld + nf + nr] = -x2;
          Sab[(long)(nf + nc) * ld + nr] = x2;
        }
        Sab[(long)nc * ld + nr] = x1;
      }
      if (free_poles) {
        free_poles[nc] = e_ph[nc];
        if (copied) free_poles[nf + nc] = e_ph[nc];
      }
    }
    if (make_b && direct) {
      for (nc = 0; nc < nhp; ++nc) {
        const int ng = n_hp[nc], kd = k_hp[nc];
        const int sg = g->clj_p[ng], sd = g->clj_h[kd];
        const double *fg = g->fqp + (size_t)ng * nspb, *fd = g->fqh + (size_t)kd * nspb;
        for (nr = 0; nr < nph; ++nr) {
          const int na = n_ph[nr], kb = k_ph[nr];
          const double x2 = ring_me(s, Jf, g->fqp + (size_t)na * nspb, g->fqh + (size_t)kb * nspb,
                                    g->clj_p[na], g->clj_h[kb], fg, fd, sg, sd, 1);
          Sab[(long)(nf + nc) * ld + nr] = x2;
          Sab[(long)nr * ld + nf + nc] = -x2;
        }
      }
    }
    if (direct) {
      for (nc = 0; nc < nhp; ++nc) {
        const int ng = n_hp[nc], kd = k_hp[nc];
        const int sg = g->clj_p[ng], sd = g->clj_h[kd];
        const double *fg = g->fqp + (size_t)ng * nspb, *fd = g->fqh + (size_t)kd * nspb;
        for (nr = 0; nr < nhp; ++nr) {
          const int na = n_hp[nr], kb = k_hp[nr];
          Sab[(long)(nc + nf) * ld + nf + nr] =
              -ring_me(s, Jf, g->fqp + (size_t)na * nspb, g->fqh + (size_t)kb * nspb, g->clj_p[na],
                       g->clj_h[kb], fg, fd, sg, sd, 0);
        }
        if (free_poles) free_poles[nc + nf] = e_hp[nc];
      }
    }
  }
  if (add_free_poles) {
    for (nc = 0; nc < nph; ++nc) {
      Sab[(long)nc * (1 + ld)] += e_ph[nc];
      if (dT == 0 && mode != 1) Sab[(long)(nf + nc) * (1 + ld)] -= e_ph[nc];
    }
    if (direct)
      for (nc = 0; nc < nhp; ++nc) Sab[(long)(nc + nf) * (ld + 1)] -= e_hp[nc];
  }
  return 0;
}
