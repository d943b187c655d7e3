/*
 * bcd_oracle.c -- CPU FP64 restatement of BoccaDorata's second-order SCGF hot
 * path.  TEST INFRASTRUCTURE ONLY (the checker); see bcd_oracle.h for which
 * test pins which function.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off,
 * no FMA, so operation order matches the reference's x86-64 build).
 */
#include "bcd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SQR2 1.4142135623730950488016887
#define KEY_AB 1024
#define KEY_J 128
#define CH_OFF (-2)
#define CH_SLOTS 5

/* ---------------------------------------------------------------- helpers */

static int *icopy(const int *src, int n) {
  int *d = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  if (src && n > 0) memcpy(d, src, sizeof(int) * n);
  return d;
}

/* Triangle rule on doubled angular momenta; ang_mom_functions::TriIneq
 * (Ang_momenta.hh:42-45) returns nonzero when any of the three sums is
 * negative or odd.  Here: 1 when the triangle is allowed. */
static int tri_ok(int a2, int b2, int c2) {
  int s1 = a2 + b2 - c2, s2 = b2 + c2 - a2, s3 = c2 + a2 - b2;
  return s1 >= 0 && s2 >= 0 && s3 >= 0 && !(s1 & 1) && !(s2 & 1) && !(s3 & 1);
}

/* Pair and channel keys of VppInt_t (Vpp_potential.cpp:74-86). */
static int key_pair(int i1, int i2, int J) { return (i1 * KEY_AB + i2) * KEY_J + J; }
static int key_chan(int ch, int ip, int J) { return ((ch - CH_OFF) * 2 + ip) * KEY_J + J; }

/* ------------------------------------------------------ model space + V */

typedef struct { int chipJ, ab, cd; double v; } me_rec;

static int me_cmp(const void *x, const void *y) {
  const me_rec *a = (const me_rec *)x, *b = (const me_rec *)y;
  if (a->chipJ != b->chipJ) return a->chipJ < b->chipJ ? -1 : 1;
  if (a->ab != b->ab) return a->ab < b->ab ? -1 : 1;
  if (a->cd != b->cd) return a->cd < b->cd ? -1 : 1;
  return 0;
}

orc_sys *orc_sys_create(int nsubsh, const int *l, const int *j2, const int *ip,
                        const int *ch, const int *no, const int *orb_n,
                        const int *orb_F, const double *orb_e,
                        long n_raw, const int *ra, const int *rb, const int *rc,
                        const int *rd, const int *rJ, const double *rv) {
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  int i, k, norb = 0;
  s->nsubsh = nsubsh;
  s->l = icopy(l, nsubsh); s->j2 = icopy(j2, nsubsh); s->ip = icopy(ip, nsubsh);
  s->ch = icopy(ch, nsubsh); s->no = icopy(no, nsubsh);
  s->orb0 = icopy(NULL, nsubsh); s->nqp = icopy(NULL, nsubsh); s->nqh = icopy(NULL, nsubsh);
  s->nspb = 0; s->Jmax = 0;
  for (i = 0; i < nsubsh; ++i) {
    s->orb0[i] = norb; norb += no[i];
    if (no[i] > s->nspb) s->nspb = no[i];
    if (j2[i] > s->Jmax) s->Jmax = j2[i];
  }
  s->norb = norb;
  s->orb_sh = icopy(NULL, norb); s->orb_n = icopy(orb_n, norb); s->orb_F = icopy(orb_F, norb);
  s->orb_e = (double *)malloc(sizeof(double) * (norb > 0 ? norb : 1));
  memcpy(s->orb_e, orb_e, sizeof(double) * norb);
  for (i = 0; i < nsubsh; ++i) {
    int nh = 0;
    for (k = 0; k < no[i]; ++k) { s->orb_sh[s->orb0[i] + k] = i; nh += (orb_F[s->orb0[i] + k] == 1); }
    s->nqh[i] = nh; s->nqp[i] = no[i] - nh;
  }

  /* read_ascii normalisation (Vpp_potential.cpp:306-360): order each pair,
   * collect the exchange phase, drop Pauli-forbidden pairs, order bra/ket. */
  me_rec *rec = (me_rec *)malloc(sizeof(me_rec) * (n_raw > 0 ? n_raw : 1));
  long n = 0, q;
  for (q = 0; q < n_raw; ++q) {
    int a = ra[q], b = rb[q], c = rc[q], d = rd[q], J = rJ[q], ph = 0, ab, cd;
    if (a < 0 || b < 0 || c < 0 || d < 0 || a >= norb || b >= norb || c >= norb || d >= norb) continue;
    if ((a == b && abs(j2[s->orb_sh[a]] - J) % 2 == 0) || (c == d && abs(j2[s->orb_sh[c]] - J) % 2 == 0)) continue;
    if (a <= b) ab = key_pair(a, b, J);
    else { ab = key_pair(b, a, J); ph += (j2[s->orb_sh[a]] + j2[s->orb_sh[b]]) / 2 - J + 1; }
    if (c <= d) cd = key_pair(c, d, J);
    else { cd = key_pair(d, c, J); ph += (j2[s->orb_sh[c]] + j2[s->orb_sh[d]]) / 2 - J + 1; }
    if (ab > cd) { int t = ab; ab = cd; cd = t; }
    {
      int sa = s->orb_sh[a], sb = s->orb_sh[b];
      int chp = ch[sa] + ch[sb], ipp = (ip[sa] + ip[sb]) % 2;
      rec[n].chipJ = key_chan(chp, ipp, J);
    }
    rec[n].ab = ab; rec[n].cd = cd;
    rec[n].v = (abs(ph) % 2 != 0) ? -rv[q] : rv[q];
    ++n;
  }
  /* prepare_for_reading (Vpp_potential.cpp:440-501): order by (ch,ip,J)
   * channel, then by the pair keys; record each channel's range. */
  qsort(rec, n, sizeof(me_rec), me_cmp);
  s->n_me = n;
  s->iabJ = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  s->icdJ = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  s->v = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  s->chipJ_str = (int *)malloc(sizeof(int) * CH_SLOTS * KEY_J * 2);
  s->chipJ_end = (int *)malloc(sizeof(int) * CH_SLOTS * KEY_J * 2);
  for (i = 0; i < CH_SLOTS * KEY_J * 2; ++i) { s->chipJ_str[i] = -10; s->chipJ_end[i] = -10; }
  for (q = 0; q < n; ++q) {
    s->iabJ[q] = rec[q].ab; s->icdJ[q] = rec[q].cd; s->v[q] = rec[q].v;
    if (q == 0 || rec[q].chipJ != rec[q - 1].chipJ) s->chipJ_str[rec[q].chipJ] = (int)q;
    s->chipJ_end[rec[q].chipJ] = (int)q;
  }
  free(rec);
  return s;
}

void orc_sys_free(orc_sys *s) {
  if (!s) return;
  free(s->l); free(s->j2); free(s->ip); free(s->ch); free(s->no); free(s->orb0);
  free(s->nqp); free(s->nqh); free(s->orb_sh); free(s->orb_n); free(s->orb_F); free(s->orb_e);
  free(s->iabJ); free(s->icdJ); free(s->v); free(s->chipJ_str); free(s->chipJ_end);
  free(s);
}

long orc_sys_n_me(const orc_sys *s) { return s->n_me; }

/* VppInt_t::get (Vpp_potential.cpp:557-607): <ab;J|V|cd;J> with the exchange
 * phase (-1)^(ja+jb-J+1) for each pair stored in swapped order; 0 when the
 * element is not stored. */
double orc_vpp_get(const orc_sys *s, int ia, int ib, int ic, int id, int J) {
  int sa = s->orb_sh[ia], sb = s->orb_sh[ib];
  int ipp = (s->ip[sa] + s->ip[sb]) % 2, chp = s->ch[sa] + s->ch[sb];
  int ph = 0, ab, cd, lo, hi;
  if (ia <= ib) ab = key_pair(ia, ib, J);
  else { ab = key_pair(ib, ia, J); ph += (s->j2[sa] + s->j2[sb]) / 2 - J + 1; }
  if (ic <= id) cd = key_pair(ic, id, J);
  else {
    int sc = s->orb_sh[ic], sd = s->orb_sh[id];
    cd = key_pair(id, ic, J); ph += (s->j2[sc] + s->j2[sd]) / 2 - J + 1;
  }
  double xph = (double)(1 - 2 * (abs(ph) % 2));
  if (cd < ab) { int t = ab; ab = cd; cd = t; }
  lo = s->chipJ_str[key_chan(chp, ipp, J)];
  hi = s->chipJ_end[key_chan(chp, ipp, J)];
  if (lo < 0 || hi < 0) return 0.0;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    int mab = s->iabJ[mid], mcd = s->icdJ[mid];
    if (mab == ab && mcd == cd) return s->v[mid] * xph;
    if (ab < mab || (ab == mab && cd < mcd)) hi = mid - 1; else lo = mid + 1;
  }
  return 0.0;
}

/* ------------------------------------------------------------ propagator */

orc_prop *orc_prop_create(const orc_sys *s, int cap_qp, int cap_qh) {
  orc_prop *p = (orc_prop *)calloc(1, sizeof(orc_prop));
  p->nsubsh = s->nsubsh; p->nspb = s->nspb;
  p->cap_qp = cap_qp > 1 ? cap_qp : 1; p->cap_qh = cap_qh > 1 ? cap_qh : 1;
  p->np = icopy(NULL, s->nsubsh); p->nh = icopy(NULL, s->nsubsh);
  p->op = icopy(NULL, s->nsubsh); p->oh = icopy(NULL, s->nsubsh);
  p->ep = (double *)calloc(p->cap_qp, sizeof(double));
  p->zp = (double *)calloc(p->cap_qp, sizeof(double));
  p->fqp = (double *)calloc((size_t)p->cap_qp * p->nspb, sizeof(double));
  p->eh = (double *)calloc(p->cap_qh, sizeof(double));
  p->zh = (double *)calloc(p->cap_qh, sizeof(double));
  p->fqh = (double *)calloc((size_t)p->cap_qh * p->nspb, sizeof(double));
  p->clj_p = icopy(NULL, p->cap_qp); p->clj_h = icopy(NULL, p->cap_qh);
  orc_prop_clean(p);
  return p;
}

/* SpProp_t::Clean_up (Sp_Propagator.cpp:244-260) */
void orc_prop_clean(orc_prop *p) {
  int i;
  for (i = 0; i < p->nsubsh; ++i) { p->np[i] = p->nh[i] = 0; p->op[i] = p->oh[i] = 0; }
  for (i = 0; i < p->cap_qp; ++i) p->clj_p[i] = -1;
  for (i = 0; i < p->cap_qh; ++i) p->clj_h[i] = -1;
  p->n_qp = p->n_qh = 0;
}

void orc_prop_free(orc_prop *p) {
  if (!p) return;
  free(p->np); free(p->nh); free(p->op); free(p->oh);
  free(p->ep); free(p->zp); free(p->fqp); free(p->eh); free(p->zh); free(p->fqh);
  free(p->clj_p); free(p->clj_h); free(p);
}

orc_prop *orc_prop_copy(const orc_prop *q) {
  orc_prop *p = (orc_prop *)calloc(1, sizeof(orc_prop));
  *p = *q;
  p->np = icopy(q->np, q->nsubsh); p->nh = icopy(q->nh, q->nsubsh);
  p->op = icopy(q->op, q->nsubsh); p->oh = icopy(q->oh, q->nsubsh);
  p->ep = (double *)malloc(sizeof(double) * q->cap_qp); memcpy(p->ep, q->ep, sizeof(double) * q->cap_qp);
  p->zp = (double *)malloc(sizeof(double) * q->cap_qp); memcpy(p->zp, q->zp, sizeof(double) * q->cap_qp);
  p->fqp = (double *)malloc(sizeof(double) * q->cap_qp * q->nspb);
  memcpy(p->fqp, q->fqp, sizeof(double) * q->cap_qp * q->nspb);
  p->eh = (double *)malloc(sizeof(double) * q->cap_qh); memcpy(p->eh, q->eh, sizeof(double) * q->cap_qh);
  p->zh = (double *)malloc(sizeof(double) * q->cap_qh); memcpy(p->zh, q->zh, sizeof(double) * q->cap_qh);
  p->fqh = (double *)malloc(sizeof(double) * q->cap_qh * q->nspb);
  memcpy(p->fqh, q->fqh, sizeof(double) * q->cap_qh * q->nspb);
  p->clj_p = icopy(q->clj_p, q->cap_qp); p->clj_h = icopy(q->clj_h, q->cap_qh);
  return p;
}

/* Bare propagator from the model space with energies from the model-space
 * spe column (SpProp_t(ModSpace_t*, i_xe=1), Sp_Propagator.cpp:292-378, as
 * used by MakeSpProp, Miscellanea.cpp:55-63): holes at spe-10, particles at
 * spe, unit overlap vectors. */
orc_prop *orc_prop_bare(const orc_sys *s) {
  int nh = 0, np = 0, i, j;
  for (i = 0; i < s->norb; ++i) { if (s->orb_F[i] == 1) ++nh; else ++np; }
  orc_prop *p = orc_prop_create(s, np, nh);
  int ip = 0, ih = 0;
  for (i = 0; i < s->nsubsh; ++i) {
    p->op[i] = ip; p->oh[i] = ih;
    for (j = 0; j < s->no[i]; ++j) {
      int o = s->orb0[i] + j, k, n = s->orb_n[o];
      double xe = s->orb_e[o];
      if (s->orb_F[o] == 1) {
        p->eh[ih] = xe - 10.0; p->zh[ih] = 1.0; p->clj_h[ih] = i;
        for (k = 0; k < p->nspb; ++k) p->fqh[(size_t)ih * p->nspb + k] = (k == n && k < s->no[i]) ? 1.0 : 0.0;
        ++ih; ++p->nh[i];
      } else {
        p->ep[ip] = xe; p->zp[ip] = 1.0; p->clj_p[ip] = i;
        for (k = 0; k < p->nspb; ++k) p->fqp[(size_t)ip * p->nspb + k] = (k == n && k < s->no[i]) ? 1.0 : 0.0;
        ++ip; ++p->np[i];
      }
    }
  }
  p->n_qp = ip; p->n_qh = ih;
  return p;
}

/* SpProp_t::add_qp / add_qh (Sp_Propagator.cpp:683-809): append the records
 * [e, Z, X(0..no-1)] of subshell ish; only an empty subshell or the last one
 * filled may grow; Z is recomputed as |X|^2; returns the number dropped for
 * lack of storage. */
static int add_frag(orc_prop *p, int ish, int nin, int ndim, const double *sol, int no_ish, int qp) {
  int *cnt = qp ? p->np : p->nh, *off = qp ? p->op : p->oh, *clj = qp ? p->clj_p : p->clj_h;
  int *ntot = qp ? &p->n_qp : &p->n_qh, cap = qp ? p->cap_qp : p->cap_qh;
  double *e = qp ? p->ep : p->eh, *z = qp ? p->zp : p->zh, *f = qp ? p->fqp : p->fqh;
  int i, k, nst, i_orb, j;
  nin = nin > 0 ? nin : 0;
  if (cnt[ish] < 1) { cnt[ish] = 0; off[ish] = *ntot > 0 ? *ntot : 0; }
  else if (*ntot < 1 || clj[*ntot - 1] != ish) return nin;
  nst = cap - *ntot; nst = nin < nst ? nin : nst; nst = nst > 0 ? nst : 0;
  i_orb = no_ish < ndim - 2 ? no_ish : ndim - 2;
  j = *ntot;
  for (i = 0; i < nst; ++i) {
    const double *r = sol + (size_t)i * ndim;
    double x1 = 0.0;
    clj[j] = ish; e[j] = r[0];
    for (k = 0; k < i_orb; ++k) { f[(size_t)j * p->nspb + k] = r[2 + k]; x1 += r[2 + k] * r[2 + k]; }
    for (k = i_orb; k < p->nspb; ++k) f[(size_t)j * p->nspb + k] = 0.0;
    z[j] = x1; ++j;
  }
  cnt[ish] += nst; *ntot += nst;
  return nin - nst;
}

int orc_prop_add_qp(orc_prop *p, int ish, int nin, int ndim, const double *sol) {
  return add_frag(p, ish, nin, ndim, sol, ndim - 2, 1);
}
int orc_prop_add_qh(orc_prop *p, int ish, int nin, int ndim, const double *sol) {
  return add_frag(p, ish, nin, ndim, sol, ndim - 2, 0);
}

/* SpProp_t::quick_n_dirty_diff (Sp_Propagator.cpp:1189-1213): sum of |de| by
 * storage index, missing entries read as zero. */
double orc_prop_diff(const orc_prop *a, const orc_prop *b) {
  double diff = 0.0;
  int i, imax = a->n_qp > b->n_qp ? a->n_qp : b->n_qp;
  for (i = 0; i < imax; ++i) diff += fabs((i < a->n_qp ? a->ep[i] : 0.0) - (i < b->n_qp ? b->ep[i] : 0.0));
  imax = a->n_qh > b->n_qh ? a->n_qh : b->n_qh;
  for (i = 0; i < imax; ++i) diff += fabs((i < a->n_qh ? a->eh[i] : 0.0) - (i < b->n_qh ? b->eh[i] : 0.0));
  return diff;
}

/* Same pairing, max instead of sum (the max-pole-energy-change criterion). */
double orc_prop_maxdiff(const orc_prop *a, const orc_prop *b) {
  double diff = 0.0, x;
  int i, imax = a->n_qp > b->n_qp ? a->n_qp : b->n_qp;
  for (i = 0; i < imax; ++i) {
    x = fabs((i < a->n_qp ? a->ep[i] : 0.0) - (i < b->n_qp ? b->ep[i] : 0.0));
    if (x > diff) diff = x;
  }
  imax = a->n_qh > b->n_qh ? a->n_qh : b->n_qh;
  for (i = 0; i < imax; ++i) {
    x = fabs((i < a->n_qh ? a->eh[i] : 0.0) - (i < b->n_qh ? b->eh[i] : 0.0));
    if (x > diff) diff = x;
  }
  return diff;
}

int orc_prop_counts(const orc_prop *p, int *n_qp, int *n_qh, int *nspb) {
  *n_qp = p->n_qp; *n_qh = p->n_qh; *nspb = p->nspb; return 0;
}

int orc_prop_export(const orc_prop *p, int *np, int *nh, double *ep, double *zp, double *fqp,
                    double *eh, double *zh, double *fqh) {
  memcpy(np, p->np, sizeof(int) * p->nsubsh); memcpy(nh, p->nh, sizeof(int) * p->nsubsh);
  memcpy(ep, p->ep, sizeof(double) * p->n_qp); memcpy(zp, p->zp, sizeof(double) * p->n_qp);
  memcpy(fqp, p->fqp, sizeof(double) * p->n_qp * p->nspb);
  memcpy(eh, p->eh, sizeof(double) * p->n_qh); memcpy(zh, p->zh, sizeof(double) * p->n_qh);
  memcpy(fqh, p->fqh, sizeof(double) * p->n_qh * p->nspb);
  return 0;
}

/* Build a propagator from subshell-ordered fragment arrays (as the text
 * reader SpProp_t(char*, ModSpace_t*) produces, Sp_Propagator.cpp:387-610);
 * Z is taken as |X|^2 like add_qp/add_qh. */
orc_prop *orc_prop_import(const orc_sys *s, const int *np, const int *nh,
                          const double *ep, const double *fqp,
                          const double *eh, const double *fqh) {
  int i, k, tp = 0, th = 0;
  for (i = 0; i < s->nsubsh; ++i) { tp += np[i]; th += nh[i]; }
  orc_prop *p = orc_prop_create(s, tp, th);
  for (i = 0; i < s->nsubsh; ++i) {
    p->op[i] = p->n_qp; p->oh[i] = p->n_qh;
    for (k = 0; k < np[i]; ++k) {
      int j = p->n_qp++, m; double z = 0.0;
      p->ep[j] = ep[j]; p->clj_p[j] = i;
      for (m = 0; m < s->nspb; ++m) { p->fqp[(size_t)j * s->nspb + m] = fqp[(size_t)j * s->nspb + m]; }
      for (m = 0; m < s->no[i]; ++m) z += fqp[(size_t)j * s->nspb + m] * fqp[(size_t)j * s->nspb + m];
      p->zp[j] = z;
    }
    p->np[i] = np[i];
    for (k = 0; k < nh[i]; ++k) {
      int j = p->n_qh++, m; double z = 0.0;
      p->eh[j] = eh[j]; p->clj_h[j] = i;
      for (m = 0; m < s->nspb; ++m) { p->fqh[(size_t)j * s->nspb + m] = fqh[(size_t)j * s->nspb + m]; }
      for (m = 0; m < s->no[i]; ++m) z += fqh[(size_t)j * s->nspb + m] * fqh[(size_t)j * s->nspb + m];
      p->zh[j] = z;
    }
    p->nh[i] = nh[i];
  }
  return p;
}

/* ---------------------------------------------------------------- sorting */

/* Heapsort of ntot rows of width ndim by column isort, ascending
 * (Sort_u_double2dim, Utilities.cpp:25-58).  The sift order is kept so that
 * rows with equal keys land where the reference puts them. */
static void heap_rows(int ntot, int ndim, int isort, double *dbl, int descending) {
  if (ntot < 2) return;
  double *hold = (double *)malloc(sizeof(double) * ndim);
  int top = ntot / 2, last = ntot - 1, n;
  for (;;) {
    int parent, child;
    if (top > 0) {
      --top;
      memcpy(hold, dbl + (size_t)top * ndim, sizeof(double) * ndim);
    } else {
      memcpy(hold, dbl + (size_t)last * ndim, sizeof(double) * ndim);
      memcpy(dbl + (size_t)last * ndim, dbl, sizeof(double) * ndim);
      if (--last == 0) { memcpy(dbl, hold, sizeof(double) * ndim); break; }
    }
    parent = top; child = 2 * top + 1;
    while (child <= last) {
      double kc = dbl[(size_t)child * ndim + isort];
      if (child < last) {
        double kn = dbl[(size_t)(child + 1) * ndim + isort];
        if (descending ? (kc > kn) : (kc < kn)) { ++child; kc = kn; }
      }
      if (descending ? (hold[isort] > kc) : (hold[isort] < kc)) {
        for (n = 0; n < ndim; ++n) dbl[(size_t)parent * ndim + n] = dbl[(size_t)child * ndim + n];
        parent = child; child = 2 * child + 1;
      } else break;
    }
    memcpy(dbl + (size_t)parent * ndim, hold, sizeof(double) * ndim);
  }
  free(hold);
}

void orc_sort_u(int ntot, int ndim, int isort, double *dbl) { heap_rows(ntot, ndim, isort, dbl, 0); }
/* descending twin (Sort_d_double2dim, Utilities.cpp:61-94) */
void orc_sort_d(int ntot, int ndim, int isort, double *dbl) { heap_rows(ntot, ndim, isort, dbl, 1); }

/* ----------------------------------------------------- Sigma(2) couplings */

/* Count_2p1h_2h1p_poles (Self_energy_ext_Op_c.cpp:35-113), icut == 0 branch. */
void orc_count_poles(const orc_sys *s, int ish, const orc_prop *g, int *nfw, int *nbk) {
  int J, a, b, c, na, nb, ng, sum;
  *nfw = *nbk = 0;
  for (J = 0; J <= s->Jmax; ++J)
    for (a = 0; a < s->nsubsh; ++a)
      for (b = a; b < s->nsubsh; ++b) {
        if (!tri_ok(s->j2[a], s->j2[b], 2 * J)) continue;
        int nJab = ((s->j2[a] + s->j2[b]) / 2 + J + 1) % 2;
        for (c = 0; c < s->nsubsh; ++c) {
          if (s->ch[a] + s->ch[b] - s->ch[c] != s->ch[ish]) continue;
          if ((s->ip[a] + s->ip[b] + s->ip[c] + s->ip[ish]) % 2 != 0) continue;
          if (!tri_ok(s->j2[c], s->j2[ish], 2 * J)) continue;
          na = g->np[a]; nb = g->np[b]; ng = g->nh[c];
          sum = na * nb * ng;
          if (a == b) { nb = na + 1; if (nJab == 1) nb -= 2; sum = (na * nb / 2) * ng; }
          *nfw += sum;
          na = g->nh[a]; nb = g->nh[b]; ng = g->np[c];
          sum = na * nb * ng;
          if (a == b) { nb = na + 1; if (nJab == 1) nb -= 2; sum = (na * nb / 2) * ng; }
          *nbk += sum;
        }
      }
}

/* Build_Dynamic_SlfEn_2ndOrder (Self_energy_ext_Op_c.cpp:323-561), icut == 0:
 * for every (Jab, a<=b, g) block the interaction part
 *   Vt(d; alpha, beta, gamma) = sqrt((2Jab+1)/(2j+1)) V^Jab(alpha beta; d gamma) [sqrt2 if gamma==d]
 * is gathered once (:396-421) and contracted with the amplitude products of
 * every 2p1h (:425-481) and 2h1p (:484-539) pole.  Rows: [E, sum m^2, m(0..n-1)]. */
int orc_build_sigma2(const orc_sys *s, int ish, const orc_prop *g,
                     double *rows_fw, int cap_fw, double *rows_bk, int cap_bk,
                     int *nfw, int *nbk) {
  const int nI = s->no[ish], ndim = 2 + nI, nspb = g->nspb;
  int J, a, b, c, ifw = 0, ibk = 0;
  double *vt = (double *)malloc(sizeof(double) * (size_t)nspb * nspb * nspb * nspb);
  for (J = 0; J <= s->Jmax; ++J)
    for (a = 0; a < s->nsubsh; ++a)
      for (b = a; b < s->nsubsh; ++b) {
        if (!tri_ok(s->j2[a], s->j2[b], 2 * J)) continue;
        int nJab = abs((s->j2[a] + s->j2[b]) / 2 - J + 1) % 2;
        double xJab = (double)(1 - 2 * nJab);
        for (c = 0; c < s->nsubsh; ++c) {
          int nJgd = abs((s->j2[c] + s->j2[ish]) / 2 - J + 1) % 2;
          if (s->ch[a] + s->ch[b] - s->ch[c] != s->ch[ish]) continue;
          if ((s->ip[a] + s->ip[b] + s->ip[c] + s->ip[ish]) % 2 != 0) continue;
          if (!tri_ok(s->j2[c], s->j2[ish], 2 * J)) continue;
          double x3 = sqrt((double)(J + J + 1) / (double)(1 + s->j2[ish]));
          int d, al, be, ga, na, nb, ng, side;
          double *pv = vt;
          for (d = 0; d < nI; ++d) {
            int od = s->orb0[ish] + d;
            for (al = 0; al < s->no[a]; ++al) {
              int oa = s->orb0[a] + al, b0 = (a == b) ? al + nJab : 0;
              for (be = b0; be < s->no[b]; ++be) {
                int ob = s->orb0[b] + be;
                for (ga = 0; ga < s->no[c]; ++ga) {
                  int og = s->orb0[c] + ga;
                  if (og == od && nJgd) continue;
                  double val = x3 * orc_vpp_get(s, oa, ob, od, og, J);
                  if (og == od) val *= SQR2;
                  *pv++ = val;
                }
              }
            }
          }
          for (side = 0; side < 2; ++side) {
            /* side 0: 2p1h from qp(a), qp(b), qh(g); side 1: 2h1p from qh(a), qh(b), qp(g) */
            int cntA = side ? g->nh[a] : g->np[a], cntB = side ? g->nh[b] : g->np[b];
            int cntG = side ? g->np[c] : g->nh[c];
            const double *eA = side ? g->eh + g->oh[a] : g->ep + g->op[a];
            const double *eB = side ? g->eh + g->oh[b] : g->ep + g->op[b];
            const double *eG = side ? g->ep + g->op[c] : g->eh + g->oh[c];
            const double *fA = side ? g->fqh + (size_t)g->oh[a] * nspb : g->fqp + (size_t)g->op[a] * nspb;
            const double *fB = side ? g->fqh + (size_t)g->oh[b] * nspb : g->fqp + (size_t)g->op[b] * nspb;
            const double *fG = side ? g->fqp + (size_t)g->op[c] * nspb : g->fqh + (size_t)g->oh[c] * nspb;
            double *rows = side ? rows_bk : rows_fw;
            int *cnt = side ? &ibk : &ifw, cap = side ? cap_bk : cap_fw;
            for (na = 0; na < cntA; ++na) {
              int nb0 = (a == b) ? na + nJab : 0;
              for (nb = nb0; nb < cntB; ++nb)
                for (ng = 0; ng < cntG; ++ng) {
                  if (*cnt >= cap) { free(vt); return -1; }
                  double *row = rows + (size_t)(*cnt) * ndim;
                  const double *xa = fA + (size_t)na * nspb, *xb = fB + (size_t)nb * nspb;
                  const double *yg = fG + (size_t)ng * nspb;
                  double x4 = (a == b && na == nb) ? 1.0 / SQR2 : 1.0;
                  row[0] = eA[na] + eB[nb] - eG[ng];
                  row[1] = 0.0;
                  pv = vt;
                  for (d = 0; d < nI; ++d) {
                    int od = s->orb0[ish] + d;
                    double x1 = 0.0;
                    for (al = 0; al < s->no[a]; ++al) {
                      int b0 = (a == b) ? al + nJab : 0;
                      for (be = b0; be < s->no[b]; ++be) {
                        double x2 = xa[al] * xb[be];
                        if (a == b) {
                          x2 += xJab * xa[be] * xb[al];
                          if (al == be) x2 /= SQR2;
                        }
                        for (ga = 0; ga < s->no[c]; ++ga) {
                          int og = s->orb0[c] + ga;
                          if (og == od && nJgd) continue;
                          x1 += (*pv) * x2 * yg[ga];
                          ++pv;
                        }
                      }
                    }
                    x1 *= x4;
                    row[2 + d] = x1;
                    row[1] += x1 * x1;
                  }
                  ++(*cnt);
                }
            }
          }
        }
      }
  free(vt);
  *nfw = ifw; *nbk = ibk;
  return 0;
}

/* ------------------------------------------------------- static self-energy */

/* SpProp_t::ExtendedHartreeFock_Pot (Sp_Propagator_a.cpp:39-130): Sigma^inf
 * of subshell ish from the quasihole density of every subshell v,
 *   U(r,c) = sum_v sum_J (2J+1)/(2j+1) sum_{m,u in v} f_rm f_cu rho_v(u,m) V^J(r m; c u),
 * f = sqrt2 for the orbit itself, and that term dropped unless (j+j_v+J) odd. */
int orc_hf_pot(const orc_sys *s, int ish, const orc_prop *g, double *u, int ndim) {
  const int nI = s->no[ish];
  int nc, nr, v, J2, m, w, k;
  for (nc = 0; nc < nI && nc < ndim; ++nc)
    for (nr = nc; nr < nI && nr < ndim; ++nr) {
      double U = 0.0;
      for (v = 0; v < s->nsubsh; ++v) {
        int J2l = abs(s->j2[ish] - s->j2[v]), J2u = s->j2[ish] + s->j2[v];
        for (J2 = J2l; J2 <= J2u; J2 += 2) {
          int nJ = ((s->j2[ish] + s->j2[v] + J2) / 2) % 2;
          double x1 = (double)(J2 + 1) / (double)(s->j2[ish] + 1), x7 = 0.0;
          for (m = 0; m < s->no[v]; ++m) {
            if (v == ish && m == nr && nJ != 1) continue;
            double x2 = (v == ish && m == nr) ? SQR2 : 1.0;
            int ia = s->orb0[ish] + nr, im = s->orb0[v] + m;
            for (w = 0; w < s->no[v]; ++w) {
              if (v == ish && w == nc && nJ != 1) continue;
              double x3 = (v == ish && w == nc) ? SQR2 : 1.0, x6 = 0.0;
              int ib = s->orb0[ish] + nc, iu = s->orb0[v] + w;
              const double *y = g->fqh + (size_t)g->oh[v] * g->nspb;
              for (k = 0; k < g->nh[v]; ++k) { x6 += y[w] * y[m]; y += g->nspb; }
              x7 = x7 + x2 * x3 * x6 * orc_vpp_get(s, ia, im, ib, iu, J2 / 2);
            }
          }
          U = U + x1 * x7;
        }
      }
      u[nr * ndim + nc] = U;
      u[nc * ndim + nr] = U;
    }
  return nI > ndim ? nI - ndim : 0;
}

/* HO kinetic energy table T_HO(na,nb,l) in units of hbar*omega
 * (fab_tkin, Utilities.cpp:136-143; table filled as Get_Tab_kin_ho, :106-124). */
void orc_kin_ho(const orc_sys *s, int ish, double htom, double *u, int ndim) {
  int r, c, l = s->l[ish];
  for (c = 0; c < ndim; ++c)
    for (r = 0; r < ndim; ++r) {
      int na = r < s->no[ish] ? s->orb_n[s->orb0[ish] + r] : r;
      int nb = c < s->no[ish] ? s->orb_n[s->orb0[ish] + c] : c;
      double x1 = 0.0;
      if (na == nb) x1 = ((double)(na + nb + l) + 1.5) / 2.0;
      if (abs(na - nb) == 1) x1 = sqrt((double)(na + nb + 1) * (double)(na + nb + l + l + 2) / 4.0) / 2.0;
      u[r * ndim + c] = htom * x1;
    }
}

/* --------------------------------------------------------------- Lanczos */

/* Reduce_Mtx_Lncz (Lanczos.cpp:44-286): multi-pivot Lanczos on diag(E) of the
 * pole rows Min[NPLS_in x (NA_in+NB_in)] (column 0 = E, columns NA_in.. = m).
 * Each pivot (normalised in place) seeds v = M u; later vectors are E*v_prev.
 * Every vector is orthogonalised against all earlier ones, latest first, then
 * yields the row [a, b, g_0..g_{npvt-1}, m^L] of the reduced self-energy.
 * Returns the number of Lanczos vectors, < 0 on errors (the reference exits
 * on some of them). pivots: nt_pvts rows of NB_in doubles. */
int orc_reduce_mtx_lncz(int NA_out, int NB_out, int NPLS_out, double *Mout, int *i_piv_out,
                        int NA_in, int NB_in, int NPLS_in, const double *Min,
                        int nt_pvts, double *pivots, int *n_itr) {
  int i, j, l, p, nLanc = 0, nact = 0;
  if (NPLS_in < 1 || NPLS_out < 1) return -1;
  if (NB_out != NB_in) return -100;
  for (i = 0; i < nt_pvts; ++i) if (n_itr[i] > 0) { nLanc += n_itr[i]; ++nact; }
  if (nLanc < 1) return -2;
  if (nLanc > NPLS_in) nLanc = NPLS_in;
  if (nLanc > NPLS_out) return -101;
  if (nact + 1 > NA_out && nLanc < NPLS_in) return -102;
  const int Tin = NA_in + NB_in, Tout = NA_out + NB_out;
  double *V = (double *)calloc((size_t)NPLS_in * nLanc, sizeof(double));
  int *last = (int *)malloc(sizeof(int) * (nt_pvts > 0 ? nt_pvts : 1));
  for (i = 0; i < Tout * nLanc; ++i) Mout[i] = 0.0;
  for (i = 0; i < nt_pvts; ++i) last[i] = -1000;
  double *erow = Mout, *mrow = Mout + Tout - NB_out;
  int i_pvt = -1, n_pvt = -1, j_next = 0;
  for (j = 0; j < nLanc; ++j) {
    double *v = V + (size_t)j * NPLS_in;
    int fresh = (j == j_next);
    if (fresh) {
      ++i_pvt;
      while (i_pvt < nt_pvts && n_itr[i_pvt] < 1) ++i_pvt;
      if (i_pvt >= nt_pvts) { free(V); free(last); return -100; }
      ++n_pvt;
      j_next += n_itr[i_pvt];
      last[n_pvt] = j_next - 1;
      double *u = pivots + (size_t)i_pvt * NB_in, x1 = 0.0;
      for (i = 0; i < NB_in; ++i) x1 += u[i] * u[i];
      x1 = sqrt(x1);
      for (i = 0; i < NB_in; ++i) u[i] /= x1;
      for (p = 0; p < NPLS_in; ++p) {
        const double *ms = Min + (size_t)p * Tin + NA_in;
        double acc = 0.0;
        for (l = 0; l < NB_in; ++l) acc += ms[l] * u[l];
        v[p] = acc;
      }
    } else {
      const double *vp = v - NPLS_in;
      for (p = 0; p < NPLS_in; ++p) v[p] = vp[p] * Min[(size_t)p * Tin];
    }
    for (i = j - 1; i >= 0; --i) {
      const double *vi = V + (size_t)i * NPLS_in;
      double x1 = 0.0, x2 = 0.0;
      for (l = 0; l < NPLS_in; ++l) { x1 += vi[l] * v[l]; x2 += vi[l] * vi[l]; }
      for (l = 0; l < NPLS_in; ++l) v[l] -= x1 * vi[l] / x2;
    }
    double xa = 0.0, xb = 0.0;
    for (p = 0; p < NPLS_in; ++p) {
      const double *ms = Min + (size_t)p * Tin;
      double x1 = v[p];
      xa += ms[0] * x1 * x1;
      xb += x1 * x1;
      for (l = 0; l < NB_out; ++l) mrow[l] += ms[NA_in + l] * x1;
    }
    xa /= xb;
    xb = sqrt(xb);
    erow[0] = xa;
    erow[1] = xb;
    for (l = 0; l < NB_out; ++l) mrow[l] /= xb;
    for (i = 0; i < n_pvt; ++i) {
      const double *vi = V + (size_t)last[i] * NPLS_in;
      double xg = 0.0;
      for (p = 0; p < NPLS_in; ++p) xg += Min[(size_t)p * Tin] * vi[p] * v[p];
      erow[2 + i] = xg / xb;
    }
    for (p = 0; p < NPLS_in; ++p) v[p] /= xb;
    if (fresh) erow[1] = erow[1 + n_pvt];
    if (n_pvt + 2 + NB_out > Tout) { free(V); free(last); return -200; }
    erow += Tout;
    mrow += Tout;
  }
  ++n_pvt;
  i_piv_out[0] = n_pvt;
  i_piv_out[1] = 1;
  for (i = 0; i < n_pvt; ++i) i_piv_out[2 + i] = last[i] + 1;
  if (last[n_pvt - 1] + 1 != nLanc && !(last[n_pvt - 1] + 1 > nLanc && NPLS_in == nLanc)) {
    free(V); free(last); return -201;
  }
  free(V); free(last);
  return nLanc;
}

/* Wide_self_energy_t::Reduce_SelfEn_Lncz (Self_energy_ext_Op_e.cpp:55-232):
 * sort the unreduced rows (fw ascending, bk descending, :119-130), rescale
 * the iteration counts when they exceed the pole count (:199-203, :216-220;
 * counts are changed in place like the reference), run Reduce_Mtx_Lncz per
 * side.  Output rows have ntot_den = max(2, nt_pvts+1) leading columns.
 * out_fw/out_bk must hold min(sum n_itr, n) rows of ntot_den+n_orb doubles;
 * i_piv_* must hold ntot_den+n_orb ints. */
int orc_reduce_selfen_lncz(int n_orb, int nfw, double *rows_fw, int nbk, double *rows_bk,
                           int nt_pvts, double *pivots, int *n_itr_fw, int *n_itr_bk,
                           double *out_fw, int *i_piv_fw, int *nfw_out,
                           double *out_bk, int *i_piv_bk, int *nbk_out, int *ntot_den) {
  int k, nLfw = 0, nLbk = 0, ndim_in = 2 + n_orb;
  int nden = nt_pvts + 1 > 2 ? nt_pvts + 1 : 2, ndim_out = nden + n_orb;
  orc_sort_u(nfw, ndim_in, 0, rows_fw);
  orc_sort_d(nbk, ndim_in, 0, rows_bk);
  for (k = 0; k < nt_pvts; ++k) {
    if (n_itr_fw[k] > 0) nLfw += n_itr_fw[k];
    if (n_itr_bk[k] > 0) nLbk += n_itr_bk[k];
  }
  const int totfw = nLfw, totbk = nLbk;
  if (nLfw > nfw) nLfw = nfw;
  if (nLbk > nbk) nLbk = nbk;
  const int alloc_fw = nLfw > 1 ? nLfw : 1, alloc_bk = nLbk > 1 ? nLbk : 1;
  for (k = 0; k < ndim_out; ++k) { i_piv_fw[k] = -100; i_piv_bk[k] = -100; }
  *ntot_den = nden; *nfw_out = 0; *nbk_out = 0;
  if (nfw >= 1 && nLfw >= 1) {
    if (totfw > nfw) {
      double x1 = (double)nfw / (double)totfw;
      for (k = 0; k < nt_pvts; ++k) n_itr_fw[k] = (int)(x1 * n_itr_fw[k] + 1.0);
    }
    int r = orc_reduce_mtx_lncz(nden, n_orb, alloc_fw, out_fw, i_piv_fw, 2, n_orb, nfw, rows_fw,
                                nt_pvts, pivots, n_itr_fw);
    if (r < 0) return r;
    *nfw_out = r;
  }
  if (nbk >= 1 && nLbk >= 1) {
    if (totbk > nbk) {
      double x1 = (double)nbk / (double)totbk;
      for (k = 0; k < nt_pvts; ++k) n_itr_bk[k] = (int)(x1 * n_itr_bk[k] + 1.0);
    }
    int r = orc_reduce_mtx_lncz(nden, n_orb, alloc_bk, out_bk, i_piv_bk, 2, n_orb, nbk, rows_bk,
                                nt_pvts, pivots, n_itr_bk);
    if (r < 0) return r;
    *nbk_out = r;
  }
  return 0;
}

/* ------------------------------------------------- symmetric eigensolver */

/* Symmetric eigendecomposition: Householder reduction to tridiagonal form
 * with the transformation accumulated (Golub & Van Loan, Matrix Computations,
 * 4th ed., Alg. 8.3.1), then implicit QL sweeps with a Wilkinson-type shift
 * on the tridiagonal (the symmetric QL algorithm of Bowdler, Martin, Reinsch
 * and Wilkinson, Numer. Math. 11 (1968) 293).  This stands in for LAPACK
 * dsyevd_ (DysonMatrix.cpp:412); eigenvalues agree with it to rounding.
 * a: n x n row-major symmetric, overwritten so that column k holds the
 * eigenvector of w[k]; w ascending.  Returns 0, or 1 on non-convergence. */
int orc_eigh(int n, double *a, double *w) {
  if (n <= 0) return 0;
  double *q = (double *)calloc((size_t)n * n, sizeof(double));
  double *e = (double *)calloc(n, sizeof(double));
  double *vv = (double *)calloc(n, sizeof(double));
  double *pp = (double *)calloc(n, sizeof(double));
  int i, j, k, m, it;
  for (i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
  for (k = 0; k + 2 < n; ++k) {
    double s2 = 0.0;
    for (i = k + 1; i < n; ++i) s2 += a[(size_t)i * n + k] * a[(size_t)i * n + k];
    double x0 = a[(size_t)(k + 1) * n + k], tail = s2 - x0 * x0;
    if (tail <= 0.0) continue;                 /* column already reduced */
    double nrm = sqrt(s2), alpha = x0 > 0.0 ? -nrm : nrm;
    for (i = k + 1; i < n; ++i) vv[i] = a[(size_t)i * n + k];
    vv[k + 1] = x0 - alpha;
    double vtv = tail + vv[k + 1] * vv[k + 1], beta = 2.0 / vtv, pv = 0.0;
    for (i = k + 1; i < n; ++i) {
      double acc = 0.0;
      for (j = k + 1; j < n; ++j) acc += a[(size_t)i * n + j] * vv[j];
      pp[i] = beta * acc;
      pv += pp[i] * vv[i];
    }
    double K = 0.5 * beta * pv;
    for (i = k + 1; i < n; ++i) pp[i] -= K * vv[i];
    for (i = k + 1; i < n; ++i)
      for (j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= vv[i] * pp[j] + pp[i] * vv[j];
    a[(size_t)(k + 1) * n + k] = alpha; a[(size_t)k * n + k + 1] = alpha;
    for (i = k + 2; i < n; ++i) { a[(size_t)i * n + k] = 0.0; a[(size_t)k * n + i] = 0.0; }
    for (i = 0; i < n; ++i) {
      double acc = 0.0;
      for (j = k + 1; j < n; ++j) acc += q[(size_t)i * n + j] * vv[j];
      acc *= beta;
      for (j = k + 1; j < n; ++j) q[(size_t)i * n + j] -= acc * vv[j];
    }
  }
  for (i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
  for (i = 0; i + 1 < n; ++i) e[i] = a[(size_t)(i + 1) * n + i];
  e[n - 1] = 0.0;
  /* implicit QL on (w, e): e[i] couples rows i and i+1 */
  for (int lo = 0; lo < n; ++lo) {
    for (it = 0;; ++it) {
      for (m = lo; m + 1 < n; ++m) {
        double dd = fabs(w[m]) + fabs(w[m + 1]);
        if (fabs(e[m]) <= 2.2204460492503131e-16 * dd) break;
      }
      if (m == lo) break;
      if (it >= 60) { free(q); free(e); free(vv); free(pp); return 1; }
      double g = (w[lo + 1] - w[lo]) / (2.0 * e[lo]);
      double r = hypot(g, 1.0);
      g = w[m] - w[lo] + e[lo] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int deflated = 0;
      for (i = m - 1; i >= lo; --i) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) { w[i + 1] -= p; e[m] = 0.0; deflated = 1; break; }
        s = f / r; c = g / r;
        g = w[i + 1] - p;
        r = (w[i] - g) * s + 2.0 * c * b;
        p = s * r;
        w[i + 1] = g + p;
        g = c * r - b;
        for (k = 0; k < n; ++k) {
          double t = q[(size_t)k * n + i + 1];
          q[(size_t)k * n + i + 1] = s * q[(size_t)k * n + i] + c * t;
          q[(size_t)k * n + i] = c * q[(size_t)k * n + i] - s * t;
        }
      }
      if (deflated) continue;
      w[lo] -= p; e[lo] = g; e[m] = 0.0;
    }
  }
  /* ascending order, vectors follow (selection sort keeps it simple) */
  for (i = 0; i + 1 < n; ++i) {
    int kmin = i;
    for (j = i + 1; j < n; ++j) if (w[j] < w[kmin]) kmin = j;
    if (kmin != i) {
      double t = w[i]; w[i] = w[kmin]; w[kmin] = t;
      for (k = 0; k < n; ++k) {
        t = q[(size_t)k * n + i]; q[(size_t)k * n + i] = q[(size_t)k * n + kmin]; q[(size_t)k * n + kmin] = t;
      }
    }
  }
  memcpy(a, q, sizeof(double) * (size_t)n * n);
  free(q); free(e); free(vv); free(pp);
  return 0;
}

/* ------------------------------------------------------------------ Dyson */

/* Seek_EFermi (DysonMatrix.cpp:619-635): midpoint between the lowest
 * quasiparticle and the highest quasihole of the same charge in g. */
double orc_seek_efermi(const orc_sys *s, const orc_prop *g, int ish) {
  int ch = s->ch[ish], i;
  double x1 = g->ep[0] + 100., x2 = g->eh[0] - 100.;
  for (i = 0; i < g->n_qp; ++i) if (g->ep[i] < x1 && ch == s->ch[g->clj_p[i]]) x1 = g->ep[i];
  for (i = 0; i < g->n_qh; ++i) if (g->eh[i] > x2 && ch == s->ch[g->clj_h[i]]) x2 = g->eh[i];
  return (x1 + x2) / 2.0;
}

/* Calc_En (DysonMatrix.cpp:542-594): Koltun partial sums over the quasiholes
 * (three magnitude buckets), scaled by 2j+1. */
static void calc_en(int n_orb, int j2, const double *u1, int nsolb, int ndim, const double *solb,
                    double *e1b, double *esum, double *asum) {
  double e1[3] = {0, 0, 0}, es[3] = {0, 0, 0}, as[3] = {0, 0, 0};
  int ne, r, c;
  for (ne = nsolb - 1; ne >= 0; --ne) {
    const double *rec = solb + (size_t)ne * ndim, *x = rec + 2;
    double x1 = 0.0, x2 = 0.0;
    int bkt;
    for (r = 0; r < n_orb; ++r) {
      x1 += x[r] * x[r];
      for (c = 0; c < n_orb; ++c) x2 += u1[r * n_orb + c] * x[r] * x[c];
    }
    bkt = x1 > 0.01 ? 0 : (x1 > 1.e-5 ? 1 : 2);
    e1[bkt] += x2; es[bkt] += x1 * rec[0]; as[bkt] += x1;
  }
  double f = (double)j2 + 1.0;
  *e1b = (e1[0] + e1[1] + e1[2]) * f;
  *esum = (es[0] + es[1] + es[2]) * f;
  *asum = (as[0] + as[1] + as[2]) * f;
}

/* DysonMatrix (DysonMatrix.cpp:123-532): assemble
 *   [ h      M_fw^T  M_bk^T ]
 *   [ M_fw   D_fw    0      ]
 *   [ M_bk   0       D_bk   ]
 * (D diagonal for an unreduced Sigma after sorting fw descending and bk
 * ascending (:154-176); fishbone for a Lanczos-reduced one (:303-333)),
 * diagonalise, split at efermi (qp iff e > efermi), sort both lists by
 * descending energy and return records [e, Z, X(0..n_orb-1)] with
 * Z = |X|^2/|v|^2.  rows_* are sorted in place for an unreduced Sigma.
 * kolt3 = {e1b, esum, asum} of Calc_En (u1 = the one-body part Sig_1b). */
int orc_dyson(int n_orb, int j2, const double *h, const double *u1,
              int ntot_den_fw, int nfw, double *rows_fw, const int *i_piv_fw, int n_piv_fw,
              int ntot_den_bk, int nbk, double *rows_bk, const int *i_piv_bk, int n_piv_bk,
              double efermi, double *sol_qp, int *nqp, double *sol_qh, int *nqh,
              double *kolt3) {
  const int dfw = nfw > 0 ? nfw : 0, dbk = nbk > 0 ? nbk : 0, n = n_orb + dfw + dbk;
  const int Tfw = ntot_den_fw + n_orb, Tbk = ntot_den_bk + n_orb, nsol = n_orb + 2;
  int i, r, side;
  if (n_piv_fw == 0 && dfw > 0) orc_sort_d(dfw, Tfw, 0, rows_fw);
  if (n_piv_bk == 0 && dbk > 0) orc_sort_u(dbk, Tbk, 0, rows_bk);
  double *A = (double *)calloc((size_t)n * n, sizeof(double));
  double *W = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (r = 0; r < n_orb; ++r)
    for (i = 0; i < n_orb; ++i) A[(size_t)r * n + i] = u1[r * n_orb + i] + h[r * n_orb + i];
  for (side = 0; side < 2; ++side) {
    int nd = side ? dbk : dfw, T = side ? Tbk : Tfw, den = side ? ntot_den_bk : ntot_den_fw;
    int npiv = side ? n_piv_bk : n_piv_fw, base = n_orb + (side ? dfw : 0);
    const double *rows = side ? rows_bk : rows_fw;
    const int *ipv = side ? i_piv_bk : i_piv_fw;
    for (i = 0; i < nd; ++i) {
      const double *row = rows + (size_t)i * T;
      A[(size_t)(base + i) * n + base + i] = row[0];
      for (r = 0; r < n_orb; ++r) {
        A[(size_t)(base + i) * n + r] = row[den + r];
        A[(size_t)r * n + base + i] = row[den + r];
      }
    }
    if (npiv > 0) {
      int nc = 1;
      for (i = 1; i < nd; ++i) {
        double bval = rows[(size_t)i * T + 1];
        A[(size_t)(base + i) * n + base + i - 1] = bval;
        A[(size_t)(base + i - 1) * n + base + i] = bval;
        if (i == ipv[nc + 1]) {   /* a new fishbone branch starts at row i */
          for (r = i; r < nd; ++r) {
            double gval = rows[(size_t)r * T + 1 + nc];
            A[(size_t)(base + r) * n + base + i - 1] = gval;
            A[(size_t)(base + i - 1) * n + base + r] = gval;
          }
          ++nc;
        }
      }
      if (nc > npiv) { free(A); free(W); return -300; }
    }
  }
  if (orc_eigh(n, A, W)) { free(A); free(W); return -1; }
  /* split at the Fermi energy: qp ascending from the front, qh filled from
   * the back (:436-471) */
  double *sol = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * nsol);
  int nf = 0, nb = 0;
  double *solb = sol + (size_t)n * nsol;
  for (i = 0; i < n; ++i) {
    double x1 = 0.0, x2 = 0.0, *rec;
    for (r = 0; r < n; ++r) {
      double v = A[(size_t)r * n + i];
      x1 += v * v;
      if (r < n_orb) x2 += v * v;
    }
    if (W[i] > efermi) rec = sol + (size_t)(nf++) * nsol;
    else { solb -= nsol; ++nb; rec = solb; }
    rec[0] = W[i]; rec[1] = x2 / x1;
    for (r = 0; r < n_orb; ++r) rec[2 + r] = A[(size_t)r * n + i];
  }
  orc_sort_d(nf, nsol, 0, sol);
  orc_sort_d(nb, nsol, 0, solb);
  memcpy(sol_qp, sol, sizeof(double) * (size_t)nf * nsol);
  memcpy(sol_qh, solb, sizeof(double) * (size_t)nb * nsol);
  *nqp = nf; *nqh = nb;
  if (kolt3) calc_en(n_orb, j2, u1, nb, nsol, solb, kolt3, kolt3 + 1, kolt3 + 2);
  free(sol); free(A); free(W);
  return 0;
}

/* ----------------------------------------------------------------- driver */

typedef struct {
  int nfw, nbk, den_fw, den_bk, npiv_fw, npiv_bk;
  double *fw, *bk;
  int *ipiv_fw, *ipiv_bk;
} chan_sigma;

/* "%18.6le" then %lf: the value DysPivots prints (Sp_Propagator.cpp:1252)
 * and Make_LancPivots reads back (Miscellanea.cpp:366-370). */
static double print_round(double x) {
  char buf[64];
  snprintf(buf, sizeof buf, "%18.6le", x);
  return strtod(buf, NULL);
}

/* DysPivots (Sp_Propagator.cpp:1225-1289): quasiholes then quasiparticles,
 * each side by ascending energy (a fragment whose energy equals the last one
 * picked is skipped), keeping those with 100*Z >= z_cut. */
int orc_make_pivots(const orc_sys *s, const orc_prop *g, int ish, double z_cut, int round_print, double *vec) {
  int npv = 0, side, j, k, nI = s->no[ish];
  for (side = 0; side < 2; ++side) {
    int cnt = side ? g->np[ish] : g->nh[ish], off = side ? g->op[ish] : g->oh[ish];
    const double *e = side ? g->ep : g->eh, *z = side ? g->zp : g->zh, *f = side ? g->fqp : g->fqh;
    double x2 = -1.0e10;
    int pick = cnt > 0 ? 1 : -100;
    while (pick >= 0) {
      double x1 = 1.0e10;
      pick = -100;
      for (j = 0; j < cnt; ++j)
        if (e[off + j] < x1 && e[off + j] > x2 && 100. * z[off + j] >= z_cut) { x1 = e[off + j]; pick = j; }
      if (pick >= 0) {
        x2 = x1;
        for (k = 0; k < nI; ++k) {
          const double v = f[(size_t)(off + pick) * g->nspb + k];
          vec[npv * nI + k] = round_print ? print_round(v) : v;
        }
        ++npv;
      }
    }
  }
  return npv;
}

double orc_fermi(const orc_sys *s, const orc_prop *g, int ish, const orc_params *par) {
  int k;
  for (k = 0; par && k < par->n_ef; ++k)
    if (par->ef_ish[k] == ish) return par->ef_val[k];
  return orc_seek_efermi(s, g, ish);
}

static void split_itr(int total, int npiv, int *itr) {
  int k;
  for (k = 0; k < npiv; ++k) itr[k] = total / npiv + (k < total % npiv ? 1 : 0);
}

/* Run_itrs (Main.cpp:105-479) for the three modes of the hot path:
 *   mode 0 (HF=):    h = U1 + Sigma^inf[g], Dyson of h alone;
 *   mode 1 (2nd):    Sigma(2) from g each sweep, one sweep unless Lanczos;
 *   mode 2 (ExtSE=): Sigma(2) built once from gin (the MakeSelfEn -> ExtSE
 *                    seam of Miscellanea.cpp:176-267 and Main.cpp:353), then
 *                    h iterated on the current g (sc0).
 * Every sweep uses E_F of the sweep's input propagator, stops when the
 * difference measure drops below conv, and returns the newest propagator. */
int orc_run(const orc_sys *s, const orc_prop *gin, const orc_params *par,
            const int *piv_n, const int *piv_off, const int *piv_ioff,
            const double *piv_vec, const int *itr_fw, const int *itr_bk,
            orc_prop *gout, orc_stats *st) {
  const int ns = s->nsubsh, nspb = s->nspb;
  int ish, itr, i, rc = 0;
  int itr_max = par->itr_max;
  if (par->mode == 1 && !par->lanczos && itr_max > 0) itr_max = 0;   /* Main.cpp:209-212 */
  chan_sigma *cs = (chan_sigma *)calloc(ns, sizeof(chan_sigma));
  double *u = (double *)calloc((size_t)nspb * nspb, sizeof(double));
  double *u1 = (double *)calloc((size_t)nspb * nspb, sizeof(double));
  double *tk = (double *)calloc((size_t)nspb * nspb, sizeof(double));
  orc_prop *g1 = orc_prop_copy(gin);
  memset(st, 0, sizeof(*st));

  long cap = 0;

  for (itr = 0; itr <= itr_max || itr_max < 0; ++itr) {
    /* build (or reuse) the dynamic self-energy per channel */
    if (par->mode != 0 && (par->mode == 1 || itr == 0)) {
      for (ish = 0; ish < ns; ++ish) {
        chan_sigma *c = &cs[ish];
        int nI = s->no[ish], nfw, nbk;
        free(c->fw); free(c->bk); free(c->ipiv_fw); free(c->ipiv_bk);
        memset(c, 0, sizeof(*c));
        orc_count_poles(s, ish, g1, &nfw, &nbk);
        double *rfw = (double *)malloc(sizeof(double) * (size_t)(nfw > 0 ? nfw : 1) * (2 + nI));
        double *rbk = (double *)malloc(sizeof(double) * (size_t)(nbk > 0 ? nbk : 1) * (2 + nI));
        orc_build_sigma2(s, ish, g1, rfw, nfw, rbk, nbk, &nfw, &nbk);
        st->n_fw_poles += nfw; st->n_bk_poles += nbk;
        if (!par->lanczos) {
          c->nfw = nfw; c->nbk = nbk; c->den_fw = c->den_bk = 2;
          c->fw = rfw; c->bk = rbk;
        } else {
          int npv, k;
          double *pv;
          int *ifw, *ibk;
          if (piv_n) {
            npv = piv_n[ish];
            pv = (double *)malloc(sizeof(double) * (size_t)(npv > 0 ? npv : 1) * nI);
            memcpy(pv, piv_vec + piv_off[ish], sizeof(double) * (size_t)npv * nI);
            ifw = (int *)malloc(sizeof(int) * (npv > 0 ? npv : 1));
            ibk = (int *)malloc(sizeof(int) * (npv > 0 ? npv : 1));
            for (k = 0; k < npv; ++k) { ifw[k] = itr_fw[piv_ioff[ish] + k]; ibk[k] = itr_bk[piv_ioff[ish] + k]; }
          } else {
            pv = (double *)malloc(sizeof(double) * (size_t)(g1->np[ish] + g1->nh[ish] + 1) * nI);
            npv = orc_make_pivots(s, g1, ish, 60.0, par->piv_round, pv);
            ifw = (int *)malloc(sizeof(int) * (npv > 0 ? npv : 1));
            ibk = (int *)malloc(sizeof(int) * (npv > 0 ? npv : 1));
            if (par->lanc_kind == 0) { for (k = 0; k < npv; ++k) ifw[k] = ibk[k] = par->lanc_n; }
            else { split_itr(par->lanc_n, npv, ifw); split_itr(par->lanc_n, npv, ibk); }
          }
          int sfw = 0, sbk = 0;
          for (k = 0; k < npv; ++k) { if (ifw[k] > 0) sfw += ifw[k]; if (ibk[k] > 0) sbk += ibk[k]; }
          int nden = npv + 1 > 2 ? npv + 1 : 2, dim = nden + nI;
          int ofw = sfw < nfw ? sfw : nfw, obk = sbk < nbk ? sbk : nbk;
          c->fw = (double *)calloc((size_t)(ofw > 0 ? ofw : 1) * dim, sizeof(double));
          c->bk = (double *)calloc((size_t)(obk > 0 ? obk : 1) * dim, sizeof(double));
          c->ipiv_fw = (int *)malloc(sizeof(int) * dim);
          c->ipiv_bk = (int *)malloc(sizeof(int) * dim);
          int r = orc_reduce_selfen_lncz(nI, nfw, rfw, nbk, rbk, npv, pv, ifw, ibk,
                                         c->fw, c->ipiv_fw, &c->nfw, c->bk, c->ipiv_bk, &c->nbk,
                                         &c->den_fw);
          c->den_bk = c->den_fw;
          c->npiv_fw = c->nfw > 0 ? c->ipiv_fw[0] : 0;
          c->npiv_bk = c->nbk > 0 ? c->ipiv_bk[0] : 0;
          free(rfw); free(rbk); free(pv); free(ifw); free(ibk);
          if (r < 0) { rc = r; goto done; }
        }
      }
    }
    /* output propagator sized for every eigenpair */
    cap = 0;
    for (ish = 0; ish < ns; ++ish) cap += s->no[ish] + cs[ish].nfw + cs[ish].nbk;
    orc_prop *g2 = orc_prop_create(s, (int)cap, (int)cap);
    double e1tot = 0.0, etot = 0.0, atot = 0.0;
    for (ish = 0; ish < ns; ++ish) {
      int nI = s->no[ish], nqp, nqh;
      chan_sigma *c = &cs[ish];
      double k3[3];
      double ef = orc_fermi(s, g1, ish, par);
      memset(u, 0, sizeof(double) * nI * nI);
      orc_hf_pot(s, ish, g1, u, nI);
      memset(u1, 0, sizeof(double) * nI * nI);
      if (par->t_fac != 0.0) {
        /* Get_Tab_kin_ho then *= U1body_fac (Main.cpp:266-270) */
        orc_kin_ho(s, ish, par->hw, tk, nI);
        for (i = 0; i < nI * nI; ++i) u1[i] += tk[i] * par->t_fac;
      }
      if (par->ext_u1 == 1) for (i = 0; i < nI; ++i) u1[i * nI + i] += s->orb_e[s->orb0[ish] + i];
      int n = nI + c->nfw + c->nbk;
      double *sq = (double *)malloc(sizeof(double) * (size_t)n * (nI + 2));
      double *sh = (double *)malloc(sizeof(double) * (size_t)n * (nI + 2));
      /* unreduced Sigma is re-sorted by orc_dyson, on a copy so that the
       * stored pole order is the build order each sweep */
      double *fw = c->fw, *bk = c->bk;
      if (c->npiv_fw == 0 && c->nfw > 0) {
        fw = (double *)malloc(sizeof(double) * (size_t)c->nfw * (c->den_fw + nI));
        memcpy(fw, c->fw, sizeof(double) * (size_t)c->nfw * (c->den_fw + nI));
      }
      if (c->npiv_bk == 0 && c->nbk > 0) {
        bk = (double *)malloc(sizeof(double) * (size_t)c->nbk * (c->den_bk + nI));
        memcpy(bk, c->bk, sizeof(double) * (size_t)c->nbk * (c->den_bk + nI));
      }
      int den_fw = c->den_fw ? c->den_fw : 2, den_bk = c->den_bk ? c->den_bk : 2;
      int r = orc_dyson(nI, s->j2[ish], u, u1, den_fw, c->nfw, fw, c->ipiv_fw, c->npiv_fw,
                        den_bk, c->nbk, bk, c->ipiv_bk, c->npiv_bk, ef, sq, &nqp, sh, &nqh, k3);
      if (fw != c->fw) free(fw);
      if (bk != c->bk) free(bk);
      if (r < 0) { free(sq); free(sh); orc_prop_free(g2); rc = r; goto done; }
      st->lost_poles += orc_prop_add_qp(g2, ish, nqp, nI + 2, sq);
      st->lost_poles += orc_prop_add_qh(g2, ish, nqh, nI + 2, sh);
      free(sq); free(sh);
      e1tot += k3[0]; etot += (k3[0] + k3[1]) / 2.0; atot += k3[2];
    }
    double diff = par->conv_kind == 1 ? orc_prop_maxdiff(g1, g2) : orc_prop_diff(g1, g2);
    st->iterations = itr + 1; st->last_diff = diff;
    st->e_kolt = etot; st->u1_tot = e1tot; st->a_tot = atot;
    orc_prop_free(g1);
    g1 = g2;
    if (diff < par->conv) break;
  }
done:
  /* hand back the newest propagator */
  if (rc == 0) {
    orc_prop *t = gout;
    free(t->np); free(t->nh); free(t->op); free(t->oh); free(t->ep); free(t->zp); free(t->fqp);
    free(t->eh); free(t->zh); free(t->fqh); free(t->clj_p); free(t->clj_h);
    *t = *g1;
    free(g1);
  } else {
    orc_prop_free(g1);
  }
  for (ish = 0; ish < ns; ++ish) { free(cs[ish].fw); free(cs[ish].bk); free(cs[ish].ipiv_fw); free(cs[ish].ipiv_bk); }
  free(cs); free(u); free(u1); free(tk);
  return rc;
}
