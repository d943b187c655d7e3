// bcd_system.cpp -- host-side packing of the model space and the interaction.
#include "bcd_system.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>

namespace bcd {

static bool tri_ok(int a2, int b2, int c2) {
  const int s1 = a2 + b2 - c2, s2 = b2 + c2 - a2, s3 = c2 + a2 - b2;
  return s1 >= 0 && s2 >= 0 && s3 >= 0 && !(s1 & 1) && !(s2 & 1) && !(s3 & 1);
}

int HostSystem::build(const bcd_modspace *ms, const bcd_vpp *vpp) {
  if (!ms || ms->nsubsh <= 0 || ms->nsubsh > kMaxSub) return BCD_ERR_ARG;
  nsubsh = ms->nsubsh;
  l.assign(ms->l, ms->l + nsubsh);
  j2.assign(ms->j2, ms->j2 + nsubsh);
  ip.assign(ms->ip, ms->ip + nsubsh);
  ch.assign(ms->ch, ms->ch + nsubsh);
  no.assign(ms->no, ms->no + nsubsh);
  orb0.resize(nsubsh);
  norb = 0;
  nspb = 0;
  jmax = 0;
  for (int i = 0; i < nsubsh; ++i) {
    if (no[i] <= 0 || no[i] > kMaxNI || ch[i] < 0 || ch[i] > 1) return BCD_ERR_ARG;
    orb0[i] = norb;
    norb += no[i];
    if (no[i] > nspb) nspb = no[i];
    if (j2[i] > jmax) jmax = j2[i];   // Calculate_Jch_bounds: Jmax = max 2j (Mod_Space.cpp:593)
  }
  if (norb >= kKeyAB || jmax >= kKeyJ) return BCD_ERR_ARG;
  orb_n.assign(ms->orb_n, ms->orb_n + norb);
  orb_F.assign(ms->orb_F, ms->orb_F + norb);
  orb_e.assign(ms->orb_e, ms->orb_e + norb);
  orb_sh.resize(norb);
  for (int i = 0; i < nsubsh; ++i)
    for (int k = 0; k < no[i]; ++k) orb_sh[orb0[i] + k] = i;
  {
    double sen = 0.0, snn = 0.0;
    nmax_est = 0;
    for (int a = 0; a < norb; ++a) {
      const int N = 2 * orb_n[a] + l[orb_sh[a]];
      if (N > nmax_est) nmax_est = N;
      sen += orb_e[a] * (N + 1.5);
      snn += (N + 1.5) * (N + 1.5);
    }
    hw_est = snn > 0.0 ? sen / snn : 0.0;
    double sv = 0.0;
    for (long k = 0; k < (vpp ? vpp->n_me : 0); ++k) sv += vpp->v[k] * vpp->v[k];
    v_rms = (vpp && vpp->n_me > 0) ? std::sqrt(sv / (double)vpp->n_me) : 0.0;
  }
  qoff.resize(nsubsh);
  nq = 0;
  for (int i = 0; i < nsubsh; ++i) { qoff[i] = nq; nq += no[i] * no[i]; }

  // pair tables: pairs a <= b allowed by the triangle rule and Pauli
  pidx.assign((size_t)(jmax + 1) * norb * norb, -1);
  std::map<long, int> chan_id;
  npair.clear();
  for (int J = 0; J <= jmax; ++J)
    for (int a = 0; a < norb; ++a)
      for (int b = a; b < norb; ++b) {
        const int sa = orb_sh[a], sb = orb_sh[b];
        if (!tri_ok(j2[sa], j2[sb], 2 * J)) continue;
        if (a == b && (J & 1)) continue;
        const long key = ((long)J * 2 + (l[sa] + l[sb]) % 2) * 8 + (ch[sa] + ch[sb]);
        auto it = chan_id.find(key);
        int c;
        if (it == chan_id.end()) {
          c = (int)npair.size();
          chan_id[key] = c;
          npair.push_back(0);
        } else {
          c = it->second;
        }
        if (npair[c] >= 0xffff) return BCD_ERR_CAPACITY;
        pidx[((size_t)J * norb + a) * norb + b] = (c << 16) | npair[c];
        ++npair[c];
      }
  nchan2b = (int)npair.size();
  voff.resize(nchan2b);
  long tot = 0;
  for (int c = 0; c < nchan2b; ++c) { voff[c] = tot; tot += (long)npair[c] * npair[c]; }
  V.assign(tot, 0.0);
  v_len = tot;
  dropped = 0;
  if (vpp && vpp->n_me > 0) {
    for (long q = 0; q < vpp->n_me; ++q) {
      const int k1 = vpp->iabJ[q], k2 = vpp->icdJ[q];
      const int J = k1 % kKeyJ, J2 = k2 % kKeyJ;
      const int a = (k1 / kKeyJ) / kKeyAB, b = (k1 / kKeyJ) % kKeyAB;
      const int c = (k2 / kKeyJ) / kKeyAB, d = (k2 / kKeyJ) % kKeyAB;
      if (J != J2 || J > jmax || a >= norb || b >= norb || c >= norb || d >= norb || a > b || c > d) {
        ++dropped;
        continue;
      }
      const int pa = pidx[((size_t)J * norb + a) * norb + b], pc = pidx[((size_t)J * norb + c) * norb + d];
      if (pa < 0 || pc < 0 || (pa >> 16) != (pc >> 16)) { ++dropped; continue; }
      const int cc = pa >> 16, ra = pa & 0xffff, rc = pc & 0xffff;
      V[voff[cc] + (long)ra * npair[cc] + rc] = vpp->v[q];
      V[voff[cc] + (long)rc * npair[cc] + ra] = vpp->v[q];
    }
  }

  // Sigma(2) blocks per subshell
  blocks.assign(nsubsh, {});
  mbound_fw.assign(nsubsh, 0);
  mbound_bk.assign(nsubsh, 0);
  max_vt = 0;
  max_dg = 0;
  for (int I = 0; I < nsubsh; ++I) {
    for (int J = 0; J <= jmax; ++J)
      for (int a = 0; a < nsubsh; ++a)
        for (int b = a; b < nsubsh; ++b) {
          if (!tri_ok(j2[a], j2[b], 2 * J)) continue;
          const int nJab = std::abs((j2[a] + j2[b]) / 2 - J + 1) % 2;
          for (int g = 0; g < nsubsh; ++g) {
            const int nJgd = std::abs((j2[g] + j2[I]) / 2 - J + 1) % 2;
            if (ch[a] + ch[b] - ch[g] != ch[I]) continue;
            if ((ip[a] + ip[b] + ip[g] + ip[I]) % 2 != 0) continue;
            if (!tri_ok(j2[g], j2[I], 2 * J)) continue;
            int np = 0;
            for (int al = 0; al < no[a]; ++al) np += no[b] - ((a == b) ? al + nJab : 0);
            if (np < 0) np = 0;
            if (np >= kMaxPairs - 1 || (long)no[I] * np * no[g] > kMaxVt) return BCD_ERR_CAPACITY;
            max_vt = std::max(max_vt, no[I] * np * no[g]);
            max_dg = std::max(max_dg, no[I] * no[g]);
            BlockDev B;
            B.gc = I; B.a = a; B.b = b; B.g = g; B.J = J; B.nJab = nJab; B.nJgd = nJgd;
            // every (alpha beta) pair of the block lies in one two-body channel
            // (same J, parity and charge); take it from the first stored pair
            B.c2b = -1;
            for (int al = 0; al < no[a] && B.c2b < 0; ++al)
              for (int be = (a == b) ? al + nJab : 0; be < no[b] && B.c2b < 0; ++be) {
                const int pa = pidx[((size_t)J * norb + orb0[a] + al) * norb + orb0[b] + be];
                if (pa >= 0) B.c2b = pa >> 16;
              }
            blocks[I].push_back(B);
            const long pairs = (a == b) ? (long)no[a] * (no[a] + 1) / 2 : (long)no[a] * no[b];
            mbound_fw[I] += pairs * no[g];
            mbound_bk[I] += pairs * no[g];
          }
        }
  }
  // blocks of one subshell grouped by (a, b, g): the same fragments of a, b, g
  // give one pole energy in every J block of the triple
  trip_start.assign(nsubsh, {});
  trip_blk.assign(nsubsh, {});
  for (int I = 0; I < nsubsh; ++I) {
    std::vector<int> order(blocks[I].size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const BlockDev &X = blocks[I][x], &Y = blocks[I][y];
      if (X.a != Y.a) return X.a < Y.a;
      if (X.b != Y.b) return X.b < Y.b;
      return X.g < Y.g;
    });
    for (size_t k = 0; k < order.size(); ++k) {
      const BlockDev &X = blocks[I][order[k]];
      if (k == 0 || X.a != blocks[I][order[k - 1]].a || X.b != blocks[I][order[k - 1]].b ||
          X.g != blocks[I][order[k - 1]].g)
        trip_start[I].push_back((int)k);
      trip_blk[I].push_back(order[k]);
    }
    trip_start[I].push_back((int)order.size());
  }
  return BCD_OK;
}

// Sig_1b of Run_itrs: Get_Tab_kin_ho fills htom * fab_tkin (Utilities.cpp:
// 106-143), the table is scaled by U1body_fac and added to zero (Main.cpp:
// 266-270); with ExtU1=1 the model-space energies are added on the diagonal
// (Main.cpp:279-285).  Trel=7 (t_fac == 0) adds no kinetic table at all.
void HostSystem::u1(int ish, double hw, double t_fac, int ext_u1, double *out) const {
  const int n = no[ish], ll = l[ish];
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      const int na = orb_n[orb0[ish] + r], nb = orb_n[orb0[ish] + c];
      double t = 0.0;
      if (na == nb) t = ((double)(na + nb + ll) + 1.5) / 2.0;
      if (std::abs(na - nb) == 1) t = std::sqrt((double)(na + nb + 1) * (double)(na + nb + ll + ll + 2) / 4.0) / 2.0;
      double v = 0.0;
      if (t_fac != 0.0) v += (hw * t) * t_fac;
      if (ext_u1 == 1 && r == c) v += orb_e[orb0[ish] + r];
      out[r * n + c] = v;
    }
}

}  // namespace bcd
