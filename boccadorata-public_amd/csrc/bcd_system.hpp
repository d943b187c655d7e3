// bcd_system.hpp -- host-side model space and interaction packing.
#pragma once

#include <vector>

#include "bcd_kernels.hpp"
#include "bcdor_amd.h"

namespace bcd {

// One system: ModSpace_t tables plus the interaction repacked from VppInt_t's
// sorted key/value lists into dense symmetric blocks per two-body channel
// (J, parity, charge), and the (J, a<=b, g) block list of the Sigma(2) build
// of every subshell in the reference's loop order (Self_energy_ext_Op_c.cpp:380-395).
struct HostSystem {
  int nsubsh = 0, norb = 0, nspb = 0, jmax = 0, nchan2b = 0, nq = 0;
  std::vector<int> l, j2, ip, ch, no, orb0, orb_sh, orb_n, orb_F, qoff;
  std::vector<double> orb_e;
  std::vector<int> pidx, npair;
  std::vector<long> voff;
  std::vector<double> V;                       // dense blocks (emptied once uploaded, see dV)
  const double *dV = nullptr;                  // the dense blocks in HBM when uploaded at add time
  long v_len = 0;                              // doubles of the dense blocks
  std::vector<std::vector<BlockDev>> blocks;   // [ish]
  // per subshell: the blocks grouped by (a, b, g) triple, each group in J
  // order; trip_start[ish] has one entry per triple plus the end
  std::vector<std::vector<int>> trip_start, trip_blk;
  std::vector<long> mbound_fw, mbound_bk;      // pole-count bounds for HF-like propagators
  int max_vt = 0;                              // largest nI * pairs * n_g of a Sigma(2) block
  int max_dg = 0;                              // largest nI * n_g of a Sigma(2) block
  long dropped = 0;                            // interaction keys outside the pair tables
  // instance descriptors of the learned start, estimated from the inputs:
  // hbar*omega as the least-squares slope of the spe column against the
  // oscillator quantum 2n + l + 3/2, the largest 2n + l, the rms of the
  // stored interaction elements
  double hw_est = 0.0, v_rms = 0.0;
  int nmax_est = 0;

  int build(const bcd_modspace *ms, const bcd_vpp *vpp);
  // one-body part U1 of subshell ish: (hw * T_HO) * t_fac, then + diag(spe)
  // when ext_u1 == 1, in the order the reference adds them (Main.cpp:266-285)
  void u1(int ish, double hw, double t_fac, int ext_u1, double *out) const;
};

}  // namespace bcd
