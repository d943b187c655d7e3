"""ctypes bindings of libbcdor_amd.so (the C ABI declared in include/bcdor_amd.h).

The shared library is built in-tree by __graft_entry__.build() (make -C
boccadorata-public_amd).  Loading fails loudly when it is missing; every
compute entry point runs on the GPU and returns BCD_ERR_DEVICE without one.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .formats import Propagator
from .synth import Interaction, ModelSpace

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(HERE, "libbcdor_amd.so")

BCD_OK = 0
ERRORS = {-1001: "invalid argument", -1002: "no usable GPU / HIP error", -1003: "capacity exceeded",
          -1004: "Lanczos failed", -1005: "eigensolver failed", -1006: "device allocation failed"}

IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)


class BcdModspace(C.Structure):
    _fields_ = [("nsubsh", C.c_int), ("l", IP), ("j2", IP), ("ip", IP), ("ch", IP), ("no", IP),
                ("orb_n", IP), ("orb_F", IP), ("orb_e", DP)]


class BcdVpp(C.Structure):
    _fields_ = [("n_me", C.c_long), ("iabJ", IP), ("icdJ", IP), ("v", DP)]


class BcdProp(C.Structure):
    _fields_ = [("nspb", C.c_int), ("np", IP), ("nh", IP), ("ep", DP), ("fqp", DP), ("eh", DP), ("fqh", DP)]


class BcdSolveParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("hf_itr_max", C.c_int), ("sc_itr_max", C.c_int), ("lanczos", C.c_int),
                ("lanc_kind", C.c_int), ("lanc_n", C.c_int), ("conv_kind", C.c_int), ("hf_conv", C.c_double),
                ("sc_conv", C.c_double), ("t_fac", C.c_double), ("ext_u1", C.c_int), ("dyson_solver", C.c_int)]


EXPORTS = ["bcd_version", "bcd_error_string", "bcd_device_count", "bcd_ctx_create", "bcd_ctx_destroy",
           "bcd_count_poles", "bcd_sigma2_build", "bcd_hf_pot", "bcd_reduce_mtx_lncz", "bcd_reduce_selfen_lncz",
           "bcd_dyson", "bcd_syevd_batched", "bcd_batch_create", "bcd_batch_destroy", "bcd_batch_solve",
           "bcd_batch_set_profile", "bcd_batch_timings", "bcd_batch_stats", "bcd_batch_converged", "bcd_batch_result",
           "bcd_batch_kernel_times", "bcd_batch_work"]

KERNEL_NAMES = ["sigma2", "lanczos", "tridiag", "ql", "density", "sigma_inf", "finalize", "wbuild", "sigma2_plan",
                "arrow", "poles"]

_LIB = None


class BcdError(RuntimeError):
    pass


def check(rc, what):
    if rc < 0:
        raise BcdError(f"{what} failed: {rc} ({ERRORS.get(rc, 'reference error code')})")
    return rc


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BcdError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.bcd_version.restype = C.c_int
        L.bcd_error_string.restype = C.c_char_p
        L.bcd_error_string.argtypes = [C.c_int]
        L.bcd_device_count.restype = C.c_int
        L.bcd_ctx_create.argtypes = [C.c_int, C.POINTER(BcdModspace), C.POINTER(BcdVpp), C.POINTER(vp)]
        L.bcd_ctx_destroy.argtypes = [vp]
        L.bcd_ctx_destroy.restype = None
        L.bcd_count_poles.argtypes = [vp, C.c_int, C.POINTER(BcdProp), IP, IP]
        L.bcd_sigma2_build.argtypes = [vp, C.c_int, C.POINTER(BcdProp), DP, C.c_int, DP, C.c_int, IP, IP]
        L.bcd_hf_pot.argtypes = [vp, C.c_int, C.POINTER(BcdProp), DP, C.c_int]
        L.bcd_reduce_mtx_lncz.argtypes = [vp, C.c_int, C.c_int, C.c_int, DP, IP, C.c_int, C.c_int, C.c_int, DP,
                                          C.c_int, DP, IP]
        L.bcd_reduce_selfen_lncz.argtypes = [vp, C.c_int, C.c_int, DP, C.c_int, DP, C.c_int, DP, IP, IP, DP, IP,
                                             IP, DP, IP, IP, IP]
        L.bcd_dyson.argtypes = [vp, C.c_int, C.c_int, DP, DP, C.c_int, C.c_int, DP, IP, C.c_int, C.c_int,
                                C.c_int, DP, IP, C.c_int, C.c_double, DP, IP, DP, IP, DP]
        L.bcd_syevd_batched.argtypes = [vp, C.c_int, IP, LP, DP, LP, DP]
        L.bcd_batch_create.argtypes = [C.c_int, C.c_int, C.POINTER(BcdModspace), C.POINTER(BcdVpp),
                                       C.POINTER(BcdProp), C.POINTER(BcdSolveParams), C.POINTER(vp)]
        L.bcd_batch_destroy.argtypes = [vp]
        L.bcd_batch_destroy.restype = None
        L.bcd_batch_solve.argtypes = [vp]
        L.bcd_batch_set_profile.argtypes = [vp, C.c_int]
        L.bcd_batch_timings.argtypes = [vp, DP, C.c_int]
        L.bcd_batch_stats.argtypes = [vp, C.c_int, IP, IP, DP, DP, LP, LP]
        L.bcd_batch_converged.argtypes = [vp, C.c_int]
        L.bcd_batch_result.argtypes = [vp, C.c_int, C.c_int, IP, IP, IP, IP, DP, DP, DP, DP, DP, DP]
        L.bcd_batch_kernel_times.argtypes = [vp, C.c_int, DP, LP]
        L.bcd_batch_work.argtypes = [vp, DP, C.c_int]
        _LIB = L
    return _LIB


def _i(a):
    return np.ascontiguousarray(a, np.int32)


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def ip(a):
    return a.ctypes.data_as(IP)


def dp(a):
    return a.ctypes.data_as(DP)


class _Keep:
    """Holds the numpy arrays a ctypes struct points into."""

    def __init__(self):
        self.arrays = []

    def i(self, a):
        a = _i(a)
        self.arrays.append(a)
        return a

    def d(self, a):
        a = _d(a)
        self.arrays.append(a)
        return a


def modspace_struct(ms: ModelSpace, keep: _Keep) -> BcdModspace:
    return BcdModspace(ms.nsubsh, ip(keep.i(ms.l)), ip(keep.i(ms.j2)), ip(keep.i(ms.ip)), ip(keep.i(ms.ch)),
                       ip(keep.i(ms.no)), ip(keep.i(ms.orb_n)), ip(keep.i(ms.orb_F)), dp(keep.d(ms.orb_e)))


def vpp_struct(vpp: Interaction, keep: _Keep) -> BcdVpp:
    iab, icd = vpp.keys()
    return BcdVpp(vpp.n, ip(keep.i(iab)), ip(keep.i(icd)), dp(keep.d(vpp.v)))


def prop_struct(g: Propagator, nspb: int, keep: _Keep) -> BcdProp:
    fqp = np.zeros((len(g.ep), nspb)); fqp[:, :min(nspb, g.fqp.shape[1])] = g.fqp[:, :nspb]
    fqh = np.zeros((len(g.eh), nspb)); fqh[:, :min(nspb, g.fqh.shape[1])] = g.fqh[:, :nspb]
    return BcdProp(nspb, ip(keep.i(g.np_)), ip(keep.i(g.nh)), dp(keep.d(np.append(g.ep, 0.0))),
                   dp(keep.d(np.append(fqp.reshape(-1), 0.0))), dp(keep.d(np.append(g.eh, 0.0))),
                   dp(keep.d(np.append(fqh.reshape(-1), 0.0))))


class Context:
    """One system on one device (bcd_ctx): stage-level entry points."""

    def __init__(self, ms: ModelSpace, vpp: Interaction, device: int = 0):
        L = lib()
        self.ms = ms
        keep = _Keep()
        h = C.c_void_p()
        check(L.bcd_ctx_create(device, C.byref(modspace_struct(ms, keep)), C.byref(vpp_struct(vpp, keep)),
                               C.byref(h)), "bcd_ctx_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().bcd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def count_poles(self, ish: int, g: Propagator):
        keep = _Keep()
        a, b = C.c_int(), C.c_int()
        check(lib().bcd_count_poles(self.h, ish, C.byref(prop_struct(g, self.ms.nspb, keep)), C.byref(a),
                                    C.byref(b)), "bcd_count_poles")
        return a.value, b.value

    def sigma2_build(self, ish: int, g: Propagator):
        nfw, nbk = self.count_poles(ish, g)
        w = 2 + int(self.ms.no[ish])
        fw = np.zeros(max(nfw, 1) * w); bk = np.zeros(max(nbk, 1) * w)
        keep = _Keep()
        a, b = C.c_int(), C.c_int()
        check(lib().bcd_sigma2_build(self.h, ish, C.byref(prop_struct(g, self.ms.nspb, keep)), dp(fw), nfw, dp(bk),
                                     nbk, C.byref(a), C.byref(b)), "bcd_sigma2_build")
        return fw[:a.value * w].reshape(a.value, w), bk[:b.value * w].reshape(b.value, w)

    def hf_pot(self, ish: int, g: Propagator):
        n = int(self.ms.no[ish])
        u = np.zeros(n * n)
        keep = _Keep()
        check(lib().bcd_hf_pot(self.h, ish, C.byref(prop_struct(g, self.ms.nspb, keep)), dp(u), n), "bcd_hf_pot")
        return u.reshape(n, n)

    def reduce_mtx_lncz(self, NA_out, NB_out, NPLS_out, Min, NA_in, NB_in, pivots, n_itr):
        Min = _d(Min)
        piv = _d(pivots).copy().reshape(-1)
        itr = _i(n_itr).copy()
        out = np.zeros(NPLS_out * (NA_out + NB_out))
        ipv = np.zeros(len(itr) + 2, np.int32)
        r = lib().bcd_reduce_mtx_lncz(self.h, NA_out, NB_out, NPLS_out, dp(out), ip(ipv), NA_in, NB_in,
                                      Min.shape[0], dp(Min.reshape(-1)), len(itr), dp(piv), ip(itr))
        return r, out.reshape(NPLS_out, NA_out + NB_out), ipv, piv.reshape(len(itr), -1)

    def reduce_selfen_lncz(self, n_orb, rows_fw, rows_bk, pivots, itr_fw, itr_bk):
        fw = _d(rows_fw).copy(); bk = _d(rows_bk).copy()
        npv = len(itr_fw)
        piv = _d(pivots).copy().reshape(-1)
        ifw = _i(itr_fw).copy(); ibk = _i(itr_bk).copy()
        den = max(npv + 1, 2); dim = den + n_orb
        cfw = max(min(int(ifw[ifw > 0].sum()), fw.shape[0]), 1)
        cbk = max(min(int(ibk[ibk > 0].sum()), bk.shape[0]), 1)
        ofw = np.zeros(cfw * dim); obk = np.zeros(cbk * dim)
        pfw = np.zeros(dim, np.int32); pbk = np.zeros(dim, np.int32)
        a, b, d = C.c_int(), C.c_int(), C.c_int()
        z = np.zeros(1)
        check(lib().bcd_reduce_selfen_lncz(self.h, n_orb, fw.shape[0], dp(fw if fw.size else z), bk.shape[0],
                                           dp(bk if bk.size else z), npv, dp(piv), ip(ifw), ip(ibk), dp(ofw),
                                           ip(pfw), C.byref(a), dp(obk), ip(pbk), C.byref(b), C.byref(d)),
              "bcd_reduce_selfen_lncz")
        return (ofw[:a.value * dim].reshape(a.value, dim), pfw, obk[:b.value * dim].reshape(b.value, dim), pbk,
                d.value, ifw, ibk, fw, bk)

    def dyson(self, n_orb, j2, h, u1, fw, ipiv_fw, npiv_fw, bk, ipiv_bk, npiv_bk, efermi, den_fw=2, den_bk=2):
        fw = _d(fw); bk = _d(bk)
        n = n_orb + fw.shape[0] + bk.shape[0]
        sq = np.zeros(n * (n_orb + 2)); sh = np.zeros(n * (n_orb + 2)); k3 = np.zeros(3)
        a, b = C.c_int(), C.c_int()
        z = np.zeros(1)
        pf = _i(ipiv_fw if ipiv_fw is not None else np.zeros(1))
        pb = _i(ipiv_bk if ipiv_bk is not None else np.zeros(1))
        check(lib().bcd_dyson(self.h, n_orb, j2, dp(_d(h).reshape(-1)), dp(_d(u1).reshape(-1)), den_fw,
                              fw.shape[0], dp(fw.reshape(-1) if fw.size else z), ip(pf), npiv_fw, den_bk,
                              bk.shape[0], dp(bk.reshape(-1) if bk.size else z), ip(pb), npiv_bk, efermi, dp(sq),
                              C.byref(a), dp(sh), C.byref(b), dp(k3)), "bcd_dyson")
        w = n_orb + 2
        return sq[:a.value * w].reshape(a.value, w), sh[:b.value * w].reshape(b.value, w), k3

    def syevd_batched(self, mats):
        ns = np.array([m.shape[0] for m in mats], np.int32)
        off = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
        woff = np.concatenate([[0], np.cumsum(ns.astype(np.int64))[:-1]]).astype(np.int64)
        a = np.concatenate([_d(m).reshape(-1) for m in mats])
        w = np.zeros(int(ns.sum()))
        check(lib().bcd_syevd_batched(self.h, len(mats), ip(ns), off.ctypes.data_as(LP), dp(a),
                                      woff.ctypes.data_as(LP), dp(w)), "bcd_syevd_batched")
        return [(w[woff[k]:woff[k] + ns[k]], a[off[k]:off[k] + ns[k] ** 2].reshape(ns[k], ns[k]))
                for k in range(len(mats))]


def solve_params(mode=2, hf_itr_max=-1, sc_itr_max=-1, lanczos=True, lanc_kind=1, lanc_n=50, conv_kind=1,
                 hf_conv=1e-8, sc_conv=1e-8, t_fac=0.0, ext_u1=1, dyson_solver=0) -> BcdSolveParams:
    """dyson_solver 0: arrowhead updates (default); 1: dense Householder + QL per sweep."""
    return BcdSolveParams(mode, hf_itr_max, sc_itr_max, int(lanczos), lanc_kind, lanc_n, conv_kind, hf_conv,
                          sc_conv, t_fac, ext_u1, dyson_solver)


class Batch:
    """A device-resident batch of independent instances (bcd_batch)."""

    def __init__(self, systems, params: BcdSolveParams, g0=None, device: int = 0):
        L = lib()
        n = len(systems)
        keep = _Keep()
        self.systems = systems
        msa = (BcdModspace * n)(*[modspace_struct(ms, keep) for ms, _ in systems])
        vpa = (BcdVpp * n)(*[vpp_struct(v, keep) for _, v in systems])
        gpa = None
        if g0 is not None:
            gpa = (BcdProp * n)(*[prop_struct(g, ms.nspb, keep) for g, (ms, _) in zip(g0, systems)])
        h = C.c_void_p()
        check(L.bcd_batch_create(device, n, msa, vpa, gpa, C.byref(params), C.byref(h)), "bcd_batch_create")
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            lib().bcd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_profile(self, on: bool):
        check(lib().bcd_batch_set_profile(self.h, int(on)), "bcd_batch_set_profile")

    def solve(self):
        check(lib().bcd_batch_solve(self.h), "bcd_batch_solve")

    def timings(self):
        t = np.zeros(6)
        check(lib().bcd_batch_timings(self.h, dp(t), 6), "bcd_batch_timings")
        return dict(zip(["wbuild", "hf", "sigma2", "lanczos", "sc0", "total"], t.tolist()))

    def stats(self, i):
        a, b = C.c_int(), C.c_int()
        d, e = C.c_double(), C.c_double()
        f, g = C.c_long(), C.c_long()
        check(lib().bcd_batch_stats(self.h, i, C.byref(a), C.byref(b), C.byref(d), C.byref(e), C.byref(f),
                                    C.byref(g)), "bcd_batch_stats")
        return dict(hf_its=a.value, sc_its=b.value, last_diff=d.value, e_kolt=e.value, n_fw=f.value, n_bk=g.value)

    def converged(self, i) -> bool:
        return check(lib().bcd_batch_converged(self.h, i), "bcd_batch_converged") == 1

    def result(self, i) -> Propagator:
        ms = self.systems[i][0]
        nq, nh = C.c_int(), C.c_int()
        check(lib().bcd_batch_result(self.h, i, ms.nspb, None, None, C.byref(nq), C.byref(nh), None, None, None,
                                     None, None, None), "bcd_batch_result")
        nq, nh = nq.value, nh.value
        ns = ms.nspb
        npa = np.zeros(ms.nsubsh, np.int32); nha = np.zeros(ms.nsubsh, np.int32)
        ep = np.zeros(max(nq, 1)); zp = np.zeros(max(nq, 1)); fqp = np.zeros(max(nq, 1) * ns)
        eh = np.zeros(max(nh, 1)); zh = np.zeros(max(nh, 1)); fqh = np.zeros(max(nh, 1) * ns)
        check(lib().bcd_batch_result(self.h, i, ns, ip(npa), ip(nha), None, None, dp(ep), dp(zp), dp(fqp), dp(eh),
                                     dp(zh), dp(fqh)), "bcd_batch_result")
        return Propagator(npa, nha, ep[:nq], zp[:nq], fqp[:nq * ns].reshape(nq, ns), eh[:nh], zh[:nh],
                          fqh[:nh * ns].reshape(nh, ns))

    def kernel_times(self):
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            a, b = C.c_double(), C.c_long()
            check(lib().bcd_batch_kernel_times(self.h, k, C.byref(a), C.byref(b)), "bcd_batch_kernel_times")
            out[name] = (a.value, b.value)
        return out

    def work(self):
        w = np.zeros(15)
        check(lib().bcd_batch_work(self.h, dp(w), 15), "bcd_batch_work")
        phases = ["load_sort", "deflation", "roots", "loewner", "eigenvectors", "reorder", "output", "norms"]
        return dict(coupling_flops=w[0], lanczos_bytes=w[1], dyson_flops=w[2], n_eig=w[3], sum_n3=w[4],
                    arrow_flops=w[5], poles_flops=w[6], arrow_phase_ms=dict(zip(phases, w[7:15])))
