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
                ("sc_conv", C.c_double), ("t_fac", C.c_double), ("ext_u1", C.c_int), ("dyson_solver", C.c_int),
                ("precision", C.c_int), ("accel", C.c_int), ("hw", C.c_double), ("mix_start", C.c_double),
                ("piv_round", C.c_int)]


EXPORTS = ["bcd_version", "bcd_error_string", "bcd_device_count", "bcd_ctx_create", "bcd_ctx_destroy",
           "bcd_count_poles", "bcd_sigma2_build", "bcd_hf_pot", "bcd_reduce_mtx_lncz", "bcd_reduce_selfen_lncz",
           "bcd_dyson", "bcd_syevd_batched", "bcd_batch_create", "bcd_batch_create_ex", "bcd_batch_set_efermi",
           "bcd_pivot_round", "bcd_reduce_mtx_lncz_noalloc", "bcd_noalloc_scratch_bytes", "bcd_batch_set_params", "bcd_batch_error_detail", "bcd_batch_begin", "bcd_batch_add", "bcd_batch_end", "bcd_batch_destroy",
           "bcd_batch_solve",
           "bcd_batch_set_profile", "bcd_batch_timings", "bcd_batch_stats", "bcd_batch_converged", "bcd_batch_result",
           "bcd_batch_kernel_times", "bcd_batch_work", "bcd_batch_set_lanczos_shapes", "bcd_batch_autotune", "bcd_batch_lanczos_shapes",
           "bcd_ladd_count", "bcd_ladd_basis", "bcd_ladd_fill",
           "bcd_ph_count", "bcd_ph_basis", "bcd_ph_fill", "bcd_t2_create", "bcd_t2_destroy", "bcd_t2_count",
           "bcd_t2_info", "bcd_t2_basis", "bcd_t2_build", "bcd_t2_matrix", "bcd_t2_set_profile",
           "bcd_t2_kernel_times", "bcd_t2_work", "bcd_mlp_create", "bcd_mlp_destroy", "bcd_mlp_num_params",
           "bcd_mlp_get_params", "bcd_mlp_set_params", "bcd_mlp_forward", "bcd_mlp_loss_grad", "bcd_mlp_fit",
           "bcd_batch_set_learned_start", "bcd_batch_set_fallback_tol", "bcd_batch_record_learning",
           "bcd_batch_num_channels", "bcd_batch_learning_data", "bcd_batch_start_info", "bcd_batch_set_graphs"]

LEARN_IN, LEARN_OUT, LEARN_NI = 80, 32, 6
LEARN_D1, LEARN_D2, LEARN_TRI = 53, 24, 21
LEARN_SCALE = 10.0

KERNEL_NAMES = ["sigma2", "lanczos", "tridiag", "ql", "density", "sigma_inf", "finalize", "wbuild", "sigma2_plan",
                "arrow", "poles", "merge"]

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
        L.bcd_reduce_mtx_lncz_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, DP, IP, C.c_int, C.c_int, C.c_int,
                                                  DP, C.c_int, DP, IP]
        L.bcd_noalloc_scratch_bytes.restype = C.c_long
        L.bcd_reduce_selfen_lncz.argtypes = [vp, C.c_int, C.c_int, DP, C.c_int, DP, C.c_int, DP, IP, IP, DP, IP,
                                             IP, DP, IP, IP, IP]
        L.bcd_dyson.argtypes = [vp, C.c_int, C.c_int, DP, DP, C.c_int, C.c_int, DP, IP, C.c_int, C.c_int,
                                C.c_int, DP, IP, C.c_int, C.c_double, DP, IP, DP, IP, DP]
        L.bcd_syevd_batched.argtypes = [vp, C.c_int, IP, LP, DP, LP, DP]
        L.bcd_batch_create.argtypes = [C.c_int, C.c_int, C.POINTER(BcdModspace), C.POINTER(BcdVpp),
                                       C.POINTER(BcdProp), C.POINTER(BcdSolveParams), C.POINTER(vp)]
        L.bcd_batch_create_ex.argtypes = [C.c_int, C.c_int, C.POINTER(BcdModspace), C.POINTER(BcdVpp),
                                          C.POINTER(BcdProp), C.POINTER(BcdSolveParams), IP, C.POINTER(vp)]
        L.bcd_batch_set_efermi.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        L.bcd_batch_set_params.argtypes = [vp, C.POINTER(BcdSolveParams)]
        L.bcd_batch_error_detail.argtypes = [vp, C.c_char_p, C.c_int]
        L.bcd_batch_begin.argtypes = [C.c_int, C.c_int, C.POINTER(BcdSolveParams), C.POINTER(vp)]
        L.bcd_batch_add.argtypes = [vp, C.c_int, C.POINTER(BcdModspace), C.POINTER(BcdVpp), C.POINTER(BcdProp),
                                    C.c_int]
        L.bcd_batch_end.argtypes = [vp]
        L.bcd_pivot_round.argtypes = [C.c_int, DP, DP, C.c_long]
        L.bcd_batch_destroy.argtypes = [vp]
        L.bcd_batch_destroy.restype = None
        L.bcd_batch_solve.argtypes = [vp]
        L.bcd_batch_set_profile.argtypes = [vp, C.c_int]
        L.bcd_batch_timings.argtypes = [vp, DP, C.c_int]
        L.bcd_batch_autotune.argtypes = [vp]
        L.bcd_batch_lanczos_shapes.argtypes = [vp, IP, C.POINTER(C.c_float), C.c_int]
        L.bcd_batch_stats.argtypes = [vp, C.c_int, IP, IP, DP, DP, LP, LP]
        L.bcd_batch_converged.argtypes = [vp, C.c_int]
        L.bcd_batch_result.argtypes = [vp, C.c_int, C.c_int, IP, IP, IP, IP, DP, DP, DP, DP, DP, DP]
        L.bcd_batch_kernel_times.argtypes = [vp, C.c_int, DP, LP]
        L.bcd_batch_work.argtypes = [vp, DP, C.c_int]
        L.bcd_batch_set_lanczos_shapes.argtypes = [vp, IP, C.c_int]
        PP = C.POINTER(BcdProp)
        for f in (L.bcd_ladd_count, L.bcd_ph_count):
            f.argtypes = [vp, PP] + [C.c_int] * 5 + [IP, IP, IP, IP]
        for f in (L.bcd_ladd_basis, L.bcd_ph_basis):
            f.argtypes = [vp, PP] + [C.c_int] * 5 + [IP, IP, C.c_int, C.c_int, IP, IP, DP, IP, IP, DP, IP, IP]
        for f in (L.bcd_ladd_fill, L.bcd_ph_fill):
            f.argtypes = [vp, PP, C.c_int, C.c_int, C.c_int, C.c_int, IP, IP, DP, C.c_int, IP, IP, DP, C.c_int, DP,
                          C.c_int, DP, C.c_int]
        FP = C.POINTER(C.c_float)
        L.bcd_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
        L.bcd_mlp_destroy.argtypes = [vp]
        L.bcd_mlp_destroy.restype = None
        L.bcd_mlp_num_params.argtypes = [vp]
        L.bcd_mlp_num_params.restype = C.c_long
        L.bcd_mlp_get_params.argtypes = [vp, FP]
        L.bcd_mlp_set_params.argtypes = [vp, FP]
        L.bcd_mlp_forward.argtypes = [vp, C.c_int, FP, FP]
        L.bcd_mlp_loss_grad.argtypes = [vp, C.c_int, FP, FP, FP, FP]
        L.bcd_mlp_fit.argtypes = [vp, C.c_int, FP, FP, C.c_int, C.c_int, C.c_float, C.c_ulonglong, FP]
        L.bcd_batch_set_learned_start.argtypes = [vp, vp]
        L.bcd_batch_set_fallback_tol.argtypes = [vp, C.c_double]
        L.bcd_batch_record_learning.argtypes = [vp, C.c_int]
        L.bcd_batch_num_channels.argtypes = [vp]
        L.bcd_batch_learning_data.argtypes = [vp, FP, FP, IP, C.c_int]
        L.bcd_batch_start_info.argtypes = [vp, C.c_int, IP, IP, IP, DP]
        L.bcd_batch_set_graphs.argtypes = [vp, C.c_int]
        L.bcd_t2_create.argtypes = [vp, PP, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.bcd_t2_destroy.argtypes = [vp]
        L.bcd_t2_destroy.restype = None
        L.bcd_t2_count.argtypes = [vp]
        L.bcd_t2_info.argtypes = [vp, C.c_int, IP]
        L.bcd_t2_basis.argtypes = [vp, C.c_int, IP, IP, DP, IP, IP, DP]
        L.bcd_t2_build.argtypes = [vp]
        L.bcd_t2_matrix.argtypes = [vp, C.c_int, DP, DP]
        L.bcd_t2_set_profile.argtypes = [vp, C.c_int]
        L.bcd_t2_kernel_times.argtypes = [vp, C.c_int, DP, LP]
        L.bcd_t2_work.argtypes = [vp, DP, C.c_int]
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


LADD_KEYS = ("n1_pp", "n2_pp", "e_pp", "k1_hh", "k2_hh", "e_hh")
PH_KEYS = ("n_ph", "k_ph", "e_ph", "n_hp", "k_hp", "e_hp")
T2_KERNELS = ("t2_zero", "t2_amp", "t2_pandya", "t2_gemm1", "t2_gemm2", "t2_diag")
_NULL_I = C.cast(None, IP)
_NULL_D = C.cast(None, DP)


def _opt_i(a, keep):
    return _NULL_I if a is None else ip(keep.i(a))


def _fill_dims(kind, mode, na, nb):
    if kind == 0:
        return (na if mode != 2 else 0) + (nb if mode != 1 else 0)
    return na + (nb if mode != 1 else 0)


class _TwoBodyMixin:
    """Ladder (gII_prop_t) and ring (Pol_prop_t) bases and matrices of one
    channel through the C ABI; bases are dicts keyed as the reference's members."""

    def _basis(self, kind, g, J, dT, pif, icut, ntcut, n1p, n1h):
        L = lib()
        keep = _Keep()
        gs = prop_struct(g, self.ms.nspb, keep)
        a, b = C.c_int(), C.c_int()
        cnt = L.bcd_ladd_count if kind == 0 else L.bcd_ph_count
        check(cnt(self.h, C.byref(gs), J, dT, pif, icut, ntcut, _opt_i(n1p, keep), _opt_i(n1h, keep),
                  C.byref(a), C.byref(b)), "basis count")
        na, nb = a.value, b.value
        arr = [np.zeros(max(na, 1), np.int32), np.zeros(max(na, 1), np.int32), np.zeros(max(na, 1)),
               np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1))]
        fn = L.bcd_ladd_basis if kind == 0 else L.bcd_ph_basis
        check(fn(self.h, C.byref(gs), J, dT, pif, icut, ntcut, _opt_i(n1p, keep), _opt_i(n1h, keep), na, nb,
                 ip(arr[0]), ip(arr[1]), dp(arr[2]), ip(arr[3]), ip(arr[4]), dp(arr[5]), C.byref(a), C.byref(b)),
              "basis build")
        keys = LADD_KEYS if kind == 0 else PH_KEYS
        return {k: v[:(na if i < 3 else nb)] for i, (k, v) in enumerate(zip(keys, arr))}

    def ladd_basis(self, g, J, dT, pif, icut=0, ntcut=0, n1p=None, n1h=None):
        return self._basis(0, g, J, dT, pif, icut, ntcut, n1p, n1h)

    def ph_basis(self, g, J, dT, pif, icut=0, ntcut=0, n1p=None, n1h=None):
        return self._basis(1, g, J, dT, pif, icut, ntcut, n1p, n1h)

    def _fill(self, kind, g, J, dT, pif, basis, mode, add, ld, fill, raw):
        keys = LADD_KEYS if kind == 0 else PH_KEYS
        b = [basis[k] for k in keys]
        na, nb = len(b[2]), len(b[5])
        ndim = _fill_dims(kind, mode, na, nb)
        ld = max(ndim, 1) if ld is None else ld
        rows = max(ndim, 1)
        sab = np.full(rows * max(ld, 1), fill)
        fp = np.full(rows, fill)
        keep = _Keep()
        gs = prop_struct(g, self.ms.nspb, keep)
        z = lambda a, f: f(keep.i(a) if f is ip else keep.d(a)) if len(a) else (_NULL_I if f is ip else _NULL_D)
        fn = lib().bcd_ladd_fill if kind == 0 else lib().bcd_ph_fill
        rc = fn(self.h, C.byref(gs), J, dT, pif, na, z(b[0], ip), z(b[1], ip), z(b[2], dp), nb, z(b[3], ip),
                z(b[4], ip), z(b[5], dp), mode, dp(sab), ld, dp(fp), int(add))
        if raw:
            return rc
        check(rc, "bcd_ladd_fill" if kind == 0 else "bcd_ph_fill")
        return sab[:ndim * ld].reshape(ndim, ld), fp[:ndim]

    def ladd_fill(self, g, J, dT, pif, basis, mode, add_free_poles=True, ld=None, fill=0.0, raw=False):
        """gII_prop_t::Fill_LaddRPA_mtx: (matrix rows [ndim x ld], free poles), both
        from buffers prefilled with `fill`; raw=True returns the status code."""
        return self._fill(0, g, J, dT, pif, basis, mode, add_free_poles, ld, fill, raw)

    def ph_fill(self, g, J, dT, pif, basis, mode, add_free_poles=True, ld=None, fill=0.0, raw=False):
        """Pol_prop_t::Fill_phRPA_mtx, as ladd_fill."""
        return self._fill(1, g, J, dT, pif, basis, mode, add_free_poles, ld, fill, raw)


class TwoBody:
    """Every ladder and ring matrix of a propagator (bcd_t2), resident on the device."""

    def __init__(self, ctx, g: Propagator, ladd_mode=3, ring_mode=3, add_free_poles=False):
        L = lib()
        keep = _Keep()
        h = C.c_void_p()
        check(L.bcd_t2_create(ctx.h, C.byref(prop_struct(g, ctx.ms.nspb, keep)), ladd_mode, ring_mode,
                              int(add_free_poles), C.byref(h)), "bcd_t2_create")
        self.h = h
        self.ctx = ctx     # keeps the context alive
        self.n = check(L.bcd_t2_count(self.h), "bcd_t2_count")

    def close(self):
        if getattr(self, "h", None):
            lib().bcd_t2_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self, k):
        v = np.zeros(10, np.int32)
        check(lib().bcd_t2_info(self.h, k, ip(v)), "bcd_t2_info")
        return dict(zip(("kind", "J", "pi", "dT", "mode", "nf", "nb", "ld", "na", "nb_basis"), (int(x) for x in v)))

    def basis(self, k):
        """The basis of matrix k as a dict keyed like the reference's members."""
        inf = self.info(k)
        na, nb = inf["na"], inf["nb_basis"]
        arr = [np.zeros(max(na, 1), np.int32), np.zeros(max(na, 1), np.int32), np.zeros(max(na, 1)),
               np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1))]
        check(lib().bcd_t2_basis(self.h, k, ip(arr[0]), ip(arr[1]), dp(arr[2]), ip(arr[3]), ip(arr[4]),
                                 dp(arr[5])), "bcd_t2_basis")
        keys = LADD_KEYS if inf["kind"] == 0 else PH_KEYS
        return {key: v[:(na if i < 3 else nb)] for i, (key, v) in enumerate(zip(keys, arr))}

    def build(self):
        check(lib().bcd_t2_build(self.h), "bcd_t2_build")

    def matrix(self, k):
        inf = self.info(k)
        ndim, ld = inf["nf"] + inf["nb"], inf["ld"]
        sab = np.zeros(max(ndim, 1) * ld)
        fp = np.full(max(ndim, 1), np.nan)
        check(lib().bcd_t2_matrix(self.h, k, dp(sab), dp(fp)), "bcd_t2_matrix")
        return sab[:ndim * ld].reshape(ndim, ld), fp[:ndim]

    def set_profile(self, on: bool):
        check(lib().bcd_t2_set_profile(self.h, int(on)), "bcd_t2_set_profile")

    def kernel_times(self):
        out = {}
        for k, name in enumerate(T2_KERNELS):
            a, n = C.c_double(), C.c_long()
            check(lib().bcd_t2_kernel_times(self.h, k, C.byref(a), C.byref(n)), "bcd_t2_kernel_times")
            out[name] = (a.value, n.value)
        return out

    def work(self):
        w = np.zeros(3)
        check(lib().bcd_t2_work(self.h, dp(w), 3), "bcd_t2_work")
        return {"flops_gemm1": w[0], "flops_gemm2": w[1], "bytes_out": w[2]}


class Context(_TwoBodyMixin):
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

    def reduce_mtx_lncz(self, NA_out, NB_out, NPLS_out, Min, NA_in, NB_in, pivots, n_itr, noalloc=False):
        """Reduce_Mtx_Lncz: (return value, Mout rows, i_piv, normalised pivots).
        noalloc=True calls bcd_reduce_mtx_lncz_noalloc (no device allocation)."""
        Min = _d(Min)
        piv = _d(pivots).copy().reshape(-1)
        itr = _i(n_itr).copy()
        out = np.zeros(max(NPLS_out, 1) * (NA_out + NB_out))
        ipv = np.zeros(len(itr) + 2, np.int32)
        fn = lib().bcd_reduce_mtx_lncz_noalloc if noalloc else lib().bcd_reduce_mtx_lncz
        r = fn(self.h, NA_out, NB_out, NPLS_out, dp(out), ip(ipv), NA_in, NB_in,
                                      Min.shape[0], dp(Min.reshape(-1) if Min.size else np.zeros(1)), len(itr),
                                      dp(piv), ip(itr))
        return r, out[:max(NPLS_out, 0) * (NA_out + NB_out)].reshape(-1, NA_out + NB_out), ipv, \
            piv.reshape(len(itr), -1)

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


# k_lanczos tile shapes (threads per job, elements per thread) and the upper
# pole counts of the buckets the autotuner picks them for
LANC_SHAPES = [(64, 4), (64, 8), (128, 8), (192, 8), (256, 8), (128, 16), (192, 16), (256, 16), (384, 8), (192, 24),
               (256, 24), (384, 16), (384, 24), (768, 24), (1024, 24), (1024, 32), (1024, 40), (512, 16), (768, 16),
               (1024, 16), (512, 24), (768, 32)]
LANC_BUCKET_MAX = [256, 512, 1024, 2048, 4096, 8192, 16384, 24576, 32768, 40960]
LANC_BUCKETS = len(LANC_BUCKET_MAX)
# a second valid shape per bucket, different from the default one (its block
# sums are split differently): the rounding probe of a solve
LANC_ALT_SHAPES = [1, 2, 3, 7, 11, 13, 14, 15, 16, 16]
# the autotuner keeps the shape of the next larger occupied bucket for a
# bucket where that shape is within this fraction of the fastest (both
# buckets then run as one launch)
LANC_MERGE_TOL = 0.05


def solve_params(mode=2, hf_itr_max=-1, sc_itr_max=-1, lanczos=True, lanc_kind=1, lanc_n=50, conv_kind=1,
                 hf_conv=1e-8, sc_conv=1e-8, t_fac=0.0, ext_u1=1, dyson_solver=0, precision=0,
                 accel=0, hw=0.0, piv_round=0, mix_start=0.0) -> BcdSolveParams:
    """dyson_solver 0: arrowhead updates (default); 1: dense Householder + QL per sweep.
    precision 0: FP64 throughout; mixed Sigma(2) coupling contraction: 1 matrix cores, FP32 operands;
    2 matrix cores, BF16 operands (both FP32 accumulation); 3 vector units, FP32 operands, FP64 sums.
    accel 0: the reference's iteration; m = 1..8: Anderson mixing of the density over m iterates,
    starting once the density residual norm is below mix_start (<= 0: 1e-3).
    t_fac = U1body_fac and hw = htom of the kinetic term (hw * T_HO) * t_fac.
    piv_round 1: pivots rounded as the pivots file carries them (%18.6le)."""
    return BcdSolveParams(mode, hf_itr_max, sc_itr_max, int(lanczos), lanc_kind, lanc_n, conv_kind, hf_conv,
                          sc_conv, t_fac, ext_u1, dyson_solver, precision, accel, hw, mix_start, int(piv_round))


def pivot_round(x, device: int = 0) -> np.ndarray:
    """The pivots-file rounding (%18.6le read back) computed on the GPU."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    out = np.zeros_like(x)
    check(lib().bcd_pivot_round(device, dp(x), dp(out), x.size), "bcd_pivot_round")
    return out


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The learned start's network (bcd_mlp): n_in -> n_hidden -> n_hidden -> n_out
    with ReLU, BF16 weights on the matrix cores, FP32 accumulation, trained on
    the device."""

    def __init__(self, n_in=LEARN_IN, n_hidden=256, n_out=LEARN_OUT, seed=0, device: int = 0):
        h = C.c_void_p()
        check(lib().bcd_mlp_create(device, n_in, n_hidden, n_out, seed, C.byref(h)), "bcd_mlp_create")
        self.h = h
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        self.n_params = lib().bcd_mlp_num_params(h)

    def close(self):
        if getattr(self, "h", None):
            lib().bcd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        check(lib().bcd_mlp_get_params(self.h, _fp(p)), "bcd_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(lib().bcd_mlp_set_params(self.h, _fp(p)), "bcd_mlp_set_params")

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        check(lib().bcd_mlp_forward(self.h, x.shape[0], _fp(x), _fp(y)), "bcd_mlp_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        check(lib().bcd_mlp_loss_grad(self.h, x.shape[0], _fp(x), _fp(t), _fp(loss), _fp(g)), "bcd_mlp_loss_grad")
        return float(loss[0]), g

    def fit(self, x, t, epochs=100, batch=1024, lr=1e-3, seed=0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        hist = np.zeros(max(epochs, 1), np.float32)
        check(lib().bcd_mlp_fit(self.h, x.shape[0], _fp(x), _fp(t), epochs, batch, lr, seed, _fp(hist)),
              "bcd_mlp_fit")
        return hist[:epochs]


class Batch:
    """A device-resident batch of independent instances (bcd_batch)."""

    def __init__(self, systems, params: BcdSolveParams, g0=None, device: int = 0, lanc_n=None):
        """lanc_n: optional Lanczos length per instance (bcd_batch_create_ex)."""
        L = lib()
        n = len(systems)
        keep = _Keep()
        self.systems = systems
        msa = (BcdModspace * n)(*[modspace_struct(ms, keep) for ms, _ in systems])
        vpa = (BcdVpp * n)(*[vpp_struct(v, keep) for _, v in systems])
        gpa = None
        if g0 is not None:
            gpa = (BcdProp * n)(*[prop_struct(g, ms.nspb, keep) for g, (ms, _) in zip(g0, systems)])
        ln = None if lanc_n is None else ip(keep.i(np.asarray(lanc_n, np.int32)))
        h = C.c_void_p()
        check(L.bcd_batch_create_ex(device, n, msa, vpa, gpa, C.byref(params), ln, C.byref(h)),
              "bcd_batch_create_ex")
        self.h = h
        self.n = n

    @classmethod
    def streamed(cls, n: int, make, params: BcdSolveParams, lanc_n=None, threads: int = 8, device: int = 0):
        """A batch built one instance at a time (bcd_batch_begin / _add / _end):
        make(i) returns the (ModelSpace, Interaction) of instance i; up to
        `threads` instances are generated and packed at once, and each one's
        arrays are dropped once its interaction is in HBM, so batches whose
        interactions together outgrow host memory can be built.  Only the
        model spaces are kept (for result())."""
        from concurrent.futures import ThreadPoolExecutor
        L = lib()
        self = cls.__new__(cls)
        h = C.c_void_p()
        check(L.bcd_batch_begin(device, n, C.byref(params), C.byref(h)), "bcd_batch_begin")
        self.h = h
        self.n = n
        spaces = [None] * n

        def add(i):
            ms, vpp = make(i)
            keep = _Keep()
            rc = L.bcd_batch_add(h, i, C.byref(modspace_struct(ms, keep)), C.byref(vpp_struct(vpp, keep)), None,
                                 0 if lanc_n is None else int(lanc_n[i]))
            spaces[i] = ms
            return rc

        with ThreadPoolExecutor(max(1, threads)) as ex:
            for i, rc in enumerate(ex.map(add, range(n))):
                check(rc, f"bcd_batch_add({i})")
        check(L.bcd_batch_end(h), "bcd_batch_end")
        self.systems = [(ms, None) for ms in spaces]
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().bcd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_graphs(self, on: bool):
        """Replay full sweep chunks from a captured HIP graph (default) or launch every sweep."""
        check(lib().bcd_batch_set_graphs(self.h, int(on)), "bcd_batch_set_graphs")

    def set_profile(self, on: bool):
        check(lib().bcd_batch_set_profile(self.h, int(on)), "bcd_batch_set_profile")

    def set_params(self, params: BcdSolveParams):
        """Precision, density mixing, ItrMax and stopping rule of later solves."""
        check(lib().bcd_batch_set_params(self.h, C.byref(params)), "bcd_batch_set_params")

    def set_efermi(self, inst: int, ish: int, ef: float):
        """Force the Fermi energy of one subshell (Set_external_EFermi)."""
        check(lib().bcd_batch_set_efermi(self.h, inst, ish, ef), "bcd_batch_set_efermi")

    def solve(self):
        rc = lib().bcd_batch_solve(self.h)
        if rc < 0:
            raise BcdError(f"bcd_batch_solve failed: {rc} ({ERRORS.get(rc, 'reference error code')}): "
                           f"{self.error_detail()}")

    def error_detail(self) -> str:
        buf = C.create_string_buffer(4096)
        lib().bcd_batch_error_detail(self.h, buf, 4096)
        return buf.value.decode()

    def autotune(self):
        """Pick the Lanczos tile shape of every pole-count bucket during the next solve."""
        check(lib().bcd_batch_autotune(self.h), "bcd_batch_autotune")

    def lanczos_shapes(self):
        """(shape index per bucket, timed ms per (bucket, shape), 0 where not timed)."""
        sh = np.zeros(LANC_BUCKETS, np.int32)
        ms = np.zeros((LANC_BUCKETS, len(LANC_SHAPES)), np.float32)
        check(lib().bcd_batch_lanczos_shapes(self.h, ip(sh), ms.ctypes.data_as(C.POINTER(C.c_float)), LANC_BUCKETS),
              "bcd_batch_lanczos_shapes")
        return sh, ms

    def set_lanczos_shapes(self, shapes):
        sh = np.asarray(shapes, np.int32)
        check(lib().bcd_batch_set_lanczos_shapes(self.h, ip(sh), len(sh)), "bcd_batch_set_lanczos_shapes")

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

    def set_learned_start(self, net):
        """Seed the sc0 iteration with the network's predicted densities (None: off)."""
        self._net = net
        check(lib().bcd_batch_set_learned_start(self.h, net.h if net is not None else None),
              "bcd_batch_set_learned_start")

    def set_fallback_tol(self, tol: float):
        check(lib().bcd_batch_set_fallback_tol(self.h, float(tol)), "bcd_batch_set_fallback_tol")

    def record_learning(self, on: bool = True):
        check(lib().bcd_batch_record_learning(self.h, int(on)), "bcd_batch_record_learning")

    def learning_data(self):
        """(inputs [n_chan x 80], targets [n_chan x 32], instance of each channel) of the last solve."""
        n = check(lib().bcd_batch_num_channels(self.h), "bcd_batch_num_channels")
        f = np.zeros((n, LEARN_IN), np.float32)
        t = np.zeros((n, LEARN_OUT), np.float32)
        ci = np.zeros(n, np.int32)
        check(lib().bcd_batch_learning_data(self.h, _fp(f), _fp(t), ip(ci), n), "bcd_batch_learning_data")
        return f, t, ci

    def start_info(self, i):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        d = C.c_double()
        check(lib().bcd_batch_start_info(self.h, i, C.byref(a), C.byref(b), C.byref(c), C.byref(d)),
              "bcd_batch_start_info")
        return dict(seeded=bool(a.value), fell_back=bool(b.value), first_its=c.value, residual=d.value)

    def kernel_times(self):
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            a, b = C.c_double(), C.c_long()
            check(lib().bcd_batch_kernel_times(self.h, k, C.byref(a), C.byref(b)), "bcd_batch_kernel_times")
            out[name] = (a.value, b.value)
        return out

    def work(self):
        w = np.zeros(17)
        check(lib().bcd_batch_work(self.h, dp(w), 17), "bcd_batch_work")
        phases = ["load_sort", "deflation", "roots", "loewner", "eigenvectors", "reorder", "output", "norms"]
        return dict(coupling_flops=w[0], lanczos_bytes=w[1], dyson_flops=w[2], n_eig=w[3], sum_n3=w[4],
                    arrow_flops=w[5], poles_flops=w[6], arrow_phase_ms=dict(zip(phases, w[7:15])),
                    sigma2_poles=int(w[15]), lanczos_poles=int(w[16]))
