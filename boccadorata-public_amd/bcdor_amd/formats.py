"""BoccaDorata file formats: model space, interaction (.bcd), propagator, pivots, bcdwk.

The layouts are the reference's, unchanged, so files move between this
package and the reference both ways.  Numbers are written with %.17g so that
reading them back (the reference reads with %lf) reproduces every bit.
"""
from __future__ import annotations

import dataclasses
import os
import struct

import numpy as np

from .synth import ModelSpace, Interaction, subshell_name


# --------------------------------------------------------------- model space
def write_model_space(path: str, ms: ModelSpace) -> None:
    """Layout of ModSpace_t::write / ModSpace_t(char*) (Mod_Space.cpp:415-504, 507-544)."""
    orb0 = ms.orb0
    with open(path, "w") as f:
        f.write("# Definition of the model space by listing its s.p. orbitals\n")
        f.write("#------------------------------------------------------------\n\n")
        f.write("# number of (ilj\\pi) subshells and total number of orbitals:\n")
        f.write(f" {ms.nsubsh:7d} {ms.norb:7d}\n")
        for i in range(ms.nsubsh):
            nh = int(ms.orb_F[orb0[i]:orb0[i] + ms.no[i]].sum())
            f.write(f"\n# subshell:  {ms.names[i]} \n")
            f.write(f" {ms.l[i]:7d} {ms.j2[i]:7d} {ms.ip[i]:7d} {ms.ch[i]:7d}\n")
            f.write(f" {ms.no[i]:7d} {ms.no[i] - nh:7d} {nh:7d}\n")
            f.write("# principal q.#,  p/h character and sp energies:\n")
            for k in range(ms.no[i]):
                o = orb0[i] + k
                f.write(f" {ms.orb_n[o]:7d} {ms.orb_F[o]:7d} {ms.orb_e[o]:.17g}     # {ms.orb_n[o]}{ms.names[i]}\n")


def read_model_space(path: str, hw: float = 0.0) -> ModelSpace:
    with open(path) as f:
        lines = f.read().split("\n")
    ns, norb = (int(x) for x in lines[4].split()[:2])
    p = 5
    l_, j2_, ip_, ch_, no_, names, on, oF, oe = [], [], [], [], [], [], [], [], []
    for _ in range(ns):
        p += 1
        names.append(lines[p].split()[2]); p += 1
        l, j2, ip, ch = (int(x) for x in lines[p].split()[:4]); p += 1
        no = int(lines[p].split()[0]); p += 2
        l_.append(l); j2_.append(j2); ip_.append(ip); ch_.append(ch); no_.append(no)
        for _k in range(no):
            t = lines[p].split(); p += 1
            on.append(int(t[0])); oF.append(int(t[1])); oe.append(float(t[2]))
    a = lambda x: np.asarray(x, np.int32)
    return ModelSpace(a(l_), a(j2_), a(ip_), a(ch_), a(no_), a(on), a(oF), np.asarray(oe, np.float64), names, hw)


# --------------------------------------------------------------- interaction
def write_vpp_bcd(path: str, ms: ModelSpace, vpp: Interaction) -> None:
    """Text layout read by VppInt_t::read_ascii (Vpp_potential.cpp:281-374)."""
    sh = ms.orb_sh
    q = lambda o: np.stack([ms.orb_n[o], ms.j2[sh[o]], ms.ip[sh[o]], ms.ch[sh[o]]], axis=1)
    cols = np.concatenate([q(vpp.a), q(vpp.b), q(vpp.c), q(vpp.d), vpp.J[:, None]], axis=1)
    with open(path, "w") as f:
        f.write(f" {vpp.n}  {vpp.n}  # tot number of mtx el. stored and to be allocated\n")
        for row, v in zip(cols.tolist(), vpp.v.tolist()):
            f.write("%d %d %d %d  %d %d %d %d  %d %d %d %d %d %d %d %d  %d %.17g\n" % (*row, v))


# ---------------------------------------------------------------- propagator
@dataclasses.dataclass
class Propagator:
    """Lehmann fragments per subshell (SpProp_t): energies, Z, overlaps [n, nspb]."""
    np_: np.ndarray       # [nsubsh] quasiparticles per subshell
    nh: np.ndarray        # [nsubsh] quasiholes per subshell
    ep: np.ndarray
    zp: np.ndarray
    fqp: np.ndarray       # [n_qp, nspb]
    eh: np.ndarray
    zh: np.ndarray
    fqh: np.ndarray

    @property
    def nspb(self) -> int:
        return int(self.fqp.shape[1])

    def offsets(self):
        op = np.concatenate([[0], np.cumsum(self.np_)[:-1]]).astype(np.int64)
        oh = np.concatenate([[0], np.cumsum(self.nh)[:-1]]).astype(np.int64)
        return op, oh

    def subshell(self, i: int):
        op, oh = self.offsets()
        sp = slice(op[i], op[i] + self.np_[i]); sh = slice(oh[i], oh[i] + self.nh[i])
        return (self.ep[sp], self.zp[sp], self.fqp[sp], self.eh[sh], self.zh[sh], self.fqh[sh])


def bare_propagator(ms: ModelSpace) -> Propagator:
    """SpProp_t(ModSpace_t*, i_xe=1) (Sp_Propagator.cpp:292-378): holes at spe-10."""
    nspb = ms.nspb
    np_, nh, ep, eh, fp, fh = [], [], [], [], [], []
    orb0 = ms.orb0
    for i in range(ms.nsubsh):
        cp = ch = 0
        for k in range(ms.no[i]):
            o = orb0[i] + k
            vec = np.zeros(nspb); vec[ms.orb_n[o]] = 1.0
            if ms.orb_F[o] == 1:
                eh.append(ms.orb_e[o] - 10.0); fh.append(vec); ch += 1
            else:
                ep.append(ms.orb_e[o]); fp.append(vec); cp += 1
        np_.append(cp); nh.append(ch)
    fp = np.asarray(fp).reshape(-1, nspb); fh = np.asarray(fh).reshape(-1, nspb)
    return Propagator(np.asarray(np_, np.int32), np.asarray(nh, np.int32), np.asarray(ep), np.ones(len(ep)),
                      fp, np.asarray(eh), np.ones(len(eh)), fh)


def write_propagator(path: str, ms: ModelSpace, g: Propagator, full_precision: bool = True) -> None:
    """SpProp_t::write layout (Sp_Propagator.cpp:613-674), wcode 0."""
    fe = (lambda x: f" {x:.17g} ") if full_precision else (lambda x: f" {x:15.6f} ")
    fz = (lambda x: f" {100.0 * x:.17g}   ") if full_precision else (lambda x: f" {100.0 * x:12.3f}   ")
    fx = (lambda x: f" {x:.17g} ") if full_precision else (lambda x: f" {x:18.6e} ")
    with open(path, "w") as f:
        f.write("#  Quasi- particle and hole fragments of the sp propagator\n")
        f.write("#----------------------------------------------------------\n")
        f.write("\n# number of (ilj\\pi) subshells, max n. of radial orbitals:\n")
        f.write(f"# {ms.nsubsh:7d} {g.nspb:7d} {0:7d}\n")
        f.write("\n# Total numbers of qp and qh stored here:\n")
        f.write(f"# {int(g.np_.sum()):7d} {int(g.nh.sum()):7d}\n")
        for i in range(ms.nsubsh):
            ep, zp, fqp, eh, zh, fqh = g.subshell(i)
            f.write("\n# Subshell:\n")
            f.write(f"# {ms.names[i]} \n")
            f.write(f"# {int(g.np_[i]):7d}  # -> tot n. of quasiparticles\n")
            for j in range(len(ep)):
                f.write(fe(ep[j]) + fz(zp[j]) + "".join(fx(x) for x in fqp[j, :ms.no[i]]) + "\n")
            f.write(f"# {int(g.nh[i]):7d}  # -> tot n. of quasiholes\n")
            for j in range(len(eh)):
                f.write(fe(eh[j]) + fz(zh[j]) + "".join(fx(x) for x in fqh[j, :ms.no[i]]) + "\n")


def read_propagator(path: str, ms: ModelSpace) -> Propagator:
    """SpProp_t(char*, ModSpace_t*) (Sp_Propagator.cpp:387-610), flags 0."""
    with open(path) as f:
        lines = f.read().split("\n")
    nspb = ms.nspb
    p = 8
    np_, nh, ep, zp, fp, eh, zh, fh = [], [], [], [], [], [], [], []
    for i in range(ms.nsubsh):
        p += 3
        n = int(lines[p].lstrip("#").split()[0]); p += 1
        for _ in range(n):
            t = lines[p].split(); p += 1
            ep.append(float(t[0])); zp.append(float(t[1]) / 100.0)
            x = np.zeros(nspb); vals = [float(s) for s in t[2:2 + ms.no[i]]]
            x[:len(vals)] = vals; fp.append(x)
        np_.append(n)
        n = int(lines[p].lstrip("#").split()[0]); p += 1
        for _ in range(n):
            t = lines[p].split(); p += 1
            eh.append(float(t[0])); zh.append(float(t[1]) / 100.0)
            x = np.zeros(nspb); vals = [float(s) for s in t[2:2 + ms.no[i]]]
            x[:len(vals)] = vals; fh.append(x)
        nh.append(n)
    return Propagator(np.asarray(np_, np.int32), np.asarray(nh, np.int32), np.asarray(ep), np.asarray(zp),
                      np.asarray(fp).reshape(-1, nspb), np.asarray(eh), np.asarray(zh),
                      np.asarray(fh).reshape(-1, nspb))


# -------------------------------------------------------------------- pivots
_PIV_HEAD = "#   Template pivots file\n#\n# number of (ilj\\pi) subshells, max n. of radial orbitals:\n"


def dys_pivots(path: str, ms: ModelSpace, g: Propagator, n_fw: int, n_bk: int, z_cut: float = 60.0) -> None:
    """The pivots file SpProp_t::DysPivots writes from g (Sp_Propagator.cpp:
    1225-1289), byte for byte: per subshell the header line with npvts = the
    subshell's orbit count, then its quasiholes and then its quasiparticles
    with 100*Z >= z_cut, each side by ascending energy (a fragment whose
    energy equals the one picked before it is skipped), every pivot with
    iteration counts n_fw n_bk, its components as %18.6le and its energy as a
    comment."""
    with open(path, "w") as f:
        f.write(_PIV_HEAD)
        f.write("# %7i %7i\n\n" % (ms.nsubsh, g.nspb))
        for i in range(ms.nsubsh):
            f.write("\n\nshell,npvts:   %i   %i  # %s \n" % (i, ms.no[i], ms.names[i]))
            ep, zp, fqp, eh, zh, fqh = g.subshell(i)
            for e, z, x, tag in ((eh, zh, fqh, "e-"), (ep, zp, fqp, "e+")):
                x2 = -1.0e10
                while True:
                    x1, pick = 1.0e10, -1
                    for j in range(len(e)):
                        if e[j] < x1 and e[j] > x2 and 100.0 * z[j] >= z_cut:
                            x1, pick = e[j], j
                    if pick < 0:
                        break
                    x2 = x1
                    f.write(" %3i  %3i  " % (n_fw, n_bk))
                    f.write("".join(" %18.6e " % v for v in x[pick, :ms.no[i]]))
                    f.write("   #    %s = %13.6f\n" % (tag, e[pick]))


def write_pivots(path: str, ms: ModelSpace, pivots: dict, full_precision: bool = True) -> None:
    """A pivots file in the layout Make_LancPivots reads (Miscellanea.cpp:
    334-390) with explicit per-pivot iteration counts.
    pivots[ish] = (vectors [npv, no], itr_fw [npv], itr_bk [npv]); the header
    gives npv.  Components are written with %.17g so that the reader gets
    every bit back, or with DysPivots' own %18.6le (full_precision=False)."""
    fx = (lambda x: f" {x:.17g} ") if full_precision else (lambda x: " %18.6e " % x)
    with open(path, "w") as f:
        f.write(_PIV_HEAD)
        f.write(f"# {ms.nsubsh:7d} {ms.nspb:7d}\n\n")
        for ish in range(ms.nsubsh):
            vec, ifw, ibk = pivots[ish]
            f.write(f"\n\nshell,npvts:   {ish}   {len(ifw)}  # {ms.names[ish]} \n")
            for k in range(len(ifw)):
                f.write(f" {int(ifw[k]):3d}  {int(ibk[k]):3d}  " + "".join(fx(x) for x in vec[k]) + "\n")


def read_pivots(path: str, ms: ModelSpace) -> dict:
    """Make_LancPivots (Miscellanea.cpp:334-390): after the 'shell,npvts: ish n'
    line of a subshell it reads n lines, keeps those that start with two
    integers (the iteration counts fw and bk) and takes the subshell's orbit
    count of numbers after them.  Returns ish -> (vectors, itr_fw, itr_bk)."""
    out = {}
    with open(path) as f:
        lines = f.read().split("\n")
    # each subshell's block is read on its own from its header on (the
    # reference reopens the file per subshell), so a block that announces
    # more lines than it has reads on into the next one
    for h, t in enumerate(lines):
        if not t.startswith("shell,npvts:"):
            continue
        ish, n = (int(x) for x in t.split(":")[1].split()[:2])
        if ish in out:
            continue
        vec, ifw, ibk = [], [], []
        p = h + 1
        for _ in range(n):
            if p >= len(lines):
                break
            s = lines[p].split()
            p += 1
            try:
                a, b = int(s[0]), int(s[1])
            except (ValueError, IndexError):
                continue
            ifw.append(a); ibk.append(b)
            vec.append([float(x) for x in s[2:2 + ms.no[ish]]])
        out[ish] = (np.asarray(vec, np.float64).reshape(len(ifw), ms.no[ish]), np.asarray(ifw, np.int32),
                    np.asarray(ibk, np.int32))
    return out


# ---------------------------------------------------------------- bcdwk Sigma
@dataclasses.dataclass
class SelfEnergyFile:
    """One channel's dynamic self-energy as Wide_self_energy_t stores it
    (SlfEn_classes.hh:14-42): pole rows [E, b or sum m^2, g..., m(0..norb-1)] of
    width ntot_den + ntot_orb per side, and i_piv = [n_pvt, 1, starts...] for a
    Lanczos-reduced side (empty for a diagonal side).  The header fields
    describe the channel (Lambda = 0: one partial wave, the subshell itself)."""
    ish: int
    ntot_den: int
    ntot_orb: int
    fw: np.ndarray        # [nfw, ntot_den + ntot_orb]
    bk: np.ndarray
    i_piv_fw: np.ndarray
    i_piv_bk: np.ndarray
    j2: int = 0
    ip: int = 0
    ch: int = 0
    alloc_fw: int = -1    # N_PLS_FW_alloc (-1: max(1, row count), as Allocate_Dynamic_SlfEn sets it)
    alloc_bk: int = -1
    i_pwv: tuple = ()
    nst_pwv: tuple = ()
    norb_pwv: tuple = ()
    isp_msp: tuple = ()   # model-space orbit of each row column
    mor_n: tuple = ()     # radial quantum number of each of those orbits

    @property
    def ntot_dim(self) -> int:
        return self.ntot_den + self.ntot_orb


def selfenergy_for_channel(ms: ModelSpace, ish: int, fw: np.ndarray, bk: np.ndarray, ntot_den: int,
                           i_piv_fw=None, i_piv_bk=None) -> SelfEnergyFile:
    """A SelfEnergyFile for subshell ish of ms with the header tables that
    set_clj(ish) fills for Lambda = 0 (Self_energy_ext_Op_a.cpp:223-282)."""
    no = int(ms.no[ish])
    o0 = int(ms.orb0[ish])
    orbs = tuple(range(o0, o0 + no))
    empty = np.zeros(0, np.int32)
    return SelfEnergyFile(ish, int(ntot_den), no, np.asarray(fw, np.float64).reshape(-1, ntot_den + no),
                          np.asarray(bk, np.float64).reshape(-1, ntot_den + no),
                          empty if i_piv_fw is None else np.asarray(i_piv_fw, np.int32),
                          empty if i_piv_bk is None else np.asarray(i_piv_bk, np.int32),
                          int(ms.j2[ish]), int(ms.ip[ish]), int(ms.ch[ish]), -1, -1, (ish,), (0,), (no,), orbs,
                          tuple(int(ms.orb_n[k]) for k in orbs))


def selfenergy_filename(storedir: str, label: str, j2: int, ip: int, ch: int, fb: str = "", order: str = "Dyn",
                        lam: int = 0, parity: int = 0, charge: int = 0) -> str:
    """Wide_self_energy_t::get_filename (Self_energy_ext_Op_d.cpp:43-52):
    <dir>/<label>_<order>[_T=<lam><+|-><charge>]_J2pi=<2j><+|->_dt=<ch><fb>.bin"""
    sign = "+-"
    stem = f"{label}_{order}" if (lam, parity, charge) == (0, 0, 0) else \
        f"{label}_{order}_T={lam}{sign[parity]}_{charge}"
    return os.path.join(storedir, f"{stem}_J2pi={j2}{sign[ip]}_dt={ch}{fb}.bin")


def _piv_len(ipv) -> int:
    return int(ipv[0]) if len(ipv) else 0


def _header_ints(sef: SelfEnergyFile, nfw: int, nbk: int) -> bytes:
    # Allocate_Dynamic_SlfEn (Self_energy_ext_Op_a.cpp:186-187) keeps at least one row
    afw = sef.alloc_fw if sef.alloc_fw >= 0 else max(1, sef.fw.shape[0])
    abk = sef.alloc_bk if sef.alloc_bk >= 0 else max(1, sef.bk.shape[0])
    npw = len(sef.i_pwv)
    h = [nfw, nbk, afw, abk, sef.ish, sef.ntot_den, sef.ntot_orb, sef.ntot_dim, npw, sef.j2, sef.ip, sef.ch]
    tab = list(sef.i_pwv) + list(sef.nst_pwv) + list(sef.norb_pwv) + list(sef.isp_msp) + list(sef.mor_n)
    if len(sef.isp_msp) != sef.ntot_orb or len(sef.mor_n) != sef.ntot_orb:
        raise ValueError("isp_msp and mor_n need one entry per orbit column")
    return struct.pack(f"<{12 + len(tab)}i", *h, *tab)


def write_bcdwk(path: str, sef: SelfEnergyFile, i_fwbk: int = 0) -> None:
    """Wide_self_energy_t::save_SelfEn_bin (Self_energy_ext_Op_d.cpp:55-133):
    12-int header and tables, the 3-int Lanczos descriptor (longest i_piv + 3,
    and whether each side is one Lanczos block), then per side the rows and,
    for a reduced side, [N_PLS, i_piv[0 .. i_piv[0]+1]].  i_fwbk > 0 writes the
    forward side only, < 0 the backward side only."""
    nfw = sef.fw.shape[0] if i_fwbk >= 0 else 0
    nbk = sef.bk.shape[0] if i_fwbk <= 0 else 0
    pf, pb = _piv_len(sef.i_piv_fw), _piv_len(sef.i_piv_bk)
    n1 = 0
    if pf > 0:
        n1 = max(n1, pf + 3)
    if pb > 0:
        n1 = max(n1, pb + 3)
    out = [_header_ints(sef, nfw, nbk), struct.pack("<3i", n1, int(pf > 0), int(pb > 0))]
    if i_fwbk >= 0:
        out.append(np.ascontiguousarray(sef.fw, "<f8").tobytes())
        if pf > 0:
            out.append(struct.pack(f"<{pf + 3}i", sef.fw.shape[0], *[int(x) for x in sef.i_piv_fw[:pf + 2]]))
    if i_fwbk <= 0:
        out.append(np.ascontiguousarray(sef.bk, "<f8").tobytes())
        if pb > 0:
            out.append(struct.pack(f"<{pb + 3}i", sef.bk.shape[0], *[int(x) for x in sef.i_piv_bk[:pb + 2]]))
    with open(path, "wb") as f:
        f.write(b"".join(out))


def _read_header(buf: bytes, off: int = 0):
    h = struct.unpack_from("<12i", buf, off)
    npw, norb = h[8], h[6]
    if npw < 0 or norb < 0 or h[7] != h[5] + h[6]:
        raise ValueError("not a bcdwk self-energy header")
    tab = struct.unpack_from(f"<{3 * npw + 2 * norb}i", buf, off + 48)
    return h, tab, off + 48 + 4 * len(tab)


def _sef_from_header(h, tab, fw, bk, ipf, ipb) -> SelfEnergyFile:
    npw, norb = h[8], h[6]
    return SelfEnergyFile(h[4], h[5], h[6], fw, bk, ipf, ipb, h[9], h[10], h[11], h[2], h[3],
                          tuple(tab[:npw]), tuple(tab[npw:2 * npw]), tuple(tab[2 * npw:3 * npw]),
                          tuple(tab[3 * npw:3 * npw + norb]), tuple(tab[3 * npw + norb:]))


def _read_side(buf, off, n, ndim, reduced):
    rows = np.frombuffer(buf, "<f8", n * ndim, off).reshape(n, ndim).copy()
    off += 8 * n * ndim
    ipv = np.zeros(0, np.int32)
    if reduced:
        # load_SelfEn_bin_c (Self_energy_ext_Op_d.cpp:183-193): N_PLS, then i_piv
        _npls, npv = struct.unpack_from("<2i", buf, off)
        if npv < 0 or npv + 2 > ndim:
            raise ValueError("bad pivot block in bcdwk file")
        ipv = np.asarray(struct.unpack_from(f"<{npv + 2}i", buf, off + 4), np.int32)
        off += 4 * (npv + 3)
    return rows, ipv, off


def read_bcdwk(path: str) -> SelfEnergyFile:
    """The layout write_bcdwk (save_SelfEn_bin) produces, as load_SelfEn_bin
    reads it (Self_energy_ext_Op_d.cpp:195-244); a side written empty comes
    back with no rows.  A one-side file (i_fwbk != 0) keeps, in its 3-int
    descriptor, the block flag and pivot length of the side it leaves out, as
    save_SelfEn_bin writes them; only the written side's pivot block is read."""
    with open(path, "rb") as f:
        buf = f.read()
    h, tab, off = _read_header(buf)
    npiv_max, bfw, bbk = struct.unpack_from("<3i", buf, off)
    off += 12
    if bfw not in (0, 1) or bbk not in (0, 1):
        raise ValueError(f"{path}: more than one Lanczos block per side is not supported")
    ndim = h[7]
    fw, ipf, off = _read_side(buf, off, h[0], ndim, bfw == 1 and h[0] > 0)
    bk, ipb, off = _read_side(buf, off, h[1], ndim, bbk == 1 and h[1] > 0)
    if off != len(buf):
        raise ValueError(f"{path}: {len(buf) - off} trailing bytes")
    return _sef_from_header(h, tab, fw, bk, ipf, ipb)


def read_bcdwk_split(path_fw: str, path_bk: str) -> SelfEnergyFile:
    """The split form load_SelfEn_bin falls back to (Self_energy_ext_Op_d.cpp:
    245-290): forward rows from the -Fw file, backward rows from the -Bk file,
    each file in the write_bcdwk layout with one side."""
    a, b = read_bcdwk(path_fw), read_bcdwk(path_bk)
    if (a.ish, a.ntot_dim) != (b.ish, b.ntot_dim):
        raise ValueError("forward and backward files belong to different channels")
    a.bk, a.i_piv_bk, a.alloc_bk = b.bk, b.i_piv_bk, b.alloc_bk
    return a


def read_bcdwk_old(path: str) -> SelfEnergyFile:
    """Older bcdwk layout shipped in the reference's examples/bcdwk: the header
    of save_SelfEn_bin (Self_energy_ext_Op_d.cpp:55-133) without the 3-int
    Lanczos descriptor; after each side's rows, [n_piv, i_piv[NTOT_DIM]]."""
    with open(path, "rb") as f:
        buf = f.read()
    h, tab, off = _read_header(buf)
    nfw, nbk, ndim = h[0], h[1], h[7]
    fw = np.frombuffer(buf, "<f8", nfw * ndim, off).reshape(nfw, ndim).copy(); off += 8 * nfw * ndim
    ipf = np.asarray(struct.unpack_from(f"<{ndim}i", buf, off + 4), np.int32); off += 4 * (1 + ndim)
    bk = np.frombuffer(buf, "<f8", nbk * ndim, off).reshape(nbk, ndim).copy(); off += 8 * nbk * ndim
    ipb = np.asarray(struct.unpack_from(f"<{ndim}i", buf, off + 4), np.int32); off += 4 * (1 + ndim)
    if off != len(buf):
        raise ValueError(f"{path}: {len(buf) - off} trailing bytes, not the old bcdwk layout")
    return _sef_from_header(h, tab, fw, bk, ipf, ipb)


def read_bcdwk_any(path: str) -> SelfEnergyFile:
    """Either layout: the current one (write_bcdwk), else the older one."""
    try:
        return read_bcdwk(path)
    except (ValueError, struct.error):
        return read_bcdwk_old(path)
