"""BoccaDorata file formats: model space, interaction (.bcd), propagator, pivots, bcdwk.

The layouts are the reference's, unchanged, so files move between this
package and the reference both ways.  Numbers are written with %.17g so that
reading them back (the reference reads with %lf) reproduces every bit.
"""
from __future__ import annotations

import dataclasses
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
def write_pivots(path: str, ms: ModelSpace, pivots: dict) -> None:
    """Pivot blocks as SpProp_t::DysPivots writes and Make_LancPivots reads
    (Sp_Propagator.cpp:1225-1289, Miscellanea.cpp:334-390).
    pivots[ish] = (vectors [npv, no], itr_fw [npv], itr_bk [npv])."""
    with open(path, "w") as f:
        f.write("#   Template pivots file\n#\n# number of (ilj\\pi) subshells, max n. of radial orbitals:\n")
        f.write(f"# {ms.nsubsh:7d} {ms.nspb:7d}\n\n")
        for ish in range(ms.nsubsh):
            vec, ifw, ibk = pivots[ish]
            f.write(f"\n\nshell,npvts:   {ish}   {len(ifw)}  # {ms.names[ish]} \n")
            for k in range(len(ifw)):
                f.write(f" {int(ifw[k]):3d}  {int(ibk[k]):3d}  " + "".join(f" {x:.17g} " for x in vec[k]) + "\n")


def read_pivots(path: str, ms: ModelSpace) -> dict:
    out = {}
    with open(path) as f:
        lines = f.read().split("\n")
    p = 0
    while p < len(lines):
        t = lines[p]
        if t.startswith("shell,npvts:"):
            ish, n = (int(x) for x in t.split(":")[1].split()[:2])
            vec, ifw, ibk = [], [], []
            for k in range(n):
                s = lines[p + 1 + k].split("#")[0].split()
                ifw.append(int(s[0])); ibk.append(int(s[1]))
                vec.append([float(x) for x in s[2:2 + ms.no[ish]]])
            out[ish] = (np.asarray(vec, np.float64).reshape(n, ms.no[ish]), np.asarray(ifw, np.int32),
                        np.asarray(ibk, np.int32))
            p += n
        p += 1
    return out


# ---------------------------------------------------------------- bcdwk Sigma
@dataclasses.dataclass
class SelfEnergyFile:
    ish: int
    ntot_den: int
    ntot_orb: int
    fw: np.ndarray        # [nfw, ntot_den + ntot_orb]
    bk: np.ndarray
    i_piv_fw: np.ndarray
    i_piv_bk: np.ndarray


def read_bcdwk_old(path: str) -> SelfEnergyFile:
    """Older bcdwk layout shipped in the reference's examples/bcdwk: the header
    of save_SelfEn_bin (Self_energy_ext_Op_d.cpp:55-133) without the 3-int
    Lanczos descriptor; after each side's rows, [n_piv, i_piv[NTOT_DIM]]."""
    with open(path, "rb") as f:
        buf = f.read()
    ints = lambda off, n: struct.unpack_from(f"<{n}i", buf, off)
    h = ints(0, 12)
    nfw, nbk, _afw, _abk, ish, den, norb, ndim, npwv = h[:9]
    off = 48 + 4 * (3 * npwv + 2 * norb)
    fw = np.frombuffer(buf, "<f8", nfw * ndim, off).reshape(nfw, ndim).copy(); off += 8 * nfw * ndim
    _npf = ints(off, 1)[0]; ipf = np.asarray(ints(off + 4, ndim), np.int32); off += 4 * (1 + ndim)
    bk = np.frombuffer(buf, "<f8", nbk * ndim, off).reshape(nbk, ndim).copy(); off += 8 * nbk * ndim
    _npb = ints(off, 1)[0]; ipb = np.asarray(ints(off + 4, ndim), np.int32); off += 4 * (1 + ndim)
    if off != len(buf):
        raise ValueError(f"{path}: {len(buf) - off} trailing bytes, not the old bcdwk layout")
    return SelfEnergyFile(ish, den, norb, fw, bk, ipf, ipb)
