"""Synthetic closed-shell systems for the SCGF hot path.

A system is a two-species harmonic-oscillator j-scheme model space truncated
at Nmax major shells, single-particle energies hw*(2n+l+3/2) plus Gaussian
noise, and a random real symmetric J-coupled two-body interaction whose
spread decays with the unperturbed energy difference of the two pairs.

Generation order (fixed, so every consumer of a seed sees the same numbers):
  1. one N(0, spe_noise^2) draw per orbit, in orbit order;
  2. for J = 0..max(2j), parity 0..1, pair charge 0..2: the allowed pairs
     a <= b (triangle(ja, jb, J), (la+lb)%2 == parity, Pauli: a == b only for
     even J) in (a, b) order, then one N(0,1) draw per unique bra <= ket,
     scaled by sigma*exp(-|E_ab - E_cd|/40).
Subshell order and naming follow the reference model-space generator
(Mod_Space.cpp:309-412): per species l = 0..Nmax with j = l-1/2 before
j = l+1/2; neutrons (ch 0) then protons (ch 1).  0s1/2, 0p1/2, 0p3/2 are
filled (8 + 8 fermions).
"""
from __future__ import annotations

import dataclasses

import numpy as np

L_NAMES = "spdfghijklmnopqrstuvwxyz"


@dataclasses.dataclass
class ModelSpace:
    l: np.ndarray        # [nsubsh]
    j2: np.ndarray
    ip: np.ndarray
    ch: np.ndarray
    no: np.ndarray
    orb_n: np.ndarray    # [norb]
    orb_F: np.ndarray
    orb_e: np.ndarray
    names: list
    hw: float = 0.0

    @property
    def nsubsh(self) -> int:
        return int(self.l.size)

    @property
    def norb(self) -> int:
        return int(self.orb_n.size)

    @property
    def orb0(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.no)[:-1]]).astype(np.int32)

    @property
    def orb_sh(self) -> np.ndarray:
        return np.repeat(np.arange(self.nsubsh, dtype=np.int32), self.no)

    @property
    def nspb(self) -> int:
        return int(self.no.max())


@dataclasses.dataclass
class Interaction:
    """Unique J-coupled matrix elements <ab;J|V|cd;J>, a<=b, c<=d, (ab)<=(cd)."""
    a: np.ndarray
    b: np.ndarray
    c: np.ndarray
    d: np.ndarray
    J: np.ndarray
    v: np.ndarray

    @property
    def n(self) -> int:
        return int(self.v.size)

    def keys(self):
        """VppInt_t packed keys (Vpp_potential.cpp:82): ((i1*1024+i2)*128+J)."""
        iab = (self.a.astype(np.int64) * 1024 + self.b) * 128 + self.J
        icd = (self.c.astype(np.int64) * 1024 + self.d) * 128 + self.J
        return iab.astype(np.int32), icd.astype(np.int32)


def subshell_name(l: int, j2: int, ch: int) -> str:
    return ("v" if ch == 0 else "p") + "_" + L_NAMES[l] + f"{j2}/2"


def make_model_space(nmax: int, hw: float, rng: np.random.Generator, spe_noise: float = 0.5,
                     filled=((0, 1), (1, 1), (1, 3))) -> ModelSpace:
    l_, j2_, ip_, ch_, no_, names = [], [], [], [], [], []
    for ch in (0, 1):
        for nlj in range(2 * nmax + 1):
            l = (nlj + 1) // 2
            if nmax - l < 0:
                continue
            j2 = 2 * (nlj - l) + 1
            l_.append(l); j2_.append(j2); ip_.append(l % 2); ch_.append(ch)
            no_.append(1 + (nmax - l) // 2)
            names.append(subshell_name(l, j2, ch))
    orb_n, orb_F, base = [], [], []
    for l, j2, no in zip(l_, j2_, no_):
        for n in range(no):
            orb_n.append(n)
            orb_F.append(1 if (n == 0 and (l, j2) in filled) else 0)
            base.append(hw * (2 * n + l + 1.5))
    base = np.asarray(base, dtype=np.float64)
    noise = rng.standard_normal(base.size) * spe_noise
    ms = ModelSpace(np.asarray(l_, np.int32), np.asarray(j2_, np.int32), np.asarray(ip_, np.int32),
                    np.asarray(ch_, np.int32), np.asarray(no_, np.int32),
                    np.asarray(orb_n, np.int32), np.asarray(orb_F, np.int32), base + noise, names, hw)
    return ms


def pair_channels(ms: ModelSpace):
    """Yield (J, parity, charge, pairs[a,b]) for every non-empty two-body channel."""
    sh = ms.orb_sh
    l = ms.l[sh]; j2 = ms.j2[sh]; ch = ms.ch[sh]
    norb = ms.norb
    ia, ib = np.triu_indices(norb)
    jmax = int(ms.j2.max())
    for J in range(jmax + 1):
        tri = (j2[ia] + j2[ib] >= 2 * J) & (np.abs(j2[ia] - j2[ib]) <= 2 * J)
        pauli = (ia != ib) | (J % 2 == 0)
        for par in (0, 1):
            for chp in (0, 1, 2):
                m = tri & pauli & (((l[ia] + l[ib]) % 2) == par) & ((ch[ia] + ch[ib]) == chp)
                if m.any():
                    yield J, par, chp, np.stack([ia[m], ib[m]], axis=1).astype(np.int32)


def make_interaction(ms: ModelSpace, rng: np.random.Generator, sigma: float = 2.0,
                     decay: float = 40.0) -> Interaction:
    A, B, C, D, JJ, VV = [], [], [], [], [], []
    e = ms.orb_e
    for J, par, chp, pairs in pair_channels(ms):
        npair = pairs.shape[0]
        i, k = np.triu_indices(npair)
        epair = e[pairs[:, 0]] + e[pairs[:, 1]]
        scale = sigma * np.exp(-np.abs(epair[i] - epair[k]) / decay)
        v = rng.standard_normal(i.size) * scale
        A.append(pairs[i, 0]); B.append(pairs[i, 1]); C.append(pairs[k, 0]); D.append(pairs[k, 1])
        JJ.append(np.full(i.size, J, np.int32)); VV.append(v)
    cat = lambda x, t: np.concatenate(x).astype(t)
    return Interaction(cat(A, np.int32), cat(B, np.int32), cat(C, np.int32), cat(D, np.int32),
                       cat(JJ, np.int32), cat(VV, np.float64))


def make_system(nmax: int, seed: int, hw: float = 20.0, sigma: float = 2.0, spe_noise: float = 0.5):
    """Model space + interaction of one synthetic instance."""
    rng = np.random.default_rng(seed)
    ms = make_model_space(nmax, hw, rng, spe_noise)
    vpp = make_interaction(ms, rng, sigma)
    return ms, vpp


@dataclasses.dataclass
class InstanceSpec:
    """One instance of a heterogeneous batch: its generator inputs and its
    Lanczos length (Krylov vectors per side)."""
    nmax: int
    seed: int
    hw: float
    sigma: float
    lanc_n: int

    def make(self):
        return make_system(self.nmax, self.seed, hw=self.hw, sigma=self.sigma)


def batched_specs(n: int = 256, seed: int = 4, nmax_set=(5, 7, 9, 11), hw_range=(10.0, 30.0),
                  sigma_range=(1.0, 3.0), lanc_range=(30, 80)) -> list:
    """The heterogeneous batched workload: n independent instances with hw
    uniform in hw_range, interaction scale sigma uniform in sigma_range, Nmax
    uniform over nmax_set and Lanczos lengths uniform over lanc_range
    (inclusive).  One generator draws the four arrays in that order from
    default_rng(seed); instance i gets interaction seed 100000 * seed + i."""
    rng = np.random.default_rng(seed)
    hw = rng.uniform(*hw_range, n)
    sigma = rng.uniform(*sigma_range, n)
    nmax = rng.choice(np.asarray(nmax_set), n)
    lanc = rng.integers(lanc_range[0], lanc_range[1] + 1, n)
    return [InstanceSpec(int(nmax[i]), 100000 * seed + i, float(hw[i]), float(sigma[i]), int(lanc[i]))
            for i in range(n)]
