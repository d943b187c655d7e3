"""The arrowhead Dyson solver (csrc/bcd_arrow.hpp, the body of the k_arrow
kernel) run on host threads, built with AddressSanitizer and
UndefinedBehaviorSanitizer, against LAPACK's dense eigensolver (numpy.linalg.eigh)
on the assembled Dyson matrix: energies, spectroscopic factors and
amplitudes, the split at E_F and the Calc_En sums.  Cases: random couplings,
exactly degenerate and nearly degenerate pole energies, zero and tiny
couplings, no poles, and real Lanczos-reduced self-energies from the golden
fixtures (fishbone blocks diagonalised first, as the GPU path does)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from helpers import load_golden

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "boccadorata-public_amd", "csrc")
BUILD = os.path.join(HERE, "arrow_host", "_build")
EXE = os.path.join(BUILD, "arrow_host")


@pytest.fixture(scope="module")
def exe():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           os.path.join(HERE, "arrow_host", "arrow_host.cpp"), "-o", EXE, "-pthread"]
    subprocess.run(cmd, check=True)
    return EXE


def run(exe, cases, with_flops=False):
    """Runs every case; a case with "h2" runs a second time with static part
    h2, warm-started from the first run's roots, and yields a pair of results.
    Each result is (nqp, nqh, failed, e, x, kolt) plus, with with_flops, the
    FLOPs the body counted (its iteration counts)."""
    buf = [struct.pack("<i", len(cases))]
    for c in cases:
        nI = c["h"].shape[0]
        warm = int("h2" in c)
        buf.append(struct.pack("<6id", nI, len(c["E0"]), len(c["E1"]), c["nt"], c["j2"], warm, c["ef"]))
        for k in ("h", "u1", "E0", "C0", "E1", "C1") + (("h2",) if warm else ()):
            buf.append(np.ascontiguousarray(c[k], np.float64).tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], input=b"".join(buf), capture_output=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    out, pos = [], 0
    raw = r.stdout

    def record(n, nI):
        nonlocal pos
        nqp, nqh, failed = struct.unpack_from("<3i", raw, pos); pos += 12
        e = np.frombuffer(raw, np.float64, n, pos); pos += 8 * n
        x = np.frombuffer(raw, np.float64, n * nI, pos).reshape(n, nI); pos += 8 * n * nI
        k = np.frombuffer(raw, np.float64, 3, pos); pos += 24
        fl = struct.unpack_from("<d", raw, pos)[0]; pos += 8
        return (nqp, nqh, failed, e, x, k) + ((fl,) if with_flops else ())
    for c in cases:
        nI = c["h"].shape[0]
        n = nI + len(c["E0"]) + len(c["E1"])
        first = record(n, nI)
        out.append((first, record(n, nI)) if "h2" in c else first)
    assert pos == len(raw)
    return out


def dense(c):
    h = c["h"] + c["u1"]
    nI = h.shape[0]
    E = np.concatenate([c["E0"], c["E1"]])
    C = np.concatenate([c["C0"].reshape(-1, nI), c["C1"].reshape(-1, nI)])
    m = len(E)
    A = np.zeros((nI + m, nI + m))
    A[:nI, :nI] = h
    A[nI:, :nI] = C
    A[:nI, nI:] = C.T
    A[nI:, nI:] = np.diag(E)
    return np.linalg.eigh(A)


def check(c, res, tol_e=1e-12, tol_z=1e-12, tol_x=1e-10):
    nqp, nqh, failed, e, x, kolt = res
    assert failed == 0
    w, V = dense(c)
    nI = c["h"].shape[0]
    n = len(w)
    ef = c["ef"]
    assert nqh == int(np.sum(w <= ef)) and nqp == n - nqh
    # records: qp descending then qh descending
    assert np.all(np.diff(e[:nqp]) <= 0) and np.all(np.diff(e[nqp:]) <= 0)
    assert np.all(e[:nqp] > ef) and np.all(e[nqp:] <= ef)
    order = np.concatenate([np.arange(n - 1, nqh - 1, -1), np.arange(nqh - 1, -1, -1)])
    scale = max(1.0, np.abs(w).max())
    assert np.max(np.abs(e - w[order])) < tol_e * scale
    X = V[:nI, order].T
    z_ref, z_got = np.sum(X * X, 1), np.sum(x * x, 1)
    assert np.max(np.abs(z_got - z_ref)) < tol_z
    # amplitudes up to sign, where the eigenvalue is isolated
    gaps = np.minimum(np.abs(np.diff(w, prepend=-np.inf)), np.abs(np.diff(w, append=np.inf)))[order]
    iso = (gaps > 1e-6 * scale) & (z_ref > 1e-8)
    sg = np.sign(np.sum(X * x, 1)); sg[sg == 0] = 1
    assert np.max(np.abs(x * sg[:, None] - X)[iso], initial=0.0) < tol_x
    # Calc_En sums over quasiholes
    f = c["j2"] + 1.0
    qh = slice(nqp, n)
    x2 = np.einsum("kr,rc,kc->k", x[qh], c["u1"], x[qh])
    assert abs(kolt[0] - f * x2.sum()) < 1e-10 * max(1, abs(kolt[0]))
    assert abs(kolt[1] - f * np.sum(z_got[qh] * e[qh])) < 1e-10 * max(1, abs(kolt[1]))
    assert abs(kolt[2] - f * np.sum(z_got[qh])) < 1e-10 * max(1, abs(kolt[2]))


def case(rng, nI, m0, m1, nt=64, scale_c=1.0, ties=False, ef=0.123):
    h = rng.standard_normal((nI, nI)) * 10; h = h + h.T
    u1 = np.diag(rng.standard_normal(nI))
    E = rng.standard_normal(m0 + m1) * 30
    if ties:
        E = np.round(E / 10.0) * 10.0
    C = rng.standard_normal((m0 + m1, nI)) * scale_c
    return dict(h=h, u1=u1, E0=E[:m0], C0=C[:m0], E1=E[m0:], C1=C[m0:], nt=nt, j2=1, ef=ef)


def test_random_and_degenerate(exe):
    rng = np.random.default_rng(7)
    cases = [case(rng, 1, 0, 0), case(rng, 3, 0, 0), case(rng, 1, 7, 3), case(rng, 4, 50, 50, nt=128),
             case(rng, 4, 200, 97, nt=256), case(rng, 6, 80, 80, nt=64), case(rng, 2, 60, 40, ties=True),
             case(rng, 4, 60, 40, ties=True, nt=128)]
    c = case(rng, 4, 60, 40)
    c["C0"][::3] = 0.0            # exactly decoupled poles
    c["C1"][::4, 1] = 1e-19       # couplings below the deflation tolerance
    cases.append(c)
    c = case(rng, 4, 60, 40)
    c["E0"][1::2] = c["E0"][0::2][:len(c["E0"][1::2])] + 1e-13   # nearly degenerate pairs
    cases.append(c)
    cases.append(case(rng, 4, 60, 40, scale_c=1e-7))              # weak coupling
    cases.append(case(rng, 16, 20, 20, nt=64))                    # largest nI
    for c, r in zip(cases, run(exe, cases)):
        check(c, r)


def fishbone(rows, ipv, den, nI):
    nd = rows.shape[0]
    D = np.diag(rows[:, 0]).astype(float)
    for i in range(1, nd):
        D[i, i - 1] = D[i - 1, i] = rows[i, 1]
    for nc in range(1, int(ipv[0]) + 1):
        st = int(ipv[nc + 1])
        if st < 1 or st >= nd:
            break
        D[st:, st - 1] = rows[st:, 1 + nc]
        D[st - 1, st:] = rows[st:, 1 + nc]
    return D, rows[:, den:den + nI]


def test_reduced_self_energies(exe):
    """Golden Lanczos-reduced self-energies: each fishbone side diagonalised,
    then the arrowhead sweeps, against the dense Dyson matrix."""
    cases = []
    for name in ("c1_hf_split50", "c2_hf_split50"):
        ms, vpp, g, z = load_golden(name)
        for i in z["channels"]:
            if f"lfw_{i}" not in z.files:
                continue
            i = int(i)
            n = int(ms.no[i])
            fw, bk = z[f"lfw_{i}"], z[f"lbk_{i}"]
            den = int(z[f"ipfw_{i}"][0]) + 1
            side = []
            for rows, ipv in ((fw, z[f"ipfw_{i}"]), (bk, z[f"ipbk_{i}"])):
                D, M = fishbone(rows, ipv, den, n)
                lam, U = np.linalg.eigh(D)
                side.append((lam, U.T @ M))
            h = np.diag(np.linspace(-30.0, 60.0, n)) + 0.5
            cases.append(dict(h=h, u1=np.zeros((n, n)), E0=side[0][0], C0=side[0][1], E1=side[1][0],
                              C1=side[1][1], nt=128, j2=int(ms.j2[i]), ef=1.0))
    for c, r in zip(cases, run(exe, cases)):
        check(c, r, tol_e=1e-12, tol_z=1e-11, tol_x=1e-9)


def test_warm_start_from_previous_roots(exe):
    """A second solve with a slightly changed static part, started from the
    first solve's roots (what the next sc0 sweep does), gives the dense
    eigenpairs of the new matrix and needs fewer secular evaluations; a large
    change, which moves roots out of their old intervals, still solves."""
    rng = np.random.default_rng(11)
    for nI, m0, m1, nt, rel, fewer in ((4, 60, 40, 128, 1e-7, True), (4, 200, 97, 256, 1e-9, True),
                                       (2, 60, 40, 64, 1e-3, False), (4, 60, 40, 128, 0.3, False)):
        c = case(rng, nI, m0, m1, nt=nt)
        d = rng.standard_normal((nI, nI)) * rel * np.abs(c["h"]).max()
        c["h2"] = c["h"] + d + d.T
        (r1, r2), = run(exe, [c], with_flops=True)
        check(c, r1[:6])
        c2 = dict(c, h=c["h2"])
        check(c2, r2[:6])
        if fewer:
            assert r2[6] < 0.75 * r1[6], (r1[6], r2[6])


# ---------------------------------------------------------------- static_dyson

STATIC_EXE = os.path.join(BUILD, "static_host")


@pytest.fixture(scope="module")
def static_exe():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           os.path.join(HERE, "arrow_host", "static_host.cpp"), "-o", STATIC_EXE]
    subprocess.run(cmd, check=True)
    return STATIC_EXE


def run_static(exe, cases):
    buf = [struct.pack("<i", len(cases))]
    for c in cases:
        nI = c["h"].shape[0]
        buf.append(struct.pack("<2id", nI, c["j2"], c["ef"]))
        buf.append(np.ascontiguousarray(c["h"], np.float64).tobytes())
        buf.append(np.ascontiguousarray(c["u1"], np.float64).tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], input=b"".join(buf), capture_output=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr.decode()[-3000:]
    raw, pos, out = r.stdout, 0, []
    for c in cases:
        nI = c["h"].shape[0]
        rc, nqp, nqh = struct.unpack_from("<3i", raw, pos); pos += 12
        e = np.frombuffer(raw, np.float64, nI, pos); pos += 8 * nI
        x = np.frombuffer(raw, np.float64, nI * nI, pos).reshape(nI, nI); pos += 8 * nI * nI
        k = np.frombuffer(raw, np.float64, 3, pos); pos += 24
        out.append((rc, nqp, nqh, e, x, k))
    assert pos == len(raw)
    return out


def static_reference(c):
    """The records and Calc_En sums from numpy's dense eigensolver on the
    static part (lower triangle of u1 + h), laid out as DysonMatrix stores them:
    quasiparticles (e > E_F) descending, then quasiholes descending."""
    a = np.tril(c["u1"] + c["h"])
    a = a + np.tril(a, -1).T
    w, v = np.linalg.eigh(a)
    qp, qh = np.flatnonzero(w > c["ef"])[::-1], np.flatnonzero(w <= c["ef"])[::-1]
    order = np.concatenate([qp, qh])
    kb = np.zeros(9)
    for i in qh:
        x1 = float(v[:, i] @ v[:, i])
        x2 = float(v[:, i] @ c["u1"] @ v[:, i])
        bk = 0 if x1 > 0.01 else (1 if x1 > 1e-5 else 2)
        kb[3 * bk:3 * bk + 3] += (x2, x1 * w[i], x1)
    f = c["j2"] + 1.0
    kolt = np.array([kb[0::3].sum(), kb[1::3].sum(), kb[2::3].sum()]) * f
    return len(qp), len(qh), w[order], v[:, order].T, kolt


def static_cases():
    rng = np.random.default_rng(5)
    cases = []
    for nI in (1, 2, 3, 4):
        for ef in (-1e3, 0.0, 1e3):
            a = rng.normal(0.0, 10.0, (nI, nI))
            u = rng.normal(0.0, 10.0, (nI, nI))
            u = u + u.T
            cases.append(dict(h=a, u1=u, ef=ef, j2=2 * nI - 1))
    # only the lower triangle counts: garbage above the diagonal
    h = np.tril(rng.normal(0.0, 5.0, (4, 4))) + np.triu(np.full((4, 4), 1e6), 1)
    cases.append(dict(h=h, u1=np.zeros((4, 4)), ef=0.0, j2=3))
    # wide spread of scales, tiny couplings, an exactly diagonal matrix
    cases.append(dict(h=np.diag([1e5, -3.0, 2e-3, 40.0]) + 1e-9 * np.tril(np.ones((4, 4)), -1),
                      u1=np.zeros((4, 4)), ef=1.0, j2=7))
    cases.append(dict(h=np.diag([3.0, -1.0, 2.0]), u1=np.diag([1.0, 1.0, 1.0]), ef=2.5, j2=1))
    return cases


def test_static_dyson_against_eigh(static_exe):
    """Distinct eigenvalues: energies to 1e-13 of the spectral radius,
    amplitudes to 1e-11 with the largest component positive, the split at
    E_F and the Calc_En sums."""
    cases = static_cases()
    for c, (rc, nqp, nqh, e, x, k) in zip(cases, run_static(static_exe, cases)):
        assert rc == 0
        rnqp, rnqh, re, rx, rk = static_reference(c)
        assert (nqp, nqh) == (rnqp, rnqh)
        scale = max(1.0, np.abs(re).max())
        assert np.abs(e - re).max() < 1e-13 * scale
        for i in range(len(e)):
            big = np.argmax(np.abs(x[i]))
            assert x[i, big] > 0 and np.argmax(np.abs(x[i]) > np.abs(x[i, big]) * (1 - 1e-12)) == big
            s = np.sign(rx[i] @ x[i])
            assert np.abs(x[i] - s * rx[i]).max() < 1e-11
        assert np.abs(k - rk).max() < 1e-11 * max(1.0, np.abs(rk).max())


def test_static_dyson_degenerate(static_exe):
    """Repeated eigenvalues (rotated diagonal matrices, the zero matrix): the
    records still hold orthonormal eigenvectors of the matrix, every energy
    counted once."""
    rng = np.random.default_rng(6)
    cases = []
    for lam in ([2.0, 2.0, -1.0, 5.0], [1.0, 1.0, 1.0, 1.0], [0.0, 0.0, 0.0, 0.0], [-3.0, 4.0, 4.0]):
        n = len(lam)
        q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        a = q @ np.diag(lam) @ q.T
        cases.append(dict(h=np.tril(a), u1=np.zeros((n, n)), ef=1.5, j2=3))
    for c, (rc, nqp, nqh, e, x, k) in zip(cases, run_static(static_exe, cases)):
        assert rc == 0
        a = np.tril(c["h"]) + np.tril(c["h"], -1).T
        lam = np.sort(np.linalg.eigvalsh(a))
        assert np.abs(np.sort(e) - lam).max() < 1e-13 * max(1.0, np.abs(lam).max())
        assert nqh == int((e <= c["ef"]).sum()) and nqp == len(e) - nqh
        assert np.all(e[:nqp] > c["ef"]) and np.all(np.diff(e[:nqp]) <= 0) and np.all(np.diff(e[nqp:]) <= 0)
        assert np.abs(x @ x.T - np.eye(len(e))).max() < 1e-13
        assert np.abs(a @ x.T - x.T * e).max() < 1e-12 * max(1.0, np.abs(lam).max())


def test_static_dyson_rejects_large_ni(static_exe):
    """nI above the compiled bound is reported, not solved."""
    c = dict(h=np.eye(5), u1=np.zeros((5, 5)), ef=0.0, j2=1)
    assert run_static(static_exe, [c])[0][0] == 1
