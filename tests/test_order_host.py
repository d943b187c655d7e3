"""bcd_order.hpp on host threads: the heapsort permutation against the
oracle's restatement of Sort_u/Sort_d (itself bit-exact against the compiled
reference's Lanczos inputs), and the pivot print rounding against the C
library's own "%18.6le" -> %lf round trip.  Built with AddressSanitizer and
UndefinedBehaviorSanitizer."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "boccadorata-public_amd", "csrc")
BUILD = os.path.join(HERE, "arrow_host", "_build")
EXE = os.path.join(BUILD, "order_host")


@pytest.fixture(scope="module")
def exe():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           os.path.join(HERE, "arrow_host", "order_host.cpp"), "-o", EXE]
    subprocess.run(cmd, check=True)
    return EXE


def run_round(exe, x):
    x = np.ascontiguousarray(x, np.float64)
    out = subprocess.run([exe], input=struct.pack("<ii", 0, x.size) + x.tobytes(), capture_output=True, check=True)
    return np.frombuffer(out.stdout, np.float64)


def run_perm(exe, key, desc):
    key = np.ascontiguousarray(key, np.float64)
    out = subprocess.run([exe], input=struct.pack("<iii", 1, key.size, int(desc)) + key.tobytes(),
                         capture_output=True, check=True)
    return np.frombuffer(out.stdout, np.int32)


def printf_round(x):
    return np.array([float("%18.6e" % v) for v in x])


def test_dys_round_random(exe):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(200000) * 10.0 ** rng.uniform(-15, 3, 200000)
    big = np.abs(x) >= 1e-16
    got, want = run_round(exe, x), printf_round(x)
    # exact where the decimal scale is an exact double (|x| >= 1e-16)
    assert np.array_equal(got[big], want[big])
    # below, within an ulp or two (no pivot component that small matters)
    assert np.all(np.abs(got[~big] - want[~big]) <= 4 * np.spacing(np.abs(want[~big])))


def test_dys_round_edges(exe):
    # exact ties of the 7th digit (dyadic values with 8 significant digits),
    # carries into the next decade, powers of ten and their neighbours
    ties = np.array([33 / 256, 0.12890625, 0.5 + 2 ** -20, 1.0 + 2 ** -21, 0.0009765625, 1.5 * 2 ** -30])
    p10 = 10.0 ** np.arange(-15, 4)
    near = np.concatenate([np.nextafter(p10, 0), p10, np.nextafter(p10, 2 * p10)])
    carry = np.array([9.9999995, 9.99999949999, 0.99999995, 0.0999999951, 999999.95, 9999999.5])
    x = np.concatenate([ties, -ties, near, -near, carry, -carry, [0.0, 1.0, -1.0]])
    assert np.array_equal(run_round(exe, x), printf_round(x))


@pytest.mark.parametrize("desc", [0, 1])
def test_heap_perm_matches_reference_sort(exe, desc):
    rng = np.random.default_rng(11 + desc)
    for n in (0, 1, 2, 3, 7, 64, 1000, 4097):
        # few distinct keys: ties everywhere, as J-degenerate poles give
        key = rng.integers(0, max(1, n // 5) + 1, n).astype(np.float64)
        perm = run_perm(exe, key, desc)
        rows = np.stack([key, np.arange(n, dtype=np.float64)], 1) if n else np.zeros((0, 2))
        ref = po.sort_rows(rows, bool(desc))
        assert np.array_equal(perm, ref[:, 1].astype(np.int32)), n
