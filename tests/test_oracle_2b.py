"""Pins of the oracle's two-body propagator matrices (oracle/bcd_oracle_2b.c):
the ladder of gII_prop_t (gII_b.cpp:62-563) and the ring of Pol_prop_t
(PolProp_b.cpp:43-614).

* Golden vectors of the compiled reference (tests/golden/c1_*_2b.npz, written
  by oracle/_ref/ref_harness_2b): bases of every two-body channel, with and
  without the n1 cut, equal exactly; the matrices of every mode, including
  the part of the buffer the reference clears and the part it leaves alone,
  and the free poles, are bit-identical.
* orc_s6j against exact rational arithmetic (the Racah sum in integers) on
  every admissible argument set with 2j <= 9 and known closed forms.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as po
from helpers import load_golden

CASES = ["c1_hf_2b", "c1_dressed_2b"]
FILL = -777.25


def exact_6j(a, b, c, d, e, f):
    """{a/2 b/2 c/2; d/2 e/2 f/2} from the Racah formula in exact arithmetic."""
    def tri(x, y, z):
        s = (x + y - z, y + z - x, z + x - y)
        return all(v >= 0 and v % 2 == 0 for v in s)
    if not (tri(a, b, c) and tri(a, e, f) and tri(d, b, f) and tri(d, e, c)):
        return 0.0
    F = math.factorial

    def delta2(x, y, z):   # Delta(x,y,z)^2 as a Fraction, arguments doubled
        return Fraction(F((x + y - z) // 2) * F((x - y + z) // 2) * F((-x + y + z) // 2), F((x + y + z) // 2 + 1))
    pre = delta2(a, b, c) * delta2(a, e, f) * delta2(d, b, f) * delta2(d, e, c)
    al = [(a + b + c) // 2, (a + e + f) // 2, (d + b + f) // 2, (d + e + c) // 2]
    be = [(a + b + d + e) // 2, (b + c + e + f) // 2, (c + a + f + d) // 2]
    tot = Fraction(0)
    for t in range(max(al), min(be) + 1):
        den = 1
        for x in al:
            den *= F(t - x)
        for x in be:
            den *= F(x - t)
        tot += Fraction((-1) ** t * F(t + 1), den)
    val = tot * tot * pre
    return math.copysign(math.sqrt(val.numerator / val.denominator) if val else 0.0, float(tot))


def test_s6j_exact():
    js = range(0, 10)
    worst, n = 0.0, 0
    for a, b, c, d, e, f in itertools.product(js, repeat=6):
        if (a + b + c) % 2 or (a + e + f) % 2 or (d + b + f) % 2 or (d + e + c) % 2:
            continue
        want = exact_6j(a, b, c, d, e, f)
        got = po.s6j(a, b, c, d, e, f)
        worst = max(worst, abs(got - want))
        n += 1
    assert n > 10000
    assert worst < 1e-14


def test_s6j_closed_forms():
    # {a b c; 0 c b} = (-1)^(a+b+c) / sqrt((2b+1)(2c+1))
    for a2, b2, c2 in itertools.product(range(0, 12), repeat=3):
        if (a2 + b2 + c2) % 2 or abs(b2 - c2) > a2 or a2 > b2 + c2:
            continue
        want = (-1) ** ((a2 + b2 + c2) // 2) / math.sqrt((b2 + 1) * (c2 + 1))
        assert abs(po.s6j(a2, b2, c2, 0, c2, b2) - want) < 1e-15
    assert abs(po.s6j(1, 1, 2, 1, 1, 2) - 1.0 / 6.0) < 1e-16
    assert po.s6j(1, 1, 4, 1, 1, 2) == 0.0   # triangle broken


def _case(name):
    ms, vpp, g, z = load_golden(name)
    S = po.System(ms, vpp)
    P = po.Prop.from_propagator(S, g)
    return ms, S, P, z


def _channels(z, pre):
    seen = set()
    for k in z.files:
        if k.startswith(pre + "_"):
            J, pi, dT = (int(x) for x in k.split("_")[1:4])
            seen.add((J, pi, dT))
    return sorted(seen)


def _basis(z, pre, J, pi, dT, keys):
    return {k: z[f"{pre}_{J}_{pi}_{dT}_{k}"] for k in keys}


@pytest.mark.parametrize("name", CASES)
def test_bases_equal_reference(name):
    ms, S, P, z = _case(name)
    n1p = np.arange(len(z["prop_ep"])) % 3
    n1h = np.arange(len(z["prop_eh"])) % 3
    for pre, fn, keys in (("lb", po.ladd_basis, po.LADD_KEYS), ("rb", po.ph_basis, po.PH_KEYS)):
        chans = _channels(z, pre)
        assert chans
        for J, pi, dT in chans:
            want = _basis(z, pre, J, pi, dT, keys)
            got = fn(S, P, J, dT, pi)
            for k in keys:
                assert np.array_equal(got[k], want[k]), (pre, J, pi, dT, k)
            for icut in (-1, 1):
                cpre = pre[0] + f"c{icut}"
                want = _basis(z, cpre, J, pi, dT, keys)
                got = fn(S, P, J, dT, pi, icut=icut, ntcut=2, n1p=n1p, n1h=n1h)
                for k in keys:
                    assert np.array_equal(got[k], want[k]), (cpre, J, pi, dT, k)


def matrix_keys(z, tag):
    for k in z.files:
        if k.startswith(tag + "m_"):
            J, pi, dT, mode, add = (int(x) for x in k.split("_")[1:6])
            yield k, J, pi, dT, mode, add


@pytest.mark.parametrize("name", CASES)
def test_matrices_equal_reference(name):
    ms, S, P, z = _case(name)
    worst, nmat, exact = 0.0, 0, 0
    for tag, keys in (("l", po.LADD_KEYS), ("r", po.PH_KEYS)):
        for k, J, pi, dT, mode, add in matrix_keys(z, tag):
            want = z[k]
            wfp = z[k.replace("m_", "f_", 1)]
            assert int(z[k.replace("m_", "rc_", 1)][0]) == 0
            b = _basis(z, tag + "b", J, pi, dT, keys)
            ld = want.shape[1]
            if tag == "l":
                got, gfp = po.ladd_fill(S, P, J, b, mode, bool(add), ld=ld, fill=FILL)
            else:
                got, gfp = po.ph_fill(S, P, J, dT, b, mode, bool(add), ld=ld, fill=FILL)
            assert got.shape == want.shape, k
            assert np.array_equal(gfp, wfp), k
            if want.size:
                scale = max(1.0, float(np.abs(want[want != FILL]).max(initial=0.0)))
                worst = max(worst, float(np.abs(got - want).max()) / scale)
            exact += int(np.array_equal(got, want))
            nmat += 1
    assert nmat > 50
    assert worst < 1e-13, worst
    # the restatement follows the reference's operation order, so every
    # matrix comes out bit for bit
    assert exact == nmat


def test_mode_errors():
    ms, S, P, z = _case("c1_hf_2b")
    b = _basis(z, "lb", 0, 0, 0, po.LADD_KEYS)
    with pytest.raises(RuntimeError, match="-100"):
        po.ladd_fill(S, P, 0, b, 5)
    with pytest.raises(RuntimeError, match="-102"):
        po.ladd_fill(S, P, 0, b, 3, ld=2)
    rb = _basis(z, "rb", 0, 0, 0, po.PH_KEYS)
    with pytest.raises(RuntimeError, match="-100"):
        po.ph_fill(S, P, 0, 0, rb, 2)
