"""Pins the oracle's FloatExp restatement (oracle/mxo_fe.h, floatexp.h:17-107):
every operation against exact rational arithmetic, and bit for bit against the
reference header itself compiled from its own source (oracle/_ref), when built."""
import ctypes as C
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import mxo

N = 4000
rng = np.random.default_rng(1234)


def rand_fe(n, emin=-3000, emax=3000, zeros=True):
    m = rng.uniform(0.5, 1.0, n) * rng.choice([-1.0, 1.0], n)
    e = rng.integers(emin, emax, n)
    if zeros:
        z = rng.random(n) < 0.03
        m[z] = 0.0
        e[z] = 0
    return m, e.astype(np.int64)


def val(m, e):
    return Fraction(float(m)) * (Fraction(2) ** int(e))


def normalised(m, e):
    return m == 0.0 and e == 0 or 0.5 <= abs(m) < 1.0


def round_fe(x):
    """exact rational -> (m, e) with m the correctly rounded double mantissa in [0.5,1)"""
    if x == 0:
        return 0.0, 0
    s = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while x / Fraction(2) ** e >= 1:
        e += 1
    while x / Fraction(2) ** e < Fraction(1, 2):
        e -= 1
    m = float(x / Fraction(2) ** e)  # float(Fraction) rounds correctly
    if m == 1.0:
        m, e = 0.5, e + 1
    return s * m, e


def test_mul_is_correctly_rounded():
    am, ae = rand_fe(N)
    bm, be = rand_fe(N)
    rm, re = mxo.fe_op(0, am, ae, bm, be)
    for k in range(N):
        assert normalised(rm[k], re[k])
        assert (rm[k], re[k]) == round_fe(val(am[k], ae[k]) * val(bm[k], be[k]))


def test_mul_d():
    am, ae = rand_fe(N)
    b = rng.normal(size=N) * 10.0 ** rng.integers(-30, 30, N)
    rm, re = mxo.fe_op(1, am, ae, b, np.zeros(N, np.int64))
    for k in range(N):
        assert (rm[k], re[k]) == round_fe(val(am[k], ae[k]) * Fraction(float(b[k])))


def test_add_sub_follow_the_54_bit_gap_rule():
    am, ae = rand_fe(N, -80, 80)
    bm, be = rand_fe(N, -80, 80)
    for op, sign in ((2, 1), (3, -1)):
        rm, re = mxo.fe_op(op, am, ae, bm, be)
        for k in range(N):
            a, b = (am[k], ae[k]), (sign * bm[k], be[k])
            if a[0] == 0.0:
                exp = b
            elif b[0] == 0.0:
                exp = a
            elif a[1] - b[1] > 54:
                exp = a
            elif a[1] - b[1] < -54:
                exp = b
            else:
                # one rounding of the aligned mantissa sum (the shifted term is exact)
                hi, lo = (a, b) if a[1] >= b[1] else (b, a)
                s = Fraction(float(hi[0])) + Fraction(float(lo[0])) * Fraction(2) ** int(lo[1] - hi[1])
                sm, se = round_fe(Fraction(float(s)))
                exp = (sm, se + hi[1]) if sm != 0.0 else (0.0, 0)
            assert (rm[k], re[k]) == exp, (k, a, b)


def test_div_and_sqrt():
    am, ae = rand_fe(N, zeros=False)
    bm, be = rand_fe(N, zeros=False)
    rm, re = mxo.fe_op(4, am, ae, bm, be)
    for k in range(N):
        assert (rm[k], re[k]) == round_fe(val(am[k], ae[k]) / val(bm[k], be[k]))
    pm = np.abs(am)
    rm, re = mxo.fe_op(5, pm, ae)
    for k in range(N):
        x = val(pm[k], ae[k])
        r = val(rm[k], re[k])
        ulp = Fraction(2) ** (int(re[k]) - 53)
        assert (r - ulp / 2) ** 2 <= x <= (r + ulp / 2) ** 2
        assert normalised(rm[k], re[k])


def test_to_double_clamps_and_rounds():
    am, ae = rand_fe(N, -1150, 1150)
    rm, _ = mxo.fe_op(7, am, ae)
    for k in range(N):
        if am[k] == 0.0 or ae[k] < -1100:
            assert rm[k] == 0.0
        elif ae[k] > 1100:
            assert rm[k] == math.copysign(math.inf, am[k])
        else:
            try:
                exp = math.ldexp(float(am[k]), int(ae[k]))
            except OverflowError:
                exp = math.copysign(math.inf, am[k])
            assert rm[k] == exp


def test_abs_less():
    am, ae = rand_fe(N, -3, 3)
    bm, be = rand_fe(N, -3, 3)
    rm, _ = mxo.fe_op(8, am, ae, bm, be)
    for k in range(N):
        assert bool(rm[k]) == (abs(val(am[k], ae[k])) < abs(val(bm[k], be[k])))


REF = os.path.join(os.path.dirname(mxo.__file__), "_ref", "libfe_ref.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (needs the reference source tree)")
@pytest.mark.parametrize("op", range(9))
def test_bit_identical_to_reference_header(op):
    L = C.CDLL(REF)
    am, ae = rand_fe(N, -1200, 1200)
    bm, be = rand_fe(N, -1200, 1200)
    if op == 1:
        bm = rng.normal(size=N)
    if op == 5:
        am = np.abs(am)
    rm, re = mxo.fe_op(op, am, ae, bm, be)
    xm, xe = np.zeros(N), np.zeros(N, np.int64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    L.fe_ref_op(op, N, p(am), p(ae), p(bm), p(be), p(xm), p(xe))
    assert np.array_equal(rm.view(np.int64), xm.view(np.int64))
    assert np.array_equal(re, xe)
