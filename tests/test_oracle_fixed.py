"""The oracle's fixed-point reference-orbit arithmetic (the floatexp path's orbit form,
src/engine/bigfixed.h) pinned against exact integer arithmetic computed here a
different way, and the fixed-point orbit against a high-precision MPFR orbit.

A value is an integer V standing for V * 2^(-64(L-1)).  The oracle forms a short
product as the exact product minus the partial products it skips; the tests below sum
the kept partial products directly, column by column."""
from fractions import Fraction

import numpy as np
import pytest

import mxo

M64 = (1 << 64) - 1


def limbs(v, L):
    return [(v >> (64 * i)) & M64 for i in range(L)]


def kept_columns(a, b, L):
    """Sum of the partial products a_i b_j 2^(64(i+j)) that the short product keeps
    (i + j >= L - 4) for 16 <= L < 72 (bigfixed.h:75-100), every one otherwise."""
    la, lb = limbs(a, L), limbs(b, L)
    lo = L - 4 if 16 <= L < 72 else 0
    return sum(la[i] * lb[j] << (64 * (i + j)) for i in range(L) for j in range(L) if i + j >= lo)


def expected(x, y, L):
    s = 64 * (L - 1)
    p = kept_columns(abs(x), abs(y), L)
    r = (p >> s) + ((p >> (s - 1)) & 1)  # round on the top dropped bit (bigfixed.h:110-121)
    return r if (x < 0) == (y < 0) else -r


def rand_value(rng, L, int_bits=2):
    """|v| < 2^int_bits, full fraction limbs"""
    v = int(rng.integers(0, 1 << 62)) >> (62 - int_bits)
    for _ in range(L - 1):
        v = (v << 64) | int(rng.integers(0, 1 << 63)) << 1 | int(rng.integers(0, 2))
    return v if rng.integers(0, 2) else -v


@pytest.mark.parametrize("L", [3, 15, 16, 29, 49, 71, 72, 80])
def test_product_equals_kept_columns_rounded(L):
    rng = np.random.default_rng(L)
    s = 64 * (L - 1)
    for _ in range(40):
        x, y = rand_value(rng, L), rand_value(rng, L)
        r = mxo.fixed_mul(L, x, y)
        assert r == expected(x, y, L)
        # within one unit of the correctly rounded exact product, never above it in
        # magnitude (the skipped partial products are nonnegative)
        exact = Fraction(abs(x) * abs(y), 1 << s)
        rn = int(exact + Fraction(1, 2))
        assert rn - 1 <= abs(r) <= rn


def test_short_product_differs_from_exact_rounding():
    """A product one unit above the rounding midpoint (y = (2^(s-1) + 1) / x mod 2^s):
    the exact product rounds up; the short product (L = 29) drops nonzero partial
    products below the midpoint and rounds down; the full product (L = 15) rounds up."""
    for L, drop in ((29, 1), (15, 0)):
        s = 64 * (L - 1)
        x = (1 << s) | 0x9E3779B97F4A7C15F39CC0605CEDC835
        y = ((1 << (s - 1)) + 1) * pow(x, -1, 1 << s) % (1 << s) | (1 << s)
        p = x * y
        assert p % (1 << s) == (1 << (s - 1)) + 1
        exact_rn = (p >> s) + 1
        r = mxo.fixed_mul(L, x, y)
        assert r == expected(x, y, L) == exact_rn - drop


def test_product_signs_and_zero():
    L = 20
    one = 1 << (64 * (L - 1))
    half = one >> 1
    assert mxo.fixed_mul(L, one, half) == half
    assert mxo.fixed_mul(L, -one, half) == -half
    assert mxo.fixed_mul(L, -one, -half) == half
    assert mxo.fixed_mul(L, 0, half) == 0
    # 2^-(64(L-1)) squared rounds to zero
    assert mxo.fixed_mul(L, 1, -1) == 0


@pytest.mark.parametrize("L", [3, 16, 40])
def test_shadow_relative_error(L):
    """The shadow from the top two nonzero limbs is within 2^-51 of the exact value,
    at every magnitude the fixed point holds."""
    rng = np.random.default_rng(100 + L)
    s = 64 * (L - 1)
    for k in range(60):
        v = rand_value(rng, L) >> int(rng.integers(0, 64 * L - 8))
        if v == 0:
            continue
        m, e = mxo.fixed_to_fe(L, v)
        assert 0.5 <= abs(m) < 1.0 and (m < 0) == (v < 0)
        got = Fraction(m) * Fraction(2) ** e
        exact = Fraction(v, 1 << s)
        assert abs(got - exact) <= abs(exact) * Fraction(1, 1 << 51)


def test_shadow_closed_forms():
    L = 5
    s = 64 * (L - 1)
    assert mxo.fixed_to_fe(L, 3 << (s - 2)) == (0.75, 0)          # 0.75
    assert mxo.fixed_to_fe(L, -(1 << s)) == (-0.5, 1)             # -1
    assert mxo.fixed_to_fe(L, 1) == (0.5, 1 - s)                  # 2^-(64(L-1))
    assert mxo.fixed_to_fe(L, 0) == (0.0, 0)
    # top limb 2^64-1 (limb 2) converts to 2^64: mantissa 2^128 = 0.5 * 2^129, scaled
    # by 2^(64 * (2 - L)), i.e. 2^-64 for (2^64 - 1) * 2^-128
    assert mxo.fixed_to_fe(L, M64 << 128) == (0.5, -63)


def _mpfr_reference(cre, cim, c_bits, bits, n):
    """orbit at `bits` from c rounded to c_bits, as the oracle holds it"""
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = c_bits
    c = mp.mpc(mp.mpf(cre), mp.mpf(cim))
    mp.mp.prec = bits
    z = c
    out = [z]
    for _ in range(n):
        z = z * z + c
        out.append(z)
    return out


@pytest.mark.parametrize("prec_bits,limbs", [(700, 13), (1100, 19)])
def test_fixed_orbit_tracks_high_precision_orbit(prec_bits, limbs):
    """The fixed-point orbit of a bounded boundary point (c held at prec_bits) stays
    within 2^-50 relative (the fixed point's absolute resolution times the
    orbit's growth) of the orbit computed at 4x the precision from the same c.
    Shadows are checked at a few iterations along a 1500-step orbit."""
    cre, cim = "-0.743643887037158704752191506114774", "0.131825904205311970493132056385139"
    n = 1500
    o = mxo.orbit(cre, cim, prec_bits, n, fx_limbs=limbs)
    assert o["ref_it"] == n
    ref = _mpfr_reference(cre, cim, prec_bits, 4 * prec_bits, n)
    worst = 0.0
    for k in (0, 1, 2, 10, 100, 500, 1000, 1500):
        zr = float(ref[k].real)
        zi = float(ref[k].imag)
        for got, want in ((o["zf"][2 * k], zr), (o["zf"][2 * k + 1], zi)):
            worst = max(worst, abs(got - want) / max(abs(want), 1e-300))
    assert worst < 2.0 ** -50


def test_fixed_orbit_of_dyadic_point_is_exact():
    """Dyadic orbits stay exact in fixed point: c = i gives the 2-cycle -1+i, -i, ...
    and c = -2 the fixed point 2."""
    o = mxo.orbit("0", "1", 300, 40, fx_limbs=7)
    z = o["zf"].reshape(-1, 2)
    assert tuple(z[0]) == (0.0, 1.0) and tuple(z[1]) == (-1.0, 1.0) and tuple(z[2]) == (0.0, -1.0)
    assert (z[3::2] == (-1.0, 1.0)).all() and (z[4::2] == (0.0, -1.0)).all()
    o = mxo.orbit("-2", "0", 300, 40, fx_limbs=7)
    z = o["zf"].reshape(-1, 2)
    assert tuple(z[0]) == (-2.0, 0.0) and (z[1:] == (2.0, 0.0)).all()
