"""Pins the oracle's cold underflow recoveries (feRescaledStepUnderflow and
feDerivativeStepUnderflow, MP.cpp:2786-2876).

They fire on the reference's own golden frames: the minibrot at 1e-875 rebases onto
z = 0 near its nucleus, where S w^2 and dc / S fall below double, and the frames that
take the recovery thousands of times reproduce minibrot875.f32 and
minibrot875_extref.f32 bit for bit.  Their formulas are checked against exact rational
arithmetic on inputs where every term has the same size (so a dropped term, a wrong
sign or a transposed operand fails): the step dz' = 2 X dz + dz^2 + dc with dz = S w,
and the derivative J' = 2 z J + 1 with J = SJ j."""
import random
from fractions import Fraction as Fr

import numpy as np
import pytest

import cases
import mxo


@pytest.mark.parametrize("name,golden", [("minibrot875", "minibrot875.f32"), ("extref875", "minibrot875_extref.f32")])
def test_step_recovery_fires_on_golden_frames(name, golden):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    v, n, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=cm)
    assert info["underflow_steps"] > 1000
    assert np.array_equal(out.view(np.int32), cases.golden(golden, H, W).view(np.int32))


def test_derivative_recovery_fires_under_ede():
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["minibrot875_ede"]
    v, n, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=cm)
    assert info["deriv_underflows"] > 1000 and info["underflow_steps"] > 1000
    # the same frame without the derivative: identical escape counts (the recovery of
    # dz/dc does not feed back into dz)
    v0, n0, _, _ = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=0)
    assert np.array_equal(n, n0)


def _fe(x):
    """A Fraction as FloatExp (m, e), 0.5 <= |m| < 1 (m rounded to double)."""
    if x == 0:
        return (0.0, 0)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    m = x / Fr(2) ** e
    while abs(m) >= 1:
        m /= 2
        e += 1
    while abs(m) < Fr(1, 2):
        m *= 2
        e -= 1
    return (float(m), e)


def _val(fe):
    m, e = fe
    return Fr(m) * Fr(2) ** e


def _unit(rng):
    a = rng.uniform(0.0, 6.283185307179586)
    return (np.cos(a), np.sin(a))


@pytest.mark.parametrize("with_x", [True, False])
def test_step_recovery_against_exact_arithmetic(with_x):
    rng = random.Random(11)
    for _ in range(200):
        k = rng.randint(1100, 4000)            # S = 2^-k: s and d = dc/S underflow double
        S = (rng.uniform(0.5, 0.999), -k)
        w = _unit(rng)
        # X of the size of S (a rebase near z = 0) and dc of the size of S^2
        Xr = _fe(Fr(rng.uniform(-1, 1)) * Fr(2) ** -k) if with_x else (0.0, 0)
        Xi = _fe(Fr(rng.uniform(-1, 1)) * Fr(2) ** -k) if with_x else (0.0, 0)
        dcr = _fe(Fr(rng.uniform(-1, 1)) * Fr(2) ** (-2 * k))
        dci = _fe(Fr(rng.uniform(-1, 1)) * Fr(2) ** (-2 * k))
        S2, w2 = mxo.step_underflow((Xr, Xi), (dcr, dci), S, w)
        dzr, dzi = _val(S) * Fr(w[0]), _val(S) * Fr(w[1])
        xr, xi = _val(Xr), _val(Xi)
        er = 2 * (xr * dzr - xi * dzi) + (dzr * dzr - dzi * dzi) + _val(dcr)
        ei = 2 * (xr * dzi + xi * dzr) + 2 * dzr * dzi + _val(dci)
        gr, gi = _val(S2) * Fr(w2[0]), _val(S2) * Fr(w2[1])
        # the rounding bound is relative to the terms (they may cancel)
        mag = 2 * (abs(xr) + abs(xi)) * (abs(dzr) + abs(dzi)) + (abs(dzr) + abs(dzi)) ** 2 + abs(_val(dcr)) + abs(_val(dci))
        assert abs(gr - er) <= mag * Fr(1, 2 ** 48) and abs(gi - ei) <= mag * Fr(1, 2 ** 48)
        assert abs(np.hypot(*w2) - 1.0) < 1e-15


def test_derivative_recovery_against_exact_arithmetic():
    rng = random.Random(12)
    for _ in range(200):
        k = rng.randint(1100, 4000)
        # z = 2^-k (a derivative step at a vanishing z) and SJ = 2^k: 2 z J and the 1 of
        # J' = 2 z J + 1 have the same size
        zr = _fe(Fr(rng.uniform(-1, 1)) * Fr(2) ** -k)
        zi = _fe(Fr(rng.uniform(-1, 1)) * Fr(2) ** -k)
        SJ = (rng.uniform(0.5, 0.999), k)
        j = _unit(rng)
        SJ2, j2, inv = mxo.deriv_underflow((zr, zi), SJ, j)
        Jr, Ji = _val(SJ) * Fr(j[0]), _val(SJ) * Fr(j[1])
        er = 2 * (_val(zr) * Jr - _val(zi) * Ji) + 1
        ei = 2 * (_val(zr) * Ji + _val(zi) * Jr)
        gr, gi = _val(SJ2) * Fr(j2[0]), _val(SJ2) * Fr(j2[1])
        mag = 2 * (abs(_val(zr)) + abs(_val(zi))) * (abs(Jr) + abs(Ji)) + 1
        assert abs(gr - er) <= mag * Fr(1, 2 ** 48) and abs(gi - ei) <= mag * Fr(1, 2 ** 48)
        assert abs(inv * float(_val(SJ2)) - 1.0) < 1e-15
