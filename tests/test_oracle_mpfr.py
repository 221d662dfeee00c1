"""Pins the brute-force MPFR pixel (accuratePointCompute, MP.cpp:719-894) with
closed forms of the quadratic map."""
import math

import numpy as np

import mxo

ESC2 = float(np.float32(1e8) * np.float32(1e8))  # escapeRadiusSquared(): the float product 1e8f*1e8f


def test_escape_radius_squared_is_the_float_product():
    assert ESC2 == 10000000272564224.0


def test_parabolic_implosion_count():
    """c = 1/4 + e escapes after about pi/sqrt(e) iterations.  The multiplier
    early-out is off (MANDEL_ORACLE_NODERIV): near the parabolic point the product
    of 2z_k shrinks below 1e-9 before the orbit leaves, the reference oracle's known
    false-interior case."""
    for e in (1e-4, 1e-6):
        v, n = mxo.point(repr(0.25 + e), "0", 200, 100000, esc2=ESC2, track_d=False)
        assert abs(n - math.pi / math.sqrt(e)) < 0.03 * math.pi / math.sqrt(e)


def test_superattracting_interior_early_out():
    """c = -1: period-2 superattracting cycle; the |d|^2 < 1e-9 early out fires."""
    v, n = mxo.point("-1", "0", 128, 1000, esc2=ESC2)
    assert v == -2.0 and n == -2


def test_misiurewicz_point_never_escapes():
    """c = i lands on the repelling cycle {-1+i, -i}: bounded, not attracting."""
    v, n = mxo.point("0", "1", 256, 3000, esc2=ESC2)
    assert v == -2.0 and n == -1


def test_exterior_count_and_smooth_match_double_iteration():
    """Far from the boundary full precision and plain double agree exactly on n."""
    for c in (0.5 + 0.5j, -2.1 + 0.01j, 0.26 + 0.0j):
        z, i = c, 1
        while abs(z) ** 2 <= ESC2 and i < 10000:
            z = z * z + c
            i += 1
        v, n = mxo.point(repr(c.real), repr(c.imag), 256, 100000, esc2=ESC2)
        assert n == i
        mu = i - math.log(math.log(abs(z) ** 2) / 2 / math.log(2)) / math.log(2)
        assert abs(v - mu) < 1e-9
