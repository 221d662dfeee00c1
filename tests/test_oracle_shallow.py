"""Pins the oracle's shallow direct path (scale <= 1e6: floatPointCompute,
MP.cpp:330-399, and solveShallowSimdList, MP.cpp:416-532, over the view's grid,
MP.cpp:1444-1519): the lane loop equals the scalar loop (the engine's own batch check,
verify.cpp:9980-9990), exact escape counts and a 50-digit smooth value for c = 1,
closed-form interior cases, the truncated grid, and the whole-set view against the
MPFR brute force at verify.cpp's "shallow (whole set)" acceptance (verify.cpp:11789,
75-92, 292-299)."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

import mxo

ESC2 = float(np.float32(1e8) * np.float32(1e8))


def test_lane_loop_equals_scalar_loop():
    # verify.cpp:9985-9990: four points, bailout 2.1, max_iter 64
    cr = [0.2522435803501477, -0.75, 0.3, -1.2]
    ci = [0.0, 0.1, 0.5, 0.0]
    a, na = mxo.shallow_points(cr, ci, 64, escape_radius=2.1, scalar=True)
    b, nb = mxo.shallow_points(cr, ci, 64, escape_radius=2.1, scalar=False)
    assert np.array_equal(na, nb)
    assert np.array_equal(a.astype(np.float32).view(np.int32), b.astype(np.float32).view(np.int32))
    # and on points all over the set, smooth and EDE
    rng = np.random.default_rng(5)
    cr = rng.uniform(-2.2, 0.8, 20000)
    ci = rng.uniform(-1.3, 1.3, 20000)
    for ede in (False, True):
        a, na = mxo.shallow_points(cr, ci, 3000, ede=ede, scalar=True)
        b, nb = mxo.shallow_points(cr, ci, 3000, ede=ede, scalar=False)
        assert np.array_equal(na, nb)
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
        assert (na > 0).sum() > 5000 and (na == -2).sum() > 1000 and (na == -1).sum() > 0


def test_c_equals_one_closed_form():
    """c = 1: z = 1, 2, 5, 26, 677, 458330, 210066388901, ... in exact integers; the
    escape count is the first i with z_i^2 > R^2 (R^2 the float product), and the smooth
    value i + 1 - log2(log|z|^2 / (2 log 2)) at 50 digits."""
    z, i = 1, 1
    while True:
        z = z * z + 1
        if z * z > ESC2:
            break
        i += 1
    v, n = mxo.shallow_points([1.0], [0.0], 1000)
    assert n[0] == i + 1 == 7
    getcontext().prec = 50
    zz = Decimal(z) * Decimal(z)
    mu = Decimal(i + 1) - (zz.ln() / 2 / Decimal(2).ln()).ln() / Decimal(2).ln()
    assert abs(v[0] - float(mu)) < 1e-12


def test_interior_and_bounded_closed_forms():
    # c = 0: d = 2 z d vanishes at the first iteration (superattracting fixed point)
    v, n = mxo.shallow_points([0.0], [0.0], 100)
    assert n[0] == -2 and v[0] == -2.0
    # c = -1: period-2 superattracting cycle 0, -1
    v, n = mxo.shallow_points([-1.0], [0.0], 100)
    assert n[0] == -2
    # c = -2: z = 2 forever (repelling fixed point), |d| grows: max_iter
    v, n = mxo.shallow_points([-2.0], [0.0], 500)
    assert n[0] == -1 and v[0] == -2.0
    # c = i: Misiurewicz, bounded and repelling: max_iter
    v, n = mxo.shallow_points([0.0], [1.0], 500)
    assert n[0] == -1


def test_ede_value_is_the_distance_formula():
    """Under EDE the raw value is |z| log|z|^2 / |dz/dc| at the escape, which for a
    point far outside (c = 3) is checked against exact-integer iterates of z and of
    dz/dc = 2 z dz/dc + 1 at 50 digits."""
    c = 3
    z, dc = c, 1
    while True:
        dc = 2 * z * dc + 1
        z = z * z + c
        if z * z > ESC2:
            break
    v, n = mxo.shallow_points([3.0], [0.0], 100, ede=True)
    getcontext().prec = 50
    zz = Decimal(z) * Decimal(z)
    de = zz.sqrt() * zz.ln() / Decimal(dc)
    assert abs(v[0] / float(de) - 1) < 1e-14


def test_grid_is_the_truncated_engine_grid():
    cre, cim, scale, W, H = "-0.75", "0.1", "37", 61, 43
    _, _, g4 = mxo.shallow(cre, cim, scale, W, H, 50)
    (dxm, dxe), (dym, dye) = mxo.grid_info(cre, cim, scale, W, H)
    getcontext().prec = 60
    dw = Decimal(2) / Decimal(scale)
    dh = dw / W * H
    assert abs(g4[0] - float(Decimal(cre) - dw)) <= 2e-16
    assert abs(g4[1] - float(Decimal(cim) - dh)) <= 2e-16
    assert abs(g4[2] / float(dw / (W - 1) * 2) - 1) < 3e-16
    assert abs(g4[3] / float(dh / (H - 1) * 2) - 1) < 3e-16
    # the same truncation as the floatexp path's FloatExp dx, dy
    assert math.ldexp(dxm, dxe) == g4[2] and math.ldexp(dym, dye) == g4[3]


@pytest.mark.parametrize("c_method", [0, 2])
def test_whole_set_view_against_mpfr(c_method):
    """verify.cpp's shallow case: -0.5 + 0i at scale 1, max_iter 5000, 120 x 80; the
    engine passes with max |diff| <= 0.5 and <= 1 % class mismatches against
    ComputeDirect (here on every 2nd pixel)."""
    W, H, mxit = 120, 80, 5000
    val, nesc, g4 = mxo.shallow("-0.5", "0", "1", W, H, mxit, ede=c_method == 2)
    f = val.astype(np.float32).reshape(H, W)
    dv, dn = mxo.direct("-0.5", "0", "1", W, H, mxit, c_method=c_method, step=2)
    m = ~np.isnan(dv)
    d = dv.astype(np.float32)
    p_int, d_int = f[m] < -1.5, d[m] < -1.5
    mism = (p_int != d_int).mean()
    both = ~p_int & ~d_int
    diff = np.abs(f[m][both].astype(np.float64) - d[m][both].astype(np.float64))
    assert diff.max() <= 0.5
    if c_method == 2:
        rel = diff / np.abs(d[m][both].astype(np.float64))
        assert np.median(rel) < 1e-6
    assert mism <= 0.01
    assert p_int.sum() > 100 and (~p_int).sum() > 100
