"""Pins the oracle's periodic-nucleus reference (the nucleus search by ball
arithmetic and Newton, the island size estimate and the choice between a periodic
reference, a full reference with BLA off and a full reference; mandel_perturbation.cpp:
76-224, 1072-1324, 1624-1629) against what the reference's tests state about their
frames, exact decimal arithmetic and a textbook case."""
import math
from decimal import Decimal, getcontext

import numpy as np

import cases
import mxo


def _orbit_mods(cr, ci, p, digits):
    """Decimal exponent of |z_k|^2 (floor log10) for k = 1..p from z_0 = 0, in decimal
    arithmetic at `digits` digits."""
    getcontext().prec = digits
    cr, ci = Decimal(cr), Decimal(ci)
    zr = zi = Decimal(0)
    out = []
    for _ in range(p):
        r2, i2 = zr * zr, zi * zi
        zr, zi = r2 - i2 + cr, 2 * zr * zi + ci
        out.append((zr * zr + zi * zi).adjusted())
    return np.array(out)


def test_period3_island_nucleus_and_size():
    """Period-3 real island: its nucleus is the real root of c^3 + 2c^2 + c + 1 = 0
    (z_3 = 0 from z_0 = 0); the size estimate is the island cardioid's scale, the
    real width of its period-1 window: -1.75 (saddle node) to -1.76853 (period
    doubling), 0.0185 against the main cardioid's 1."""
    p, (nre, nim), lsize = mxo.nucleus("-1.754", "0", "200", 64, 64, 100, prec_bits=256)
    assert p == 3
    getcontext().prec = 80
    c = Decimal("-1.75")
    for _ in range(60):
        c -= (c ** 3 + 2 * c * c + c + 1) / (3 * c * c + 4 * c + 1)
    assert abs(Decimal(nre) - c) < Decimal("1e-60") and abs(Decimal(nim)) < Decimal("1e-60")
    assert 0.0125 < 10 ** lsize < 0.025


def test_deep1_nucleus_at_1e875():
    """verify.cpp:11833-11843: the deep1 view at 1e875 holds the resolved period-14164
    minibrot.  Its size estimate agrees with the golden frame's 81 interior pixels: a
    minibrot's interior is the Mandelbrot set's area, 1.5066, times size^2 (in pixels),
    to within a factor 2 on an 18 x 24 sampling.  The nucleus lies in the search disk
    (radius 2/scale about the view centre) and its orbit returns to 0 at step 14164 and
    not before, in 950-digit decimal arithmetic."""
    cre, cim, scale, W, H, mxit, _ = cases.FRAMES["minibrot875"]
    p, (nre, nim), lsize = mxo.nucleus(cre, cim, scale, W, H, mxit)
    assert p == 14164
    lpx = math.log10(4.0 / (W - 1)) - (len(scale) - 1)
    interior = (cases.golden("minibrot875.f32", H, W) < -1.5).sum()
    assert interior == 81
    assert 0.5 < interior / (1.5066 * 10 ** (2 * (lsize - lpx))) < 2.0
    getcontext().prec = 950
    d = ((Decimal(nre) - Decimal(cre)) ** 2 + (Decimal(nim) - Decimal(cim)) ** 2).sqrt()
    assert d <= 2 / Decimal(scale)
    lz = _orbit_mods(nre, nim, p, 950)
    assert lz[-1] < -800                      # |z_p|^2: 0 to far below any pixel
    assert lz[:-1].min() > lz[-1] + 600       # no earlier return


def test_reference_choice_on_the_reference_frames():
    """verify.cpp:11818-11861: at 1e875 the nucleus is the reference (81 interior
    pixels); at 1e737 it is over 130 orders below a pixel, so the centre's full reference
    runs with BLA off and every pixel is exterior; shifted 0.55 widths right the
    nucleus leaves the search disk and the centre's full reference is used."""
    cre, cim, scale, W, H, mxit, _ = cases.FRAMES["minibrot875"]
    _, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    assert (info["period"], info["subpixel"], info["ref_virtual"]) == (14164, 0, 1)
    assert (out < -1.5).sum() == 81
    cre, cim, scale, W, H, mxit, _ = cases.FRAMES["subpixel737"]
    p, _, lsize = mxo.nucleus(cre, cim, scale, W, H, mxit)
    lpix = math.log10(4.0 / (W - 1)) - (len(scale) - 1)
    assert p == 14164 and lsize - lpix < -130
    _, nesc, _, info = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    assert (info["period"], info["subpixel"], info["bla_applies"]) == (0, 1, 0)
    assert (nesc > 0).all()
    cre, cim, scale, W, H, mxit, _ = cases.FRAMES["extref875"]
    assert mxo.nucleus(cre, cim, scale, W, H, mxit)[0] == 0
    _, _, _, info = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    assert (info["period"], info["subpixel"], info["ref_virtual"]) == (0, 0, 0)


def test_periodic_and_full_reference_agree():
    """The same resolved-nucleus frame with the periodic reference and with a full
    reference from the centre pixel: identical classes (the reference's golden frame
    for both), smooth values within verify.cpp's tolerance of 1.0."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["minibrot875"]
    g = cases.golden("minibrot875.f32", H, W)
    _, _, o1, i1 = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    _, _, o0, i0 = mxo.compute_fe(cre, cim, scale, W, H, mxit, periodic=False)
    assert i1["period"] == 14164 and i0["period"] == 0
    assert np.array_equal(o1 < -1.5, g < -1.5) and np.array_equal(o0 < -1.5, g < -1.5)
    ext = g > -1.5
    assert np.abs(o1[ext] - g[ext]).max() <= 1.0 and np.abs(o0[ext] - g[ext]).max() <= 1.0
