"""The colour pass (tools/render.cpp:21-35, 69-91, 252-302; engine/interpolate.cpp:
7-65): the palette against an independent monotone cubic Hermite evaluation, and the
colouring and linear-light downsampling against closed forms."""
import numpy as np
import pytest

import mxhip
import mxo

POS = np.array([0.0, 0.16, 0.42, 0.6425, 0.8575, 1.0])
COL = np.array([[0, 32, 237, 255, 0, 0], [70, 107, 255, 170, 2, 70], [100, 203, 255, 0, 0, 100]], float)


def test_palette_product_equals_oracle():
    assert np.array_equal(mxhip.palette().view(np.int32), mxo.palette().view(np.int32))


def test_palette_is_the_monotone_hermite_interpolant():
    """Interior knot slopes are the weighted harmonic mean of the neighbouring secants
    (zero where they change sign) - scipy's PCHIP rule - and the end slopes the end
    secants; the cubic Hermite spline with those slopes, in float64, evaluated at k/2048."""
    from scipy.interpolate import CubicHermiteSpline, PchipInterpolator

    pal = mxo.palette()
    x = np.arange(2048) / 2048.0
    for c in range(3):
        d = PchipInterpolator(POS, COL[c]).derivative()(POS)
        d[0] = (COL[c][1] - COL[c][0]) / (POS[1] - POS[0])
        d[-1] = (COL[c][-1] - COL[c][-2]) / (POS[-1] - POS[-2])
        ref = CubicHermiteSpline(POS, COL[c], d)(x)
        assert np.abs(pal[c] - ref).max() < 2e-3
        assert pal[c][0] == COL[c][0]


def _gam(l):
    return np.floor(255.0 * np.maximum(l, 0.0) ** (1 / 2.2) + 0.5)


def test_colorize_closed_forms():
    pal = mxo.palette().astype(np.float64)
    # interior (-2) and empty (-10) samples are black
    assert (mxo.colorize(np.full((4, 6), -2.0, np.float32)) == 0).all()
    assert (mxo.colorize(np.full((4, 6), -10.0, np.float32)) == 0).all()
    # a constant smooth value: the palette entry at (int)(2048 * log(v+2)^(log(v+2)^2 / 60)) mod 2048
    for v in (0.0, 3.5, 117.25, 2.0e4, 1.5e6):
        L = np.log(np.float32(v) + 2)
        idx = int(np.float32(L ** (L * L * 60 / 3600)) * 2048) % 2048
        rgb = mxo.colorize(np.full((2, 3), v, np.float32))
        want = _gam((pal[:, idx] / 255.0) ** 2.2)
        assert np.abs(rgb.astype(float) - want).max() <= 1, v
    # distance estimate: tanh(v 60 / 3600 * 5) = 1/2 at v = atanh(1/2) * 12 -> entry 1024
    v = np.arctanh(0.5) * 12.0
    rgb = mxo.colorize(np.full((1, 1), v, np.float32), ede=True)
    assert np.abs(rgb[0, 0].astype(float) - _gam((pal[:, 1024] / 255.0) ** 2.2)).max() <= 1
    # 2x2 blocks of two black and two coloured samples: half the colour in linear light
    it = np.full((4, 4), -2.0, np.float32)
    it[0::2, 0::2] = 117.25
    it[1::2, 1::2] = 117.25
    rgb = mxo.colorize(it, ss=2)
    L = np.log(np.float32(117.25) + 2)
    idx = int(np.float32(L ** (L * L * 60 / 3600)) * 2048) % 2048
    want = _gam(((pal[:, idx] / 255.0) ** 2.2) / 2)
    assert np.abs(rgb.astype(float) - want).max() <= 1


@pytest.mark.parametrize("density", [60.0, 200.0])
def test_colorize_ss_average_equals_single_sample_blocks(density):
    """A frame whose ss x ss blocks are constant colours like its one-sample frame."""
    rng = np.random.default_rng(5)
    small = rng.uniform(0, 5e4, (6, 7)).astype(np.float32)
    small[rng.uniform(size=small.shape) < 0.2] = -2.0
    big = np.repeat(np.repeat(small, 3, axis=0), 3, axis=1)
    a = mxo.colorize(small, density=density)
    b = mxo.colorize(big, ss=3, density=density)
    assert np.abs(a.astype(int) - b.astype(int)).max() <= 1
