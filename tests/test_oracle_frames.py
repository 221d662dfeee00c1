"""Pins the oracle's whole-frame restatement (reference orbit, BLA, probe gate,
pixelRescaled) against the reference's own saved frames and known answers."""
import numpy as np
import pytest

import cases
import mxo


@pytest.mark.parametrize("name,golden", [("minibrot875", "minibrot875.f32"), ("extref875", "minibrot875_extref.f32")])
def test_reference_golden_frames_bit_exact(name, golden):
    """verify.cpp:11838-11857 accept class 0% / max diff 1.0; the restatement
    reproduces the saved float frames exactly."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    val, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=cm)
    g = cases.golden(golden, H, W)
    assert np.array_equal(out < -1.5, g < -1.5)
    assert np.array_equal(out.view(np.int32), g.view(np.int32))


def test_exterior1000_known_answer():
    """c = i is a Misiurewicz point; the 1e1000 frame around it is entirely exterior
    and escapes after about 2650 iterations (verify.cpp:11802-11805)."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["exterior1000"]
    val, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    assert (nesc > 0).all()
    assert 2600 < nesc.min() and nesc.max() < 2700
    # smooth value n - log2(log2|z|) with ESC2 < |z|^2 <= ~ESC2^2: log2 log2|z| in (4.73, 5.74]
    assert ((val < nesc - 4.73) & (val >= nesc - 5.74)).all()


def test_subpixel737_all_exterior():
    """verify.cpp:11818-11831: the deep1 nucleus is 139 orders sub-pixel at 1e737, so
    every pixel is exterior."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["subpixel737"]
    val, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    assert (nesc > 0).all()


def test_perturbation_agrees_with_mpfr_brute_force():
    """The rescaled loop against full-precision iteration of every pixel (the
    engine's own ComputeDirect oracle), on a small frame: classes identical, smooth
    values within the engine's default tolerance 0.5 (verify.cpp:85)."""
    cre, cim, scale = "0", "1", cases.pow10(300)
    W, H, mxit = 16, 12, 5000
    val, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit)
    dv, dn = mxo.direct(cre, cim, scale, W, H, mxit)
    assert np.array_equal(nesc > 0, dn > 0)
    ext = nesc > 0
    assert np.abs(val[ext] - dv[ext]).max() < 0.5


def test_ede_against_mpfr_brute_force():
    """Exterior distance estimate (pixel-normalised, mandel_perturbation.cpp:3506-3511)
    against the full-precision derivative (accuratePointCompute's dc, 831-851)."""
    cre, cim, scale = "0", "1", cases.pow10(300)
    W, H, mxit = 12, 8, 5000
    val, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=2)
    dv, dn = mxo.direct(cre, cim, scale, W, H, mxit, c_method=2)
    ext = nesc > 0
    assert np.array_equal(ext, dn > 0)
    rel = np.abs(val[ext] - dv[ext]) / np.abs(dv[ext])
    assert rel.max() < 1e-6


def test_bla_on_matches_bla_off_classification():
    """BLA skipping must not change any class (the method of pbench.cpp:569-595);
    smooth values agree except on a chaotic tail (the reference bounds the same
    tail at 4.0 for deep51, verify.cpp:11793-11799)."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["extref875"]
    v1, n1, _, _ = mxo.compute_fe(cre, cim, scale, W, H, mxit, use_bla=True)
    v0, n0, _, i0 = mxo.compute_fe(cre, cim, scale, W, H, mxit, use_bla=False)
    assert i0["bla_applies"] == 0
    assert np.array_equal(n1 > 0, n0 > 0)
    ext = n1 > 0
    d = np.abs(v1[ext] - v0[ext])
    assert np.median(d) < 1e-6 and (d < 1e-3).mean() > 0.95 and d.max() < 10.0
