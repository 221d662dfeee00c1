"""Pins the oracle's cutoff-sensitive fallback (MP.cpp:2173-2183, set at 4144 and
4177): a view-centre reference escaping within 16 of max_iter, or bounded but escaping
by max_iter + 64, marks the frame; every sample whose float value is negative or above
max_iter - 16 is then the MPFR brute force at its own c (checked against the oracle's
separate brute-force frame).  A reference bounded well past max_iter + 64 leaves the
frame alone."""
import json
import os

import numpy as np
import pytest

import mxo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WALK = json.load(open(os.path.join(ROOT, "tests", "golden", "walk_de_seed7.json")))["50"]
SCALE = "1" + "0" * 48
W, H = 24, 16
REF_IT = 1835  # escape count of the view centre of this 24 x 16 view


def frame(mxit, c_method=0, sub=1):
    return mxo.compute_fe(WALK["re"], WALK["im"], SCALE, W, H, mxit, c_method=c_method, periodic=False, sub=sub)


def test_view_centre_escape_count():
    v, n, o, info = frame(400000)
    assert info["ref_virtual"] == 1 and info["ref_escaped"] == 1 and info["ref_it"] == REF_IT


@pytest.mark.parametrize("mxit,hint", [(REF_IT + 5, 1), (REF_IT - 30, 2)])
@pytest.mark.parametrize("c_method", [0, 2])
def test_fallback_cells_are_the_brute_force(mxit, hint, c_method):
    v, n, o, info = frame(mxit, c_method)
    assert info["cutoff_sensitive_hint"] == hint and info["cutoff_sensitive"] == 1
    fb = (o < 0) | (o > mxit - 16)
    assert info["gmp_fallback"] == int(fb.sum()) > 10
    idx = np.nonzero(fb.ravel())[0].astype(np.int32)
    dv, dn = mxo.direct_list(WALK["re"], WALK["im"], SCALE, W, H, mxit, idx, c_method=c_method)
    assert np.array_equal(n.ravel()[idx], dn)
    assert np.array_equal(v.ravel()[idx][dn > 0], dv[dn > 0])


def test_far_bounded_reference_is_not_sensitive():
    v, n, o, info = frame(REF_IT - 200)
    assert info["ref_escaped"] == 0 and info["cutoff_sensitive_hint"] == 2
    assert info["cutoff_sensitive"] == 0 and info["gmp_fallback"] == 0


def test_supersampled_cells_take_the_fallback():
    mxit = REF_IT + 5
    v, n, o, info = frame(mxit, c_method=1, sub=3)
    cf, cv, cn = info["cells"]
    done = cn != np.iinfo(np.int32).min
    fb = done & ((cf < 0) | (cf > mxit - 16))
    assert info["ss_pixels"] > 0 and info["gmp_fallback"] >= int(fb.sum()) > 0
