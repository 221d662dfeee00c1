"""The cutoff-sensitive fallback through mx_compute on the GPU (MP.cpp:2173-2183, set
at 4144 and 4177): a 24 x 16 view at 1e-48 around the seed-7 boundary-walk centre,
referenced at the exact view centre (which escapes at 1835), with max_iter just above
and just below that escape; the frame is flagged, the same cells as the oracle's are
recomputed in arbitrary precision, and every escape count equals the oracle's."""
import json
import os

import numpy as np
import pytest

import mxhip
import mxo

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WALK = json.load(open(os.path.join(ROOT, "tests", "golden", "walk_de_seed7.json")))["50"]
SCALE = "1" + "0" * 48
W, H = 24, 16
OPTS = mxhip.MX_OPT_VIEW_CENTER_REF | mxhip.MX_OPT_FLOATEXP | mxhip.MX_OPT_NO_PERIODIC
REL = 1e-9


@pytest.mark.parametrize("mxit", [1840, 1805, 1635])
@pytest.mark.parametrize("c_method,sub", [(0, 1), (2, 1), (1, 3)])
def test_cutoff_fallback_matches_oracle(ctx, mxit, c_method, sub):
    print(f"[cutoff] mxit {mxit} c_method {c_method} sub {sub}", flush=True)
    it, val, nesc, info = ctx.compute(WALK["re"], WALK["im"], SCALE, W, H, mxit, c_method=c_method, sub=sub,
                                      options=OPTS, form=mxhip.MX_FORM_FE64)
    v, n, o, oi = mxo.compute_fe(WALK["re"], WALK["im"], SCALE, W, H, mxit, c_method=c_method, periodic=False, sub=sub)
    assert info["ref_virtual"] == 1
    assert info["cutoff_sensitive"] == oi["cutoff_sensitive"] == (mxit != 1635)
    assert info["gmp_fallback"] == oi["gmp_fallback"]
    if mxit != 1635:
        assert info["gmp_fallback"] > 0
    if sub > 1:
        cf, cv, cn = oi["cells"]
    else:
        cv, cn = v, n
    assert np.array_equal(nesc, cn)
    ext = cn > 0
    assert (np.abs(val[ext] - cv[ext]) <= REL * np.abs(cv[ext])).all()
