"""The nucleus search below 1e280 (the engine's MANDEL_PERIODIC=1 with MANDEL_FE=1:
mandel_perturbation.cpp:1240-1245, 1573-1576) on the GPU: the period-524 minibrot near
the deep1 location, resolved in a 48 x 32 view at 1e-36, is found and taken as the
reference (its orbit one period long, read modulo the period); the frame, smooth and
exterior DE, equals the oracle's with the same switches.  Without MX_OPT_FLOATEXP such
a request is refused (the plain-double path has no periodic reference here)."""
import numpy as np
import pytest

import cases
import mxhip
import mxo

pytestmark = pytest.mark.gpu
REL = 1e-9


def _view():
    p, nuc, size = mxo.nucleus(cases.LOC["deep1_x"], cases.LOC["deep1_y"], "1" + "0" * 20, 48, 32, 200000)
    assert p == 524
    return nuc[0], nuc[1], "1" + "0" * 36


@pytest.mark.parametrize("c_method", [0, 2])
def test_forced_periodic_reference_matches_oracle(ctx, c_method):
    cre, cim, scale = _view()
    print(f"[periodic] 1e36 view, c_method {c_method}", flush=True)
    opts = mxhip.MX_OPT_FLOATEXP | mxhip.MX_OPT_PERIODIC
    it, val, nesc, info = ctx.compute(cre, cim, scale, 48, 32, 200000, c_method=c_method, options=opts)
    ov, on, _, oi = mxo.compute_fe(cre, cim, scale, 48, 32, 200000, c_method=c_method, force_fe=True, force_periodic=True)
    assert info["period"] == oi["period"] == 524 and info["subpixel"] == 0
    assert np.array_equal(nesc, on)
    ext = on > 0
    assert (np.abs(val[ext] - ov[ext]) <= REL * np.abs(ov[ext])).all()
    assert (on < 0).sum() > 50 and ext.sum() > 500


def test_periodic_on_the_double_path_is_refused(ctx):
    cre, cim, scale = _view()
    with pytest.raises(mxhip.MxError) as e:
        ctx.compute(cre, cim, scale, 48, 32, 200000, options=mxhip.MX_OPT_PERIODIC)
    assert e.value.rc == mxhip.MX_E_UNSUPPORTED
