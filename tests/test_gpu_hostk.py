"""The GPU kernels against the host build of their own per-sample code
(tests/host_kernels.h): same escape counts bit for bit, same step counters, values
equal up to the device's logarithm (the final smooth / distance value is the only
call into a math library).  This ties every host-side finding of
tests/test_host_kernels.py to the device."""
import numpy as np
import pytest

import cases
import hostk
import mxhip

pytestmark = pytest.mark.gpu

FORMS = {"fe64": mxhip.MX_FORM_FE64, "fe32": mxhip.MX_FORM_FE32, "sc64": mxhip.MX_FORM_SC64}
FRAMES = ["minibrot875", "extref875", "subpixel737", "exterior1000", "minibrot875_ede", "exterior1000_ede"]


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("name", FRAMES)
def test_device_equals_host_build(ctx, name, form):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    print(f"[hostk] {name} {W}x{H} form {form}", flush=True)
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm)
    ede = cm == 2
    ref = ctx.upload(fv)
    try:
        for interior in (True, False):
            nd, vd, _, std = ctx.render(ref, ctx.tile(fv, form=FORMS[form], ede=ede, interior=interior))
            nh, vh, sth = hostk.run(fv, FORMS[form], ede=ede, interior=interior)
            assert np.array_equal(nd, nh), interior
            assert np.array_equal(std, sth.sum(axis=0)), interior
            ext = nh > 0
            assert np.max(np.abs(vd[ext] - vh[ext]) / np.abs(vh[ext])) < 1e-13, interior
    finally:
        ref.release()
