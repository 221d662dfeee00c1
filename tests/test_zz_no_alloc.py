"""mx_tile_render_noalloc, the entry point that allocates nothing: on two calls in a row
(the second with a smaller tile, against another reference) it gives, sample for
sample, what mx_tile_render gives for the same tile, in every floatexp form; and it
refuses tiles beyond its fixed scratch area's limits."""
import ctypes as C

import numpy as np
import pytest

import cases
import mxhip

pytestmark = pytest.mark.gpu

MAX_SAMPLES = 1 << 20
MAX_TABLE = 16384


def _same(a, b, what):
    nesc, val, out, st = a
    n2, v2, o2, s2 = b
    assert np.array_equal(nesc, n2), what
    assert np.array_equal(val.view(np.int64), v2.view(np.int64)), what
    assert np.array_equal(out.view(np.int32), o2.view(np.int32)), what
    assert np.array_equal(st, s2), what


def _frame(name, W=None, H=None):
    cre, cim, scale, W0, H0, mxit, cm = cases.FRAMES[name]
    return mxhip.frame_prepare(cre, cim, scale, W or W0, H or H0, mxit, c_method=cm)


@pytest.mark.parametrize("form,ede", [(mxhip.MX_FORM_FE64, False), (mxhip.MX_FORM_FE32, False),
                                      (mxhip.MX_FORM_SC64, False), (mxhip.MX_FORM_FE64, True)])
def test_noalloc_equals_render_twice(ctx, form, ede):
    # first a 256 x 192 view at 1e-1000, then the 37 x 23 view of a minibrot at 1e-875
    # with its reference pixel inside the tile
    big = mxhip.frame_prepare("0", "1", cases.pow10(1000), 256, 192, 10000, c_method=2 if ede else 0)
    small = _frame("minibrot875_odd")
    for k, fv in enumerate((big, small)):
        print(f"[noalloc] form {form} ede {ede} call {k} {fv['W']}x{fv['H']}", flush=True)
        ref = ctx.upload(fv)
        try:
            t = ctx.tile(fv, form=form, ede=ede, interior=True)
            if not fv.get("ref_virtual", 1):
                t.skip_sample = int(fv["ref_y"]) * fv["W"] + int(fv["ref_x"])
            want = ctx.render(ref, t)
            got = ctx.render_noalloc(ref, t)
            _same(got, want, (form, ede, k))
            assert want[0].max() > 0
        finally:
            ref.release()


def test_noalloc_probe_grid_and_no_bla(ctx):
    """A stride-8 probe grid and a BLA-off tile (the engine's no-BLA loop) as well."""
    fv = _frame("subpixel737")
    ref = ctx.upload(fv)
    try:
        for kw in (dict(stride=8), dict(bla=False), dict(stride=1)):
            print(f"[noalloc] subpixel737 {kw}", flush=True)
            t = ctx.tile(fv, form=mxhip.MX_FORM_FE64, interior=True, **kw)
            _same(ctx.render_noalloc(ref, t), ctx.render(ref, t), kw)
    finally:
        ref.release()


def test_noalloc_refuses_beyond_its_limits(ctx):
    fv = _frame("exterior1000")
    ref = ctx.upload(fv)
    try:
        def tile(W, H):
            t = ctx.tile(fv, form=mxhip.MX_FORM_FE64)
            cm, ce = np.zeros(W), np.zeros(W, np.int32)
            rm, re = np.zeros(H), np.zeros(H, np.int32)
            t._keep = [cm, ce, rm, re]
            t.width, t.height = W, H
            t.dcol_m, t.dcol_e = cm.ctypes.data_as(C.c_void_p), ce.ctypes.data_as(C.c_void_p)
            t.drow_m, t.drow_e = rm.ctypes.data_as(C.c_void_p), re.ctypes.data_as(C.c_void_p)
            return t

        for W, H in ((1025, 1024), (MAX_TABLE, 2)):
            with pytest.raises(mxhip.MxError) as e:
                ctx.render_noalloc(ref, tile(W, H))
            assert "limits" in str(e.value)
        t = ctx.tile(fv, form=mxhip.MX_FORM_DOUBLE) if "dcol_d" in fv else None
        if t is not None:
            with pytest.raises(mxhip.MxError):
                ctx.render_noalloc(ref, t)
        # a refused call leaves the entry point working
        t = ctx.tile(fv, form=mxhip.MX_FORM_FE64)
        _same(ctx.render_noalloc(ref, t), ctx.render(ref, t), "after refusal")
    finally:
        ref.release()
