"""Adaptive supersampling on the floatexp path: the product's (H sub) x (W sub) cells
(which pixels are refined, every sub-sample's escape count and value) against the
oracle's, and the explicit-sample tile mode against the grid it reproduces."""
import numpy as np
import pytest

import cases
import mxhip
import mxo

pytestmark = pytest.mark.gpu

INT_MIN = np.iinfo(np.int32).min


@pytest.mark.parametrize("name,c_method,sub", [("subpixel737", 1, 3), ("minibrot875", 3, 3), ("extref875", 1, 5)])
def test_supersampled_frame_matches_oracle(ctx, name, c_method, sub):
    cre, cim, scale, W, H, mxit, _ = cases.FRAMES[name]
    print(f"[ss] {name} c_method {c_method} sub {sub}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, c_method=c_method, sub=sub, form=mxhip.MX_FORM_FE64)
    _, _, _, oi = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=c_method, sub=sub)
    cf, cv, cn = oi["cells"]
    assert info["ss_pixels"] == oi["ss_pixels"] > 0
    assert np.array_equal(nesc, cn)
    done = cn != INT_MIN
    ext = done & (cn > 0)
    rel = np.abs(val[ext] - cv[ext]) / np.maximum(np.abs(cv[ext]), 1e-300)
    assert rel.max() < 1e-9
    assert np.array_equal(it[~done], np.full((~done).sum(), -10.0, np.float32))


def test_explicit_samples_equal_grid(ctx):
    """A tile given as an explicit list of (dcr, dci) equal to a grid's dc tables
    gives that grid's results sample for sample."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["subpixel737"]
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm)
    ref = ctx.upload(fv)
    try:
        for form in (mxhip.MX_FORM_FE64, mxhip.MX_FORM_FE32):
            t = ctx.tile(fv, form=form, interior=True)
            grid = ctx.render(ref, t)
            jj, ii = np.meshgrid(np.arange(W), np.arange(H))
            dm = np.empty(2 * W * H)
            de = np.empty(2 * W * H, np.int32)
            dm[0::2], de[0::2] = fv["dcol_m"][jj.ravel()], fv["dcol_e"][jj.ravel()]
            dm[1::2], de[1::2] = fv["drow_m"][ii.ravel()], fv["drow_e"][ii.ravel()]
            ts = ctx.tile(fv, form=form, interior=True)
            ts.n_samples = W * H
            ts.sample_dc_m, ts.sample_dc_e = mxhip._p(dm), mxhip._p(de)
            ts._keep2 = (dm, de)
            lst = ctx.render(ref, ts)
            for a, b in zip(grid[:3], lst[:3]):
                assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    finally:
        ref.release()


def test_adaptive_gui_golden(ctx):
    """The reference's adaptive-supersampling golden (runAdaptiveGuiCase,
    verify.cpp:307-388) through mx_compute: its acceptance, and every cell bit for bit."""
    from test_oracle_ss import gui875_acceptance, gui875_cells

    print("[ss] gui875 4x4 sub 5", flush=True)

    def compute(*a, **kw):
        return ctx.compute(*a, form=mxhip.MX_FORM_FE64, **kw)[0]

    cells = gui875_cells(compute)
    g, done = gui875_acceptance(cells)
    assert done.sum() == 400
    assert np.array_equal(cells.view(np.int32), g.view(np.int32))
