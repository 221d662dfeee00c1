"""Parity of the HIP kernels (through the C ABI) with the oracle's CPU restatement of
the engine, on the reference's own test frames.  Escape counts bit-exact; values
(smooth iteration / distance estimate) within 1e-9 relative: the only op that may
differ is log(), computed by the device math library instead of the host's."""
import numpy as np
import pytest

import cases
import mxhip
import mxo

REL = 1e-9
pytestmark = pytest.mark.gpu


def _check(val_g, nesc_g, val_o, nesc_o, label):
    assert nesc_g.shape == nesc_o.shape
    bad = np.argwhere(nesc_g != nesc_o)
    assert bad.size == 0, f"{label}: {len(bad)} escape counts differ, first {bad[:5].tolist()} gpu={nesc_g[tuple(bad[0])]} cpu={nesc_o[tuple(bad[0])]}"
    ext = nesc_o > 0
    d = np.abs(val_g[ext] - val_o[ext]) / np.maximum(np.abs(val_o[ext]), 1e-300)
    assert (d <= REL).all(), f"{label}: max rel value diff {d.max()}"
    assert (val_g[~ext] == -2.0).all()


@pytest.mark.parametrize("name", ["extref875", "minibrot875", "minibrot875_odd", "minibrot875_ede", "exterior1000",
                                  "exterior1000_ede", "subpixel737_ede"])
def test_frame_matches_oracle(ctx, name):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    print(f"[parity] {name} {W}x{H} mxit={mxit} c_method={cm}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, c_method=cm)
    val_o, nesc_o, out_o, info_o = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=cm)
    assert info["ref_it"] == info_o["ref_it"] and info["ref_virtual"] == info_o["ref_virtual"]
    assert (info["period"], info["subpixel"]) == (info_o["period"], info_o["subpixel"])
    assert info["interior_used"] == info_o["interior_used"]
    _check(val, nesc, val_o, nesc_o, name)
    # the engine's float buffer
    assert np.array_equal(it < -1.5, out_o < -1.5)
    assert info["steps"] == info_o["steps"] and info["bla_applies"] == info_o["bla_applies"]
    assert info["bla_skipped"] == info_o["bla_skipped"] and info["rebases"] == info_o["rebases"]


@pytest.mark.parametrize("name,golden", [("minibrot875", "minibrot875.f32"), ("extref875", "minibrot875_extref.f32")])
def test_frame_matches_reference_golden(ctx, name, golden):
    """The reference's own saved frames (tests/golden), same tolerances as verify.cpp:11838-11857."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, c_method=cm)
    g = cases.golden(golden, H, W)
    assert np.array_equal(it < -1.5, g < -1.5)
    ext = ~(g < -1.5)
    assert np.abs(it[ext].astype(np.float64) - g[ext]).max() <= 1.0


def test_bla_off_and_interior_off(ctx):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["extref875"]
    for bla, interior in ((0, 1), (1, 0), (0, 0)):
        it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, use_bla=bla, interior=interior)
        val_o, nesc_o, _, _ = mxo.compute_fe(cre, cim, scale, W, H, mxit, use_bla=bool(bla), interior=bool(interior))
        _check(val, nesc, val_o, nesc_o, f"bla={bla} interior={interior}")


def test_tile_on_product_arrays(ctx):
    """Kernel alone: the same orbit/BLA/dc arrays into the GPU and into the oracle loop."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["minibrot875_odd"]
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit)
    ref = ctx.upload(fv)
    try:
        t = ctx.tile(fv, interior=True)
        nesc, val, out, st = ctx.render(ref, t)
    finally:
        ref.release()
    orb = dict(zf=fv["zf"], zre_m=fv["zfe_m"][0::2], zre_e=fv["zfe_e"][0::2], zim_m=fv["zfe_m"][1::2], zim_e=fv["zfe_e"][1::2])
    assert fv["period"] == 14164
    b = mxo.bla_tables(fv["zf"], fv["ref_it"], fv["dcmax"])
    jj, ii = np.meshgrid(np.arange(W), np.arange(H))
    v2, n2, st2 = mxo.pixels_fe(orb, dict(nlev=b["nlev"], off=b["off"], size=b["size"], r2=b["r2"], l=b["l"], coef=b["coef"], rmax2=b["rmax2"]),
                                fv["dcol_m"][jj.ravel()], fv["dcol_e"][jj.ravel()], fv["drow_m"][ii.ravel()], fv["drow_e"][ii.ravel()],
                                mxit, fv["esc2"], fv["dx_m"], fv["dx_e"], use_interior=True, bla_minlevel=fv["bla_minlevel"],
                                period=fv["period"])
    _check(val, nesc, v2, n2, "minibrot875_odd tile")
    assert int(st[0]) == st2["steps"] and int(st[2]) == st2["bla_skipped"]


def test_periodic_reference_and_full_reference(ctx):
    """The resolved nucleus frame with its periodic reference (the engine's default)
    and with the nucleus search switched off (a full reference from the centre pixel):
    each equals the oracle run the same way; both give the reference's golden classes."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["minibrot875"]
    g = cases.golden("minibrot875.f32", H, W)
    for periodic in (True, False):
        print(f"[parity] minibrot875 periodic {periodic}", flush=True)
        opts = 0 if periodic else mxhip.MX_OPT_NO_PERIODIC
        it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, options=opts)
        val_o, nesc_o, out_o, info_o = mxo.compute_fe(cre, cim, scale, W, H, mxit, periodic=periodic)
        assert info["period"] == info_o["period"] == (14164 if periodic else 0)
        _check(val, nesc, val_o, nesc_o, f"minibrot875 periodic={periodic}")
        assert np.array_equal(it < -1.5, g < -1.5)
