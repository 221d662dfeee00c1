"""The colour pass on the device against the oracle's CPU restatement of render.cpp's
(tools/render.cpp:252-302).  Device powf/logf/tanhf may differ from the host's in the
last place, which moves a palette index or a byte by one at most on a few pixels:
bytes within 2, at least 99 % identical."""
import numpy as np
import pytest

import cases
import mxhip
import mxo

pytestmark = pytest.mark.gpu


def _close(a, b, label):
    d = np.abs(a.astype(int) - b.astype(int))
    assert d.max() <= 2, f"{label}: max byte difference {d.max()}"
    assert (d == 0).mean() >= 0.99, f"{label}: identical fraction {(d == 0).mean()}"


def test_colorize_golden_frame(ctx):
    g = cases.golden("minibrot875.f32", 18, 24)
    _close(ctx.colorize(g), mxo.colorize(g), "minibrot875.f32")


@pytest.mark.parametrize("name,c_method,sub", [("subpixel737", 1, 3), ("subpixel737_ede", 3, 3)])
def test_colorize_supersampled_frame(ctx, name, c_method, sub):
    cre, cim, scale, W, H, mxit, _ = cases.FRAMES[name]
    print(f"[color] {name} sub {sub}", flush=True)
    it, _, _, _ = ctx.compute(cre, cim, scale, W, H, mxit, c_method=c_method, sub=sub)
    for density in (60.0, 250.0):
        _close(ctx.colorize(it, ss=sub, c_method=c_method, density=density),
               mxo.colorize(it, ss=sub, ede=bool(c_method & 2), density=density), f"{name} density {density}")


def test_plan_colorize_equals_host_buffer_colorize(ctx):
    """Colouring a plan's device output gives the bytes that colouring its read-back
    float buffer gives (ss 1 and 3)."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["exterior1000_ede"]
    W, H = 48, 36
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm)
    ref = ctx.upload(fv)
    try:
        t = ctx.tile(fv, ede=True, interior=True)
        p = ctx.plan(ref, t)
        p.launch()
        p.wait()
        _, _, out, _ = p.read()
        for ss in (1, 3):
            a = p.colorize(ss=ss, c_method=cm)
            b = ctx.colorize(out.reshape(H, W), ss=ss, c_method=cm)
            assert np.array_equal(a, b), ss
        p.destroy()
    finally:
        ref.release()
