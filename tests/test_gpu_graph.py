"""Captured sequences: the plans of several frames recorded as one HIP graph; every
replay, with eager launches and synchronisations of the program's own between the
replays, gives sample for sample what launching each plan alone gives."""
import numpy as np
import pytest

import cases
import mxhip

pytestmark = pytest.mark.gpu

NAMES = ["exterior1000", "subpixel737", "extref875", "exterior1000_ede"]


@pytest.mark.parametrize("form", [mxhip.MX_FORM_FE32, mxhip.MX_FORM_FE64])
def test_graph_replays_equal_single_launches(ctx, form):
    print(f"[graph] form {form}", flush=True)
    refs, tiles = [], []
    try:
        for name in NAMES:
            cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
            fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm)
            refs.append(ctx.upload(fv))
            tiles.append(ctx.tile(fv, form=form, ede=bool(cm & 2), interior=True))
        alone = []
        for ref, t in zip(refs, tiles):
            p = ctx.plan(ref, t)
            p.launch()
            p.wait()
            alone.append(p.read())
            p.destroy()
        plans = [ctx.plan(ref, t) for ref, t in zip(refs, tiles)]
        g = ctx.graph(plans)
        other = ctx.plan(refs[0], tiles[0])
        for rep in range(3):
            g.launch()
            assert g.wait() > 0
            for k, p in enumerate(plans):
                nesc, val, out, st = p.read()
                a = alone[k]
                assert np.array_equal(nesc, a[0]), (k, rep)
                assert np.array_equal(val.view(np.int64), a[1].view(np.int64)), (k, rep)
                assert np.array_equal(out.view(np.int32), a[2].view(np.int32)), (k, rep)
                assert np.array_equal(st, a[3]), (k, rep)
            # eager work between the replays
            other.launch()
            other.wait()
            assert np.array_equal(other.read()[0], alone[0][0])
        g.destroy()
        other.destroy()
        for p in plans:
            p.destroy()
    finally:
        for r in refs:
            r.release()
