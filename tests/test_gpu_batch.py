"""Batches: tiles of several references, sizes and iteration limits in one persistent
launch give, sample for sample, what launching each tile alone gives; a batch
relaunches with the same results; tiles with and without interior detection mix
(the BLA-off tile runs the engine's no-BLA loop, which has none); mixed forms or EDE
flags and repeated plans are refused."""
import numpy as np
import pytest

import cases
import mxhip

pytestmark = pytest.mark.gpu

# smooth frames of different depth, size and max_iter; a probe-grid tile (stride 8)
# adds an uneven work shape
NAMES = ["exterior1000", "subpixel737", "extref875", "minibrot875_odd"]


def _tiles(ctx, form, ede=False):
    out = []
    for name in NAMES:
        cre, cim, scale, W, H, mxit, _ = cases.FRAMES[name]
        fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=2 if ede else 0)
        ref = ctx.upload(fv)
        out.append((ref, ctx.tile(fv, form=form, ede=ede, interior=True)))
        if name == "subpixel737":
            out.append((ref, ctx.tile(fv, stride=8, form=form, ede=ede, interior=True)))
    return out


def _release(tiles):
    seen = set()
    for ref, _ in tiles:
        if id(ref) not in seen:
            seen.add(id(ref))
            ref.release()


@pytest.mark.parametrize("form,ede", [(mxhip.MX_FORM_FE32, False), (mxhip.MX_FORM_FE64, False),
                                      (mxhip.MX_FORM_FE32, True)])
def test_batch_equals_single_launches(ctx, form, ede):
    print(f"[batch] form {form} ede {ede}", flush=True)
    tiles = _tiles(ctx, form, ede)
    try:
        alone = []
        for ref, t in tiles:
            p = ctx.plan(ref, t)
            p.launch()
            p.wait()
            alone.append(p.read())
            p.destroy()
        plans = [ctx.plan(ref, t) for ref, t in tiles]
        # largest work first, as a caller orders a batch
        order = sorted(range(len(plans)), key=lambda k: -tiles[k][1].width * tiles[k][1].height * tiles[k][1].max_iter)
        b = ctx.batch([plans[k] for k in order])
        for rep in range(2):
            b.launch()
            ms = b.wait()
            assert ms > 0
            for k, p in enumerate(plans):
                nesc, val, out, st = p.read()
                a = alone[k]
                assert np.array_equal(nesc, a[0]), (k, rep)
                assert np.array_equal(val.view(np.int64), a[1].view(np.int64)), (k, rep)
                assert np.array_equal(out.view(np.int32), a[2].view(np.int32)), (k, rep)
                assert np.array_equal(st, a[3]), (k, rep)
        b.destroy()
        for p in plans:
            p.destroy()
    finally:
        _release(tiles)


def test_batch_refusals(ctx):
    tiles = _tiles(ctx, mxhip.MX_FORM_FE32)
    try:
        (r0, t0), (r1, t1) = tiles[0], tiles[1]
        p0 = ctx.plan(r0, t0)
        p1 = ctx.plan(r1, t1)
        with pytest.raises(mxhip.MxError) as e:
            ctx.batch([p0, p0])
        assert e.value.rc == mxhip.MX_E_INVALID
        t64 = ctx.tile(mxhip.frame_prepare(*cases.FRAMES["exterior1000"][:6]), form=mxhip.MX_FORM_FE64, interior=True)
        p2 = ctx.plan(r0, t64)
        with pytest.raises(mxhip.MxError):
            ctx.batch([p0, p2])
        tno = ctx.tile(mxhip.frame_prepare(*cases.FRAMES["exterior1000"][:6]), form=mxhip.MX_FORM_FE32, ede=True)
        p3 = ctx.plan(r0, tno)
        with pytest.raises(mxhip.MxError):
            ctx.batch([p0, p3])
        for p in (p0, p1, p2, p3):
            p.destroy()
    finally:
        _release(tiles)
