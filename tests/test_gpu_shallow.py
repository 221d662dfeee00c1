"""The shallow direct path (scale <= 1e6, the engine's method 0,
mandel_perturbation.cpp:1444-1519, 1933-1965) through mx_compute on the GPU: bit for bit
the oracle's floatPointCompute / solveShallowSimdList values on verify.cpp's whole-set
view (smooth and EDE), an odd-size boundary view, strips against the full frame,
adaptive supersampling (sub 3 and 5), and the shallow frames of the zoom sequence
(1280 x 720 around the seed-42 boundary-walk centre, scale 1, 124, 15457)."""
import json
import os

import numpy as np
import pytest

import mxhip
import mxo

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INT_MIN = np.iinfo(np.int32).min
WALK42 = json.load(open(os.path.join(ROOT, "tests", "golden", "walk_de_seed42.json")))["500"]


REL = 1e-12


def _same(it, val, nesc, oval, onesc):
    """Escape counts bit-exact; values within REL relative (log() is the device math
    library's, not the host's); the engine floats equal, or one float ulp apart where
    the values differ."""
    nesc, onesc, val, oval, it = nesc.ravel(), onesc.ravel(), val.ravel(), oval.ravel(), it.ravel()
    assert np.array_equal(nesc, onesc)
    ext = onesc > 0
    a, b = val[ext], oval[ext]
    diff = a != b
    assert (np.abs(a[diff] - b[diff]) <= REL * np.abs(b[diff])).all()
    assert (val[~ext] == -2.0).all() and (oval[~ext] == -2.0).all()
    ulps = np.abs(it.view(np.int32).astype(np.int64) - oval.astype(np.float32).view(np.int32).astype(np.int64))
    assert ulps.max() <= 1
    assert (ulps[val == oval] == 0).all()


@pytest.mark.parametrize("c_method", [0, 2])
def test_whole_set_view(ctx, c_method):
    W, H, mxit = 120, 80, 5000
    print(f"[shallow] whole set c_method {c_method}", flush=True)
    it, val, nesc, info = ctx.compute("-0.5", "0", "1", W, H, mxit, c_method=c_method)
    assert info["path"] == 0
    oval, onesc, _ = mxo.shallow("-0.5", "0", "1", W, H, mxit, ede=c_method == 2)
    _same(it, val, nesc, oval, onesc)
    assert (nesc == -2).sum() > 100 and (nesc == -1).sum() > 0 and (nesc > 0).sum() > 1000
    assert info["steps"] > 0


def test_odd_size_boundary_view_and_strips(ctx):
    cre, cim, scale, W, H, mxit = WALK42["re"], WALK42["im"], "15468", 37, 23, 42832
    print("[shallow] seed-42 centre at 15468, 37x23", flush=True)
    it, val, nesc, _ = ctx.compute(cre, cim, scale, W, H, mxit)
    oval, onesc, _ = mxo.shallow(cre, cim, scale, W, H, mxit)
    _same(it, val, nesc, oval, onesc)
    # rows 8..15 as a strip of the 23-row frame: the strip's own origin c0 + 8 dy is
    # formed in mpf before the truncation to double (MP.cpp:1417-1423), so its c differ
    # from the full frame's c0 + dy i in the last bits; the strip equals the oracle's
    # strip, and the full frame's rows away from chaotic pixels
    its, vals, nescs, _ = ctx.compute(cre, cim, scale, W, 8, mxit, full_h=H, row_base=8)
    sv, sn, _ = mxo.shallow(cre, cim, scale, W, 8, mxit, full_h=H, row_base=8)
    _same(its, vals, nescs, sv, sn)
    assert (nescs == nesc[8:16]).mean() > 0.9


@pytest.mark.parametrize("c_method,sub", [(1, 3), (3, 5)])
def test_adaptive_supersampling(ctx, c_method, sub):
    """Base cells, then the sub*sub - 1 cells of each flagged pixel; unflagged cells
    keep the empty-pixel value -10."""
    cre, cim, scale, W, H, mxit = "-0.7435", "0.1314", "300", 48, 32, 20000
    ede = (c_method & 2) != 0
    print(f"[shallow] ss c_method {c_method} sub {sub}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, c_method=c_method, sub=sub)
    bval, bnesc, _ = mxo.shallow(cre, cim, scale, W, H, mxit, ede=ede)
    base = bval.astype(np.float32).reshape(H, W)
    flags = mxo.ss_flags(base, mxit, ede).ravel()
    flagged = np.nonzero(flags)[0].astype(np.int32)
    assert info["ss_pixels"] == len(flagged) > 0
    cval, cnesc, _ = mxo.shallow(cre, cim, scale, W, H, mxit, ede=ede, sub=sub, flagged=flagged)
    ev = np.full((H * sub, W * sub), -10.0)
    en = np.full((H * sub, W * sub), INT_MIN, np.int32)
    h = sub // 2
    ev[h::sub, h::sub] = bval.reshape(H, W)
    en[h::sub, h::sub] = bnesc.reshape(H, W)
    per = sub * sub - 1
    for q in range(len(flagged) * per):
        pix, t = flagged[q // per], q % per
        if t >= h * sub + h:
            t += 1
        xi, yi = t // sub - h, t % sub - h
        i, j = pix // W, pix % W
        ev[i * sub + h + yi, j * sub + h + xi] = cval[q]
        en[i * sub + h + yi, j * sub + h + xi] = cnesc[q]
    done = en != INT_MIN
    assert np.array_equal(nesc, en)
    _same(it[done], val[done], nesc[done], ev[done], en[done])
    assert (it[~done] == -10.0).all()


@pytest.mark.parametrize("k,scale,mxit", [(0, "1", 1000), (1, "124", 21916), (2, "15457", 42833)])
def test_zoom_sequence_shallow_frames(ctx, k, scale, mxit):
    """Frames 0-2 of the 240-frame zoom (scale <= 1e6), full size."""
    W, H = 1280, 720
    print(f"[shallow] zoom frame {k} scale {scale}", flush=True)
    it, val, nesc, info = ctx.compute(WALK42["re"], WALK42["im"], scale, W, H, mxit)
    oval, onesc, _ = mxo.shallow(WALK42["re"], WALK42["im"], scale, W, H, mxit)
    _same(it, val, nesc, oval, onesc)
    print(f"[shallow] frame {k}: kernel {info['kernel_ms']:.2f} ms, {info['steps']} iterations", flush=True)
