"""Edge sizes through mx_compute on every path, against the oracle: the smallest frames
the request validation admits (2 x 2; one-row strips of taller frames,
compute_backend.cpp:263-266), column pairs, ragged sizes (edge tiles of the 8x8 work
items cut on both axes), max_iter 1 and 2, and the largest supersampling factor (sub 15).
Escape counts bit-exact, values within 1e-9 relative."""
import numpy as np
import pytest

import cases
import mxhip
import mxo

pytestmark = pytest.mark.gpu
REL = 1e-9
DEEP = (cases.LOC["deep1_x"], cases.LOC["deep1_y"], cases.pow10(875))     # floatexp path
DBL = ("-0.743643887037151", "0.131825904205330", "40000000000")           # plain-double path
SHALLOW = ("-0.5", "0", "1")                                              # shallow path


def _check(val, nesc, oval, onesc, label):
    assert nesc.shape == onesc.shape, label
    bad = np.argwhere(nesc != onesc)
    assert bad.size == 0, f"{label}: {len(bad)} escape counts differ, first {bad[:4].tolist()}"
    ext = onesc > 0
    rel = np.abs(val[ext] - oval[ext]) / np.maximum(np.abs(oval[ext]), 1e-300)
    assert rel.max(initial=0.0) <= REL, f"{label}: value rel diff {rel.max()}"


def _oracle(path, cre, cim, scale, W, H, mxit, c_method=0, sub=1, full_h=0, row_base=0):
    if path == "fe":
        v, n, _, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=c_method, sub=sub, full_h=full_h,
                                       row_base=row_base)
        return (info["cells"][1], info["cells"][2]) if sub > 1 else (v, n)
    if path == "dbl":
        v, n, _, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit, c_method=c_method, sub=sub)
        return (info["cell_val"], info["cell_nesc"]) if sub > 1 else (v, n)
    v, n, _ = mxo.shallow(cre, cim, scale, W, H, mxit, ede=bool(c_method & 2), full_h=full_h, row_base=row_base)
    return v, n


@pytest.mark.parametrize("path,frame,mxit", [("fe", DEEP, 300000), ("dbl", DBL, 8192), ("shallow", SHALLOW, 5000)])
@pytest.mark.parametrize("W,H", [(2, 2), (3, 2), (2, 33), (13, 67), (97, 3)])
def test_small_and_ragged_frames(ctx, path, frame, mxit, W, H):
    cre, cim, scale = frame
    print(f"[edges] {path} {W}x{H}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit)
    ov, on = _oracle(path, cre, cim, scale, W, H, mxit)
    _check(val, nesc, np.asarray(ov).reshape(H, W), np.asarray(on).reshape(H, W), f"{path} {W}x{H}")
    assert not (it == -10).any()


@pytest.mark.parametrize("path,frame,mxit", [("fe", DEEP, 300000), ("shallow", SHALLOW, 5000)])
@pytest.mark.parametrize("W,full_h,row_base", [(97, 9, 4), (2, 2, 1), (2, 2, 0)])
def test_one_row_strips(ctx, path, frame, mxit, W, full_h, row_base):
    """A strip one row high of a taller frame (the engine's fullHeight / rowBase)."""
    cre, cim, scale = frame
    print(f"[edges] {path} strip {W}x1 of {full_h} rows at {row_base}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, 1, mxit, full_h=full_h, row_base=row_base)
    ov, on = _oracle(path, cre, cim, scale, W, 1, mxit, full_h=full_h, row_base=row_base)
    _check(val, nesc, np.asarray(ov).reshape(1, W), np.asarray(on).reshape(1, W), f"{path} strip")


@pytest.mark.parametrize("path,frame", [("fe", DEEP), ("dbl", DBL), ("shallow", SHALLOW)])
@pytest.mark.parametrize("mxit", [1, 2])
def test_tiny_iteration_limits(ctx, path, frame, mxit):
    cre, cim, scale = frame
    print(f"[edges] {path} max_iter {mxit}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, 24, 16, mxit)
    ov, on = _oracle(path, cre, cim, scale, 24, 16, mxit)
    _check(val, nesc, np.asarray(ov).reshape(16, 24), np.asarray(on).reshape(16, 24), f"{path} max_iter {mxit}")


@pytest.mark.parametrize("path,frame,mxit,cm", [("fe", DEEP, 300000, 1), ("dbl", DBL, 8192, 1)])
def test_largest_supersampling(ctx, path, frame, mxit, cm):
    """sub 15, the largest the request accepts: the (15 H) x (15 W) cells, every flagged
    pixel's 224 other cells."""
    cre, cim, scale = frame
    W, H = 6, 4
    print(f"[edges] {path} sub 15", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, c_method=cm, sub=15)
    ov, on = _oracle(path, cre, cim, scale, W, H, mxit, c_method=cm, sub=15)
    done = on != np.iinfo(np.int32).min
    assert np.array_equal(it != -10, done) and info["ss_pixels"] > 0
    _check(val[done], nesc[done], ov[done], on[done], f"{path} sub 15")
