"""Adaptive supersampling in the oracle: the pixel test (mandel_perturbation.cpp:
1796-1883) on hand-built fields with known answers, and the sub-sample cells against
frames computed directly at the sub-sample positions."""
import math

import numpy as np

import cases
import mxo


def test_flags_constant_field_and_spike():
    mxit = 10000
    f = np.full((9, 11), 5.0, np.float32)
    assert not mxo.ss_flags(f, mxit, False).any()
    # threshold log(mxit)/8: a jump of exp(thr)*1.01 flags, exp(thr)*0.99 does not
    thr = math.exp(math.log(mxit) / 8)
    for k, expect in ((1.01, True), (0.99, False)):
        g = f.copy()
        g[4, 5] += thr * k
        fl = mxo.ss_flags(g, mxit, False)
        assert fl[3:6, 4:7].all() == expect and fl.sum() == (9 if expect else 0)
    # a spike in a corner flags itself and its three neighbours
    g = f.copy()
    g[0, 0] += 100
    fl = mxo.ss_flags(g, mxit, False)
    assert fl.sum() == 4 and fl[:2, :2].all()


def test_flags_exterior_distance_sign_change():
    f = np.full((7, 7), 0.25, np.float32)
    assert not mxo.ss_flags(f, 1000, True).any()
    f[3, 3] = -2.0                      # an interior pixel among exterior ones
    fl = mxo.ss_flags(f, 1000, True)
    assert fl.sum() == 9 and fl[2:5, 2:5].all()
    f[:, :] = -2.0                      # all interior: no sign change
    assert not mxo.ss_flags(f, 1000, True).any()


def test_subsample_cells_equal_direct_samples():
    """Every cell of a refined pixel is the per-pixel loop at dc = dx (j + v/3 - ref_x),
    dy (i + u/3 - ref_y) (the FloatExp product of the double offset, mandel_
    perturbation.cpp:2160-2163), formed here independently and run through the
    oracle's per-pixel loop on the frame's own orbit; unrefined pixels keep only their
    centre cell, which is the base frame."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["subpixel737"]
    W, H = 9, 7
    _, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=1, sub=3, use_bla=False,
                                        want_orbit=True)
    cf, cv, cn = info["cells"]
    filled = cn != np.iinfo(np.int32).min
    assert info["ss_pixels"] > 0 and not info["ref_virtual"]
    blocks = filled.reshape(H, 3, W, 3).sum(axis=(1, 3))
    assert set(np.unique(blocks)) <= {1, 9} and (blocks == 9).sum() == info["ss_pixels"]
    assert np.array_equal(cn[1::3, 1::3], nesc)
    # the refined pixels' sub-samples, recomputed
    (dxm, dxe), (dym, dye) = mxo.grid_info(cre, cim, scale, W, H)
    rx, ry = info["ref_x"], info["ref_y"]
    ii, jj = np.nonzero(blocks == 9)
    pos = [(i, j, u, v) for i, j in zip(ii, jj) for u in (-1, 0, 1) for v in (-1, 0, 1) if u or v]
    offx = np.array([j + v / 3 - rx for i, j, u, v in pos])
    offy = np.array([i + u / 3 - ry for i, j, u, v in pos])
    n = len(pos)
    am, ae = mxo.fe_op(1, np.full(n, dxm), np.full(n, dxe, np.int64), offx, None)
    bm, be = mxo.fe_op(1, np.full(n, dym), np.full(n, dye, np.int64), offy, None)
    o = info["orbit"]
    orb = dict(zf=o["zf"], zre_m=o["zfe_m"][0::2], zre_e=o["zfe_e"][0::2], zim_m=o["zfe_m"][1::2], zim_e=o["zfe_e"][1::2])
    esc2 = float(np.float32(1e8) * np.float32(1e8))  # the default bailout's float square
    v2, n2, _ = mxo.pixels_fe(orb, None, am, ae, bm, be, mxit, esc2, dxm, dxe, use_interior=bool(info["interior_used"]))
    cells = [(3 * i + 1 + u, 3 * j + 1 + v) for i, j, u, v in pos]
    assert np.array_equal(n2, np.array([cn[c] for c in cells]))
    assert np.array_equal(v2, np.array([cv[c] for c in cells]))


def gui875_cells(compute):
    """runAdaptiveGuiCase (verify.cpp:307-388): deep1 at 1e875, 4x4 pixels, sub 5,
    max_iter 300000, SUPER_SAMPLING; returns the (20, 20) float cells"""
    return compute(cases.LOC["deep1_x"], cases.LOC["deep1_y"], cases.pow10(875), 4, 4, 300000, c_method=1, sub=5)


def gui875_acceptance(cells):
    """verify.cpp:338-378's acceptance against minibrot875_gui_ss5.f32: every base cell
    computed, some sub-cells computed, no class mismatch, max difference <= 1.0"""
    g = cases.golden("minibrot875_gui_ss5.f32", 20, 20)
    done = cells != -10.0
    base = np.zeros_like(done)
    base[2::5, 2::5] = True
    assert done[base].all() and (done & ~base).any()
    gd = g[done]
    ours = cells[done]
    assert np.array_equal(ours < 0, gd < 0)
    ext = ours >= 0
    assert np.abs(ours[ext].astype(np.float64) - gd[ext]).max() <= 1.0
    return g, done


def test_oracle_reproduces_the_adaptive_gui_golden():
    """The golden of the engine's production GUI path with adaptive supersampling: the
    nucleus in view is sub-pixel at this size, so BLA is off and the engine's no-BLA
    vector loop computes every sample (solveRescaledSimd8NoBla, mandel_perturbation.cpp:
    3620-3816).  The oracle passes verify.cpp's acceptance and reproduces all 400 cells
    bit for bit, including the sub-samples whose step-underflow recovery follows that
    loop's rule (|w'|^2 == 0) rather than pixelRescaled's."""
    def compute(*a, **kw):
        return mxo.compute_fe(*a, **kw)[3]["cells"][0]

    cells = gui875_cells(compute)
    g, done = gui875_acceptance(cells)
    assert done.sum() == 400
    assert np.array_equal(cells.view(np.int32), g.view(np.int32))
