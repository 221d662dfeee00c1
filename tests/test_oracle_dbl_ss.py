"""Pins the oracle's adaptive supersampling on the plain-double path
(Mandel::Compute, mandel_perturbation.cpp:1789-2005): the flagged pixels' other cells
run through stepParallel's double loop against the current reference -- the first pass
with its BLA table, later passes after a glitch re-reference at a sub-sample with BLA
off -- with dc = trunc53(d (cell - reference cell) / sub) (2300-2312).

Pinned by: the base cells equal the frame without supersampling; every refined cell
against the MPFR brute force at that sub-sample's own c (formed here in decimal
arithmetic from the grid formulas, MP.cpp:1403-1416), at verify.cpp's default
acceptance (maxDiff 0.5, at most 1 % class mismatches); a frame that re-references
among its sub-samples; the EDE sign-change test."""
from decimal import Decimal, getcontext

import numpy as np
import pytest

import cases
import mxo

TICKTOCK140 = (cases.LOC["ticktock_x"], cases.LOC["ticktock_y"], cases.pow10(140), 24, 16, 100000)
# the period-106 minibrot the nucleus search finds in the ticktock view at scale 1e7
# (island size ~5e-10), seen at scale 2e9: interior and exterior pixels side by side
MINI106 = ("-1.4799461355795131264511335866615963781", "0.00090132482780272173726842509958825902", "2000000000",
           24, 16, 20000)


def cell_c(cre, cim, scale, W, H, i, j, u, v, sub):
    """c of sample (i, j, u, v): c0 + d (k + offset / sub), the engine's grid
    (dw = 2 / scale, dh = dw H / W, c0 = centre - (dw, dh), d = 2 dw / (W - 1),
    2 dh / (H - 1))."""
    getcontext().prec = len(scale) + 60
    dw = Decimal(2) / Decimal(scale)
    dh = dw / W * H
    dx, dy = dw / (W - 1) * 2, dh / (H - 1) * 2
    re = Decimal(cre) - dw + dx * j + dx * v / sub
    im = Decimal(cim) - dh + dy * i + dy * u / sub
    return format(re, "f"), format(im, "f")


def refined_cells(info, W, H, sub):
    """(i, j, u, v) of every computed cell that is not a pixel centre"""
    cn = info["cell_nesc"]
    h = sub // 2
    out = []
    for r, c in zip(*np.nonzero(cn != np.iinfo(np.int32).min)):
        i, u = divmod(int(r), sub)
        j, v = divmod(int(c), sub)
        if u != h or v != h:
            out.append((i, j, u - h, v - h))
    return out


def test_base_cells_equal_the_plain_frame():
    cre, cim, scale, W, H, mxit = TICKTOCK140
    _, _, out1, _ = mxo.compute_dbl(cre, cim, scale, W, H, mxit)
    _, _, out3, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit, c_method=1, sub=3)
    assert np.array_equal(info["cells"][1::3, 1::3], out1) and np.array_equal(out3, out1)
    # refined pixels have all 9 cells, the others only their centre
    filled = (info["cell_nesc"] != np.iinfo(np.int32).min).reshape(H, 3, W, 3).sum(axis=(1, 3))
    assert set(np.unique(filled)) <= {1, 9} and (filled == 9).sum() == info["ss_pixels"] > 0


def test_refined_cells_against_brute_force_with_a_sub_sample_reference():
    """ticktock at 1e140: the first sub-sample pass leaves glitched sub-samples, the
    engine re-references at the one with the smallest glitch metric and runs a second
    pass with BLA off; sampled refined cells against the MPFR brute force."""
    cre, cim, scale, W, H, mxit = TICKTOCK140
    _, _, _, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit, c_method=1, sub=3)
    assert info["ss_passes"] >= 2
    cells = refined_cells(info, W, H, 3)
    rng = np.random.default_rng(3)
    pick = [cells[k] for k in rng.choice(len(cells), 48, replace=False)]
    wp = mxo.precision_bits(scale) + 64
    mis, dmax = 0, 0.0
    for i, j, u, v in pick:
        cr, ci = cell_c(cre, cim, scale, W, H, i, j, u, v, 3)
        bv, bn = mxo.point(cr, ci, wp, mxit)
        ours = float(info["cells"][3 * i + 1 + u, 3 * j + 1 + v])
        if (bn > 0) != (ours >= 0):
            mis += 1
        elif bn > 0:
            dmax = max(dmax, abs(ours - bv))
    assert mis <= 0.01 * len(pick) and dmax <= 0.5, (mis, dmax)


def test_distance_estimate_refines_sign_changes_only():
    """Under EDE a pixel is refined when a neighbour has the opposite sign (MP.cpp:
    1836-1845): the period-106 minibrot's edge; the refined cells' classes agree with
    the MPFR brute force."""
    cre, cim, scale, W, H, mxit = MINI106
    _, nesc, out, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit, c_method=3, sub=3)
    flags = mxo.ss_flags(out, mxit, True).reshape(H, W)
    assert info["ss_pixels"] == int(flags.sum()) > 0
    neg = out < 0
    for i, j in zip(*np.nonzero(flags)):
        nb = neg[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2]
        assert nb.any() and (~nb).any()
    cells = refined_cells(info, W, H, 3)
    wp = mxo.precision_bits(scale) + 64
    rng = np.random.default_rng(5)
    mis = 0
    for k in rng.choice(len(cells), min(40, len(cells)), replace=False):
        i, j, u, v = cells[k]
        cr, ci = cell_c(cre, cim, scale, W, H, i, j, u, v, 3)
        _, bn = mxo.point(cr, ci, wp, mxit)
        mis += (bn > 0) != (info["cells"][3 * i + 1 + u, 3 * j + 1 + v] >= 0)
    assert mis <= 1
