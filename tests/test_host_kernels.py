"""The floatexp kernels' own per-sample code, built for the host (tests/host_kernels.h),
on CPU threads: the lean-step loop against the full step alone, the plain form
against the oracle, and the accuracy of the mixed form's representation choices.

The device runs this same code (tests/test_gpu_hostk.py checks device == host bit for
bit), so what is shown here holds for the GPU kernels."""
import numpy as np
import pytest

import cases
import hostk
import mxhip
import mxo

FE64, FE32, SC64 = mxhip.MX_FORM_FE64, mxhip.MX_FORM_FE32, mxhip.MX_FORM_SC64
_FV = {}


def prepared(name):
    if name not in _FV:
        cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
        _FV[name] = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm)
    return _FV[name]


def cls(n):
    """n_esc with the two non-escaping outcomes (max_iter, detected interior) merged:
    the engine's float buffer holds -2 for both"""
    return np.where(n > 0, n, -1)


def identical(a, b):
    return float((cls(a) == cls(b)).mean())


# frames: a periodic-nucleus reference (minibrot875), a full reference with rebases
# (extref875), a sub-pixel nucleus with BLA off (subpixel737), a Misiurewicz view
FRAMES = ["minibrot875", "extref875", "subpixel737", "exterior1000"]


@pytest.mark.parametrize("name", FRAMES + ["minibrot875_ede", "exterior1000_ede"])
def test_host_plain_form_equals_oracle(name):
    """The plain form's code on the host gives the oracle's escape counts and values
    bit for bit (both call the host libm for the final logarithm)."""
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    fv = prepared(name)
    val_o, nesc_o, _, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=cm)
    ne, v, _ = hostk.run(fv, FE64, ede=cm == 2, interior=bool(info["interior_used"]))
    m = np.ones(W * H, bool)
    if not fv["ref_virtual"]:
        m[(H // 2) * W + W // 2] = False  # the reference pixel comes from the orbit itself
    assert np.array_equal(ne[m], nesc_o.ravel()[m])
    ext = m & (ne > 0)
    assert np.array_equal(v[ext], val_o.ravel()[ext])


@pytest.mark.parametrize("form", [FE64, FE32, SC64])
@pytest.mark.parametrize("name", FRAMES)
@pytest.mark.parametrize("ede", [False, True])
def test_lean_steps_change_nothing(form, name, ede):
    """Full step + lean steps (the kernel's loop) against full steps alone: every escape
    count, every value bit and every step counter equal, for the interior-detection
    variants and several lean-run lengths."""
    fv = prepared(name)
    for interior in (True, False):
        ref = hostk.run(fv, form, ede=ede, interior=interior, lean=False)
        for chunk in (1, 7, 32):
            got = hostk.run(fv, form, ede=ede, interior=interior, lean=True, chunk=chunk)
            assert np.array_equal(got[0], ref[0]), (interior, chunk)
            assert np.array_equal(got[1].view(np.uint64), ref[1].view(np.uint64)), (interior, chunk)
            assert np.array_equal(got[2], ref[2]), (interior, chunk)


@pytest.mark.parametrize("name", FRAMES + ["minibrot875_ede", "subpixel737_ede", "exterior1000_ede"])
def test_scaled_algorithm_in_fp64_meets_the_bar(name):
    """The mixed form's algorithm (power-of-two scaled delta, FP64 radii and rebase
    test, fused multiply-adds) with an FP64 delta reproduces the plain form's escape
    counts on >= 99.9 % of samples, values within 1e-6 relative: its decisions are
    sound, so what separates the FP32 form from the bar is the FP32 delta alone.
    Exception: the distance estimates around the resolved minibrot (periodic
    reference) move by up to ~17 % between two FP64 roundings of the same algorithm
    (the derivative is carried through repeated near-zero returns), so there only
    the counts are compared."""
    cm = cases.FRAMES[name][6]
    fv = prepared(name)
    a = hostk.run(fv, FE64, ede=cm == 2)
    b = hostk.run(fv, SC64, ede=cm == 2)
    assert identical(a[0], b[0]) >= 0.999
    if name == "minibrot875_ede":
        return
    ok = (a[0] == b[0]) & (a[0] > 0)
    assert np.max(np.abs(b[1][ok] - a[1][ok]) / np.abs(a[1][ok])) <= 1e-6


def _dc_rounded(fv, bits):
    g = dict(fv)
    q = 2.0 ** bits
    g["dcol_m"] = np.round(fv["dcol_m"] * q) / q
    g["drow_m"] = np.round(fv["drow_m"] * q) / q
    return g


@pytest.mark.parametrize("name", ["minibrot875", "extref875"])
def test_counts_are_ill_conditioned_at_fp32_resolution(name):
    """Why no FP32-mantissa delta can meet >= 99.9 % identical counts on these frames:
    rounding each sample's dc (its offset from the reference) to 24 significant bits,
    a relative change below FP32's resolution of that same offset, and then computing
    everything in the plain FP64 form already changes more than 0.1 % of the escape
    counts.  At 40 bits nothing changes.  The FP32 form rounds its delta to 24 bits at
    every step, and its disagreement is larger still."""
    fv = prepared(name)
    base = hostk.run(fv, FE64)[0]
    assert identical(hostk.run(_dc_rounded(fv, 24), FE64)[0], base) < 0.999
    assert identical(hostk.run(_dc_rounded(fv, 40), FE64)[0], base) == 1.0
    f32 = identical(hostk.run(fv, FE32)[0], base)
    assert f32 <= identical(hostk.run(_dc_rounded(fv, 24), FE64)[0], base)
    assert f32 >= 0.9  # measured: 0.942 (minibrot875), 0.949 (extref875)


@pytest.mark.xfail(strict=True, reason="the workload's FP32 target (>= 99.9 % identical escape counts) is out of "
                   "reach of an FP32 delta on views near a minibrot: see "
                   "test_counts_are_ill_conditioned_at_fp32_resolution")
@pytest.mark.parametrize("name", ["minibrot875", "extref875"])
def test_fp32_form_target(name):
    fv = prepared(name)
    assert identical(hostk.run(fv, FE32)[0], hostk.run(fv, FE64)[0]) >= 0.999
