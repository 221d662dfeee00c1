"""The mixed FP32-floatexp form and the power-of-two-scaled FP64 variant against the
plain form (escape counts identical to the engine's) on the reference's test frames.

Accuracy bar of the workload for the chosen form: >= 99.9 % of samples at identical
escape count and smooth error <= 1e-6 (relative).  The FP64 variant meets it; the
FP32 form does not (test_fp32_target, expected to fail).  Why FP32 cannot
meet it next to them is shown on the host: rounding dc alone to 24 bits in the plain
form already moves more than 0.1 % of the counts
(tests/test_host_kernels.py::test_counts_are_ill_conditioned_at_fp32_resolution)."""
import numpy as np
import pytest

import cases
import mxhip

pytestmark = pytest.mark.gpu

FRAMES = ["extref875", "minibrot875", "exterior1000", "exterior1000_ede", "subpixel737_ede"]
# frames next to a resolved or sub-pixel minibrot, where FP32 misses the bar
NEAR_MINIBROT = {"extref875", "minibrot875", "subpixel737_ede"}


def cls(n):
    return np.where(n > 0, n, -1)  # max_iter and detected interior both give -2 in the engine buffer


def frames(ctx, name, forms):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES[name]
    print(f"[forms] {name} {W}x{H}", flush=True)
    return cm, [ctx.compute(cre, cim, scale, W, H, mxit, c_method=cm, form=f)[1:3] for f in forms]


def rel_err(v, v0, n, n0):
    ext = (n0 > 0) & (n == n0)
    return np.abs(v[ext] - v0[ext]) / np.abs(v0[ext])


@pytest.mark.parametrize("name", FRAMES)
def test_scaled_fp64_meets_the_bar(ctx, name):
    cm, [(v0, n0), (v3, n3)] = frames(ctx, name, [mxhip.MX_FORM_FE64, mxhip.MX_FORM_SC64])
    assert np.mean(cls(n0) == cls(n3)) >= 0.999
    e = rel_err(v3, v0, n3, n0)
    assert np.median(e) < 1e-9
    if not cm:
        assert e.max() <= 1e-6


@pytest.mark.parametrize("name", FRAMES)
def test_fp32_measured(ctx, name):
    """What the FP32 form achieves: every class kept, >= 90 % identical counts next to
    a minibrot (all of them elsewhere), median value error below 1e-4 relative."""
    cm, [(v0, n0), (v1, n1)] = frames(ctx, name, [mxhip.MX_FORM_FE64, mxhip.MX_FORM_FE32])
    assert np.array_equal(n0 > 0, n1 > 0)
    ident = np.mean(cls(n0) == cls(n1))
    assert ident >= (0.9 if name in NEAR_MINIBROT else 0.999)
    assert np.median(rel_err(v1, v0, n1, n0)) < 1e-4


@pytest.mark.xfail(strict=False, reason="FP32 delta: counts next to a minibrot are ill-conditioned at FP32 "
                   "resolution, and single samples elsewhere carry smooth errors of a few 1e-6")
@pytest.mark.parametrize("name", FRAMES)
def test_fp32_target(ctx, name):
    """The workload's bar for the FP32 form, as the assertion."""
    cm, [(v0, n0), (v1, n1)] = frames(ctx, name, [mxhip.MX_FORM_FE64, mxhip.MX_FORM_FE32])
    assert np.mean(cls(n0) == cls(n1)) >= 0.999
    if not cm:
        assert rel_err(v1, v0, n1, n0).max() <= 1e-6
