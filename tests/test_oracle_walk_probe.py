"""The boundary-walk probe's shortcut (tests/golden/make_walk.py): mxo.direct with
skip_bulbs marks the pixels inside the main cardioid or the period-2 disc interior
without iterating them.  The walk ranks escaping pixels only, so the shortcut must leave
every escaping pixel's count and value exactly as the full MPFR iteration gives them,
and claim no pixel that escapes -- checked on views that straddle both boundaries, the
cusp and the contact point of the two components."""
import numpy as np
import pytest

import mxo

VIEWS = [
    ("-0.75", "0", "1.3333333333", 2048),          # the whole set
    ("-0.75", "0.01", "200", 8192),                  # the contact point of the cardioid and the disc
    ("0.25", "0", "400", 8192),                      # the cusp
    ("0.266322544642", "-0.48330543154", "1365.3333333", 4096),  # the period-4 bulb's root
]


@pytest.mark.parametrize("view", VIEWS, ids=["whole", "contact", "cusp", "root4"])
@pytest.mark.parametrize("c_method", [0, 2])
def test_skip_bulbs_keeps_every_escaping_pixel(view, c_method):
    cx, cy, sc, mxit = view
    kw = dict(prec_bits=96, wp_bits=96, threads=4, c_method=c_method)
    v0, n0 = mxo.direct(cx, cy, sc, 48, 48, mxit, **kw)
    v1, n1 = mxo.direct(cx, cy, sc, 48, 48, mxit, skip_bulbs=True, **kw)
    esc = n0 > 0
    assert np.array_equal(esc, n1 > 0)
    assert np.array_equal(n0[esc], n1[esc]) and np.array_equal(v0[esc], v1[esc])
    # every pixel clearly inside either component (by 1e-9, tested here in doubles on
    # the engine's grid) is marked, so the shortcut is exercised
    W = H = 48
    dw = 2 / float(sc)
    d = 2 * dw / (W - 1)
    x = float(cx) - dw + d * np.arange(W)[None, :]
    y = float(cy) - dw + d * np.arange(H)[:, None]
    q = (x - 0.25) ** 2 + y * y
    inside = (q * (q + x - 0.25) < y * y / 4 - 1e-9) | ((x + 1) ** 2 + y * y < 0.0625 - 1e-9)
    assert inside.any() and (n1[inside] == -2).all() and not esc[inside].any()
