"""The oracle's plain-double path (scale 1e6..1e280: series-approximation seed, double
BLA, glitch detection with re-referencing passes) pinned against the reference's own
golden frames with the reference's tolerances (src/tools/verify.cpp:11860-11881),
against the MPFR brute force on the reference's double-path stress frames
(verify.cpp:11789-11800), and its series seed against an independent mpmath
iteration."""
import numpy as np
import pytest

import cases
import mxo

POINT31 = ("-0.749139567333446841955467474699747367338762518832278501811",
           "0.040823298514634751035521346975478853963578400940553676068", "7354177" + "0" * 25, 48, 32, 500000)
SLOWPOINT = ("-1.1758621450236620370", "-0.2447677973532398022", "782679600", 48, 32, 500000)
# 640x360 at scale 4e10, mxit 8192, bailout 65536, 128-bit precision
C1 = ("-0.743643887037151", "0.131825904205330", "40000000000", 640, 360, 8192)


def _stats(out, ref):
    """verify.cpp runCase: class mismatches, then |diff| over pixels exterior in both."""
    pi, di = out < -1.5, ref < -1.5
    assert not (out == -10).any(), "unresolved pixels left"
    mis = int((pi != di).sum())
    both = ~pi & ~di
    d = np.abs(out[both].astype(np.float64) - ref[both].astype(np.float64))
    if d.size == 0:
        return mis, 0.0, 0.0, 0.0
    q = int(np.ceil(0.99 * d.size)) - 1
    return mis, float(d.max()), float(d.mean()), float(np.partition(d, q)[q])


@pytest.mark.parametrize("frame,golden,tol", [
    (POINT31, "point_7e31.f32", (12000.0, 200.0, 4000.0)),      # verify.cpp:11876-11881
    (SLOWPOINT, "slowpoint_7e8.f32", (90000.0, 300.0, 3000.0)),  # verify.cpp:11864-11870
])
def test_double_path_matches_reference_golden(frame, golden, tol):
    cre, cim, scale, W, H, mxit = frame
    val, nesc, out, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit)
    g = cases.golden(golden, H, W)
    mis, mx, mean, p99 = _stats(out, g)
    assert mis == 0, f"{mis} class mismatches"
    assert mx <= tol[0] and mean <= tol[1] and p99 <= tol[2], (mx, mean, p99)
    assert info["escalate"] == 0


def test_point31_uses_glitch_passes():
    """point31 glitches against the centre-pixel reference: a second reference is chosen
    at the minimum glitch metric and resolves the rest (MP.cpp:1729-1761, 2556-2594)."""
    cre, cim, scale, W, H, mxit = POINT31
    _, nesc, out, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit)
    assert info["passes"] >= 2
    assert info["sa_it0"] > 0 and 5 <= info["sa_order0"] < 20


@pytest.mark.parametrize("name,scale_exp", [("ticktock", 141), ("flake", 157)])
def test_double_path_matches_brute_force(name, scale_exp):
    """The reference's glitch-stress frames at 120x80, brute force sampled every 12
    pixels, default tolerances (maxDiff 0.5, at most 1 % class mismatches)."""
    cre, cim = cases.LOC[name + "_x"], cases.LOC[name + "_y"]
    scale = cases.pow10(scale_exp)
    _, _, out, info = mxo.compute_dbl(cre, cim, scale, 120, 80, 200000)
    dv, _ = mxo.direct(cre, cim, scale, 120, 80, 200000, step=12)
    m = ~np.isnan(dv)
    mis, mx, _, _ = _stats(out[m], dv[m].astype(np.float32))
    assert mis <= 0.01 * m.sum() and mx <= 0.5
    assert info["bla_applies"] > 0


def test_c1_frame_matches_brute_force():
    cre, cim, scale, W, H, mxit = C1
    _, nesc, out, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=65536.0, prec_bits=128)
    dv, _ = mxo.direct(cre, cim, scale, W, H, mxit, escape_radius=65536.0, step=8, prec_bits=128, wp_bits=256)
    m = ~np.isnan(dv)
    mis, mx, mean, p99 = _stats(out[m], dv[m].astype(np.float32))
    assert mis == 0 and mean < 1.0
    assert info["passes"] == 1 and info["sa_it0"] > 500


def test_c1_distance_estimate_matches_brute_force():
    """Exterior distance estimate on the double path (derivative delta seeded by its
    own series, carried through BLA skips and rebases) against the MPFR brute force's
    pixel-normalised |z| ln|z| / (dx |dz/dc|)."""
    cre, cim, scale, W, H, mxit = C1
    val, nesc, out, info = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=65536.0, prec_bits=128, c_method=2)
    dv, dn = mxo.direct(cre, cim, scale, W, H, mxit, c_method=2, escape_radius=65536.0, step=8, prec_bits=128)
    m = ~np.isnan(dv)
    pi, di = nesc[m] < 0, dn[m] < 0
    assert (pi == di).all()
    both = ~pi & ~di
    rel = np.abs(val[m][both] - dv[m][both]) / np.abs(dv[m][both])
    assert np.median(rel) < 1e-7 and np.mean(rel < 1e-3) > 0.95
    # same classification and escape counts as the smooth run
    _, n0, _, _ = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=65536.0, prec_bits=128)
    assert np.array_equal(n0, nesc)


def test_bla_on_off_agree():
    """BLA changes the rounding path, not the classification (the reference's
    BLA-on/off comparison method, tests/perf_bench.cpp:569-595)."""
    cre, cim = cases.LOC["flake_x"], cases.LOC["flake_y"]
    scale = cases.pow10(157)
    _, n1, o1, i1 = mxo.compute_dbl(cre, cim, scale, 60, 40, 200000, use_bla=True)
    _, n0, o0, i0 = mxo.compute_dbl(cre, cim, scale, 60, 40, 200000, use_bla=False)
    assert i1["bla_applies"] > 0 and i0["bla_applies"] == 0
    assert np.array_equal(o1 < -1.5, o0 < -1.5)
    assert np.abs(o1 - o0).max() < 0.5


def test_series_seed_matches_mpmath():
    """The truncated series evaluated at dc equals z_{n+1}(c+dc) - z_{n+1}(c) (n = the
    last accepted series iteration) from an independent 256-bit mpmath iteration."""
    mp = pytest.importorskip("mpmath")
    cre, cim = "-0.743643887037151", "0.131825904205330"
    delta = 1e-10 + 0.5625e-10j
    A, it, order = mxo.sa_coeffs(cre, cim, 128, 8192, delta, 65536.0)
    assert it > 100 and 5 <= order < 20
    rng = np.random.default_rng(3)
    dc = (rng.uniform(-1, 1, 12) * delta.real + 1j * rng.uniform(-1, 1, 12) * delta.imag)
    seed = mxo.sa_seed(A, order, delta, dc)
    mp.mp.prec = 256
    # the reference point as the oracle's 128-bit grid holds it
    c = mp.mpc(mp.mpf(cre), mp.mpf(cim))
    c = mp.mpc(mp.mpf(mp.mpf(c.real), prec=128), mp.mpf(mp.mpf(c.imag), prec=128))

    def z_n(cc, n):
        z = mp.mpc(0)
        for _ in range(n):
            z = z * z + cc
        return z

    zref = z_n(c, it + 1)
    for d, s in zip(dc, seed):
        dd = mp.mpc(float(d.real), float(d.imag))
        exact = complex(z_n(c + dd, it + 1) - zref)
        assert abs(s - exact) <= 1e-9 * abs(exact), (d, s, exact)


def test_complex_division():
    """The Smith division the series seed uses: within 3 ulp of the exact quotient."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(5)
    a = rng.standard_normal(200) + 1j * rng.standard_normal(200)
    b = (rng.standard_normal(200) + 1j * rng.standard_normal(200)) * 10.0 ** rng.integers(-200, 200, 200)
    q = mxo.cdiv(a, b)
    for x, y, r in zip(a, b, q):
        xr, xi, yr, yi = Fr(x.real), Fr(x.imag), Fr(y.real), Fr(y.imag)
        den = yr * yr + yi * yi
        er = (xr * yr + xi * yi) / den
        ei = (xi * yr - xr * yi) / den
        mag = max(abs(er), abs(ei))
        assert abs(Fr(r.real) - er) <= 3 * mag * Fr(2) ** -52
        assert abs(Fr(r.imag) - ei) <= 3 * mag * Fr(2) ** -52


# deep1's period-14164 nucleus at 1e200, odd size so that the centre pixel is the view
# centre, 3000-bit centre: the reference's |X| underflows double near iteration 14164
ESCALATE = (cases.LOC["deep1_x"], cases.LOC["deep1_y"], cases.pow10(200), 25, 17, 30000, 3000)


def test_near_zero_reference_escalates_to_floatexp():
    """A double-path frame (1e150 < scale <= 1e280) whose reference passes too close to
    zero runs on the floatexp path from the exact view centre (MP.cpp:1336-1355); the
    result matches the MPFR brute force."""
    cre, cim, scale, W, H, mxit, pb = ESCALATE
    val, nesc, out, info = mxo.compute(cre, cim, scale, W, H, mxit, prec_bits=pb)
    assert info["path"] == 2 and info["escalated"] == 1 and info["ref_virtual"] == 1
    assert (info["ref_x"], info["ref_y"]) == ((W - 1) / 2, (H - 1) / 2)
    dv, dn = mxo.direct(cre, cim, scale, W, H, mxit, step=4, prec_bits=pb)
    m = ~np.isnan(dv)
    assert ((nesc[m] < 0) == (dn[m] < 0)).all()
    ext = m & (nesc > 0)
    assert np.abs(val[ext] - dv[ext]).max() < 1e-6
