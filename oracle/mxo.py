"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  ctypes bindings for oracle/libmxo.so, the CPU restatement of the
reference engine's perturbation path (src/engine/mandel_perturbation.cpp)."""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_NAME = "libmxo.so"


def use_build(name):
    """Load another build of the oracle (libmxo_clang.so: the clang + AVX2 + libomp
    build that times the CPU baseline) from the next call on."""
    global _LIB, _NAME
    _LIB, _NAME = None, name


class Frame(C.Structure):
    _fields_ = [
        ("cre", C.c_char_p), ("cim", C.c_char_p), ("scale", C.c_char_p),
        ("W", C.c_int), ("H", C.c_int), ("full_h", C.c_int), ("row_base", C.c_int),
        ("mxit", C.c_int), ("c_method", C.c_int), ("escape_radius", C.c_float),
        ("prec_bits", C.c_int), ("use_bla", C.c_int), ("bla_minskip", C.c_int),
        ("bla_eps", C.c_double), ("interior", C.c_int), ("probe", C.c_int), ("threads", C.c_int),
        ("val", C.POINTER(C.c_double)), ("n_esc", C.POINTER(C.c_int32)), ("out", C.POINTER(C.c_float)),
        ("ref_it", C.c_int), ("ref_escaped", C.c_int), ("ref_virtual", C.c_int),
        ("ref_x", C.c_double), ("ref_y", C.c_double), ("interior_used", C.c_int),
        ("cutoff_sensitive_hint", C.c_int), ("seconds_pixels", C.c_double),
        ("steps", C.c_longlong), ("bla_applies", C.c_longlong), ("bla_skipped", C.c_longlong),
        ("rebases", C.c_longlong),
        ("orbit_zf", C.c_void_p), ("orbit_m", C.c_void_p), ("orbit_e", C.c_void_p),
        ("sub", C.c_int), ("cells", C.c_void_p), ("cell_val", C.c_void_p), ("cell_nesc", C.c_void_p),
        ("ss_pixels", C.c_int), ("no_periodic", C.c_int), ("period", C.c_int), ("subpixel", C.c_int),
        ("cutoff_sensitive", C.c_int), ("gmp_fallback", C.c_int), ("view_center", C.c_int),
        ("rescales", C.c_longlong), ("underflow_steps", C.c_longlong), ("deriv_underflows", C.c_longlong),
        ("force_fe", C.c_int), ("force_periodic", C.c_int),
    ]


class PixelsIn(C.Structure):
    _fields_ = [
        ("zf", C.c_void_p), ("zre_m", C.c_void_p), ("zre_e", C.c_void_p),
        ("zim_m", C.c_void_p), ("zim_e", C.c_void_p), ("mx_ref_it", C.c_int),
        ("nlev", C.c_int), ("lev_off", C.c_void_p), ("lev_size", C.c_void_p),
        ("r2", C.c_void_p), ("l", C.c_void_p), ("coef", C.c_void_p), ("rmax2", C.c_double),
        ("bla_minlevel", C.c_int), ("esc2", C.c_double), ("use_interior", C.c_int),
        ("ede", C.c_int), ("dx_m", C.c_double), ("dx_e", C.c_int64), ("mxit", C.c_int),
        ("threads", C.c_int), ("period", C.c_int),
    ]


class DFrame(C.Structure):
    _fields_ = [
        ("cre", C.c_char_p), ("cim", C.c_char_p), ("scale", C.c_char_p),
        ("W", C.c_int), ("H", C.c_int), ("mxit", C.c_int), ("escape_radius", C.c_float),
        ("prec_bits", C.c_int), ("use_bla", C.c_int), ("bla_minskip", C.c_int), ("interior", C.c_int),
        ("threads", C.c_int),
        ("val", C.POINTER(C.c_double)), ("n_esc", C.POINTER(C.c_int32)), ("out", C.POINTER(C.c_float)),
        ("passes", C.c_int), ("escalate", C.c_int), ("sa_it0", C.c_int), ("sa_order0", C.c_int),
        ("steps", C.c_longlong), ("bla_applies", C.c_longlong), ("bla_skipped", C.c_longlong),
        ("rebases", C.c_longlong), ("ede", C.c_int),
        ("sub", C.c_int), ("ss", C.c_int), ("cells", C.POINTER(C.c_float)), ("cell_val", C.POINTER(C.c_double)),
        ("cell_nesc", C.POINTER(C.c_int32)), ("ss_pixels", C.c_int), ("ss_passes", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _NAME)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle/{_NAME} is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.mxo_compute_fe.argtypes = [C.POINTER(Frame)]
        L.mxo_compute_dbl.argtypes = [C.POINTER(DFrame)]
        L.mxo_sa_coeffs.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.POINTER(C.c_int)]
        L.mxo_sa_coeffs.restype = C.c_int
        L.mxo_sa_seed.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.mxo_cdiv.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.mxo_pixels_dbl.argtypes = [C.POINTER(PixelsIn)] + [C.c_void_p] * 6 + [C.c_int, C.c_int, C.c_double, C.c_double,
                                                                                C.c_double, C.c_int] + [C.c_void_p] * 6
        L.mxo_dbl_setup.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int] + \
            [C.c_void_p] * 4 + [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_int] + [C.c_void_p] * 3
        L.mxo_dbl_setup.restype = C.c_int
        L.mxo_fe_op.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6
        L.mxo_orbit_shadows.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int] + [C.c_void_p] * 7
        L.mxo_orbit_shadows.restype = C.c_int
        L.mxo_ss_flags.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.mxo_ss_flags.restype = C.c_int
        L.mxo_fixed_mul.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int)]
        L.mxo_fixed_to_fe.argtypes = [C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.mxo_bla_total.argtypes = [C.c_int]
        L.mxo_bla_total.restype = C.c_int64
        L.mxo_bla_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int] + [C.c_void_p] * 7
        L.mxo_bla_tables.restype = C.c_int
        L.mxo_pixels_fe.argtypes = [C.POINTER(PixelsIn), C.c_int] + [C.c_void_p] * 7
        L.mxo_direct.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.mxo_direct_skip_bulbs.argtypes = L.mxo_direct.argtypes
        L.mxo_direct_list.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_float, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.mxo_point.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_char_p, C.c_void_p]
        L.mxo_step_underflow.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int64, C.c_double,
                                         C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.mxo_deriv_underflow.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int64, C.c_double, C.c_double,
                                          C.c_void_p, C.c_void_p, C.c_void_p]
        L.mxo_shallow_points.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                         C.c_void_p]
        L.mxo_shallow.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_float, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p]
        L.mxo_point.restype = C.c_double
        L.mxo_grid_info.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p] + [C.c_int] * 5 + [C.c_void_p] * 4
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def precision_bits(scale_str):
    """R.cpp:170 / verify.cpp:157: digits(scale) * log2(10) + 64."""
    return int(len(scale_str) * math.log(10) / math.log(2)) + 64


def compute_fe(cre, cim, scale, W, H, mxit, c_method=0, escape_radius=1e8, prec_bits=None,
               use_bla=True, bla_minskip=8, interior=True, probe=True, threads=None, full_h=0, row_base=0,
               want_orbit=False, sub=1, periodic=True, view_center=False, force_fe=False, force_periodic=False):
    """Whole floatexp frame (Compute -> reference -> BLA -> probe -> pass).  With
    want_orbit, info["orbit"] holds the reference orbit's shadows (zf, zfe_m, zfe_e,
    (re, im) interleaved).  With sub > 1, info["cells"] holds the (H sub) x (W sub)
    float, value and escape-count cells (-10, 0, INT_MIN where nothing was computed),
    with the adaptive sub-samples when c_method has the super-sampling bit (1).
    periodic=False skips the search for a nucleus in view (full reference always);
    info["period"] is the nucleus period used as the reference (0: none) and
    info["subpixel"] whether a nucleus was found below 4 pixels (BLA then off)."""
    f = Frame()
    if sub > 1:
        nc = W * H * sub * sub
        cf = np.full(nc, -10.0, np.float32)
        cv = np.zeros(nc)
        cn = np.full(nc, np.iinfo(np.int32).min, np.int32)
        f.sub, f.cells, f.cell_val, f.cell_nesc = sub, _p(cf), _p(cv), _p(cn)
    if want_orbit:
        ozf = np.zeros(2 * (mxit + 1))
        om = np.zeros(2 * (mxit + 1))
        oe = np.zeros(2 * (mxit + 1), np.int64)
        f.orbit_zf, f.orbit_m, f.orbit_e = _p(ozf), _p(om), _p(oe)
    f.cre, f.cim, f.scale = cre.encode(), cim.encode(), scale.encode()
    f.W, f.H, f.full_h, f.row_base = W, H, full_h, row_base
    f.mxit, f.c_method, f.escape_radius = mxit, c_method, escape_radius
    f.prec_bits = prec_bits or precision_bits(scale)
    f.use_bla, f.bla_minskip, f.bla_eps = int(use_bla), bla_minskip, 0.0
    f.interior, f.probe = int(interior), int(probe)
    f.no_periodic = int(not periodic)
    f.view_center = int(view_center)
    f.force_fe, f.force_periodic = int(force_fe), int(force_periodic)
    f.threads = threads or os.cpu_count()
    val = np.zeros(W * H, np.float64)
    nesc = np.zeros(W * H, np.int32)
    out = np.zeros(W * H, np.float32)
    f.val = val.ctypes.data_as(C.POINTER(C.c_double))
    f.n_esc = nesc.ctypes.data_as(C.POINTER(C.c_int32))
    f.out = out.ctypes.data_as(C.POINTER(C.c_float))
    lib().mxo_compute_fe(C.byref(f))
    info = {k: getattr(f, k) for k in ("ref_it", "ref_escaped", "ref_virtual", "ref_x", "ref_y", "interior_used",
                                        "cutoff_sensitive_hint", "seconds_pixels", "steps", "bla_applies",
                                        "bla_skipped", "rebases", "period", "subpixel", "cutoff_sensitive",
                                        "gmp_fallback", "rescales", "underflow_steps", "deriv_underflows")}
    if want_orbit:
        n = 2 * (f.period if f.period else f.ref_it + 1)
        info["orbit"] = dict(zf=ozf[:n].copy(), zfe_m=om[:n].copy(), zfe_e=oe[:n].copy())
    if sub > 1:
        info["cells"] = (cf.reshape(H * sub, W * sub), cv.reshape(H * sub, W * sub), cn.reshape(H * sub, W * sub))
        info["ss_pixels"] = f.ss_pixels
    return val.reshape(H, W), nesc.reshape(H, W), out.reshape(H, W), info


def compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=1e8, prec_bits=None, use_bla=True, bla_minskip=8,
                interior=True, threads=None, c_method=0, sub=1):
    """Whole plain-double frame (Compute's double path: centre-pixel reference with its
    series approximation, BLA, probe gate, passes re-referenced on glitches).  With
    sub > 1 the info dict carries the (H sub) x (W sub) cells ("cells", "cell_val",
    "cell_nesc") and, when c_method has the super-sampling bit (1), the adaptive
    sub-samples' passes ("ss_pixels", "ss_passes")."""
    f = DFrame()
    f.cre, f.cim, f.scale = cre.encode(), cim.encode(), scale.encode()
    f.W, f.H, f.mxit, f.escape_radius = W, H, mxit, escape_radius
    f.prec_bits = prec_bits or precision_bits(scale)
    f.use_bla, f.bla_minskip, f.interior = int(use_bla), bla_minskip, int(interior)
    f.threads = threads or os.cpu_count()
    f.ede = int((c_method & 2) != 0)
    val = np.zeros(W * H, np.float64)
    nesc = np.zeros(W * H, np.int32)
    out = np.zeros(W * H, np.float32)
    f.val = val.ctypes.data_as(C.POINTER(C.c_double))
    f.n_esc = nesc.ctypes.data_as(C.POINTER(C.c_int32))
    f.out = out.ctypes.data_as(C.POINTER(C.c_float))
    f.sub, f.ss = sub, int(sub > 1 and (c_method & 1) != 0)
    if sub > 1:
        ncell = W * H * sub * sub
        cells = np.full(ncell, -10.0, np.float32)
        cell_val = np.full(ncell, -10.0, np.float64)
        cell_nesc = np.full(ncell, np.iinfo(np.int32).min, np.int32)
        f.cells = cells.ctypes.data_as(C.POINTER(C.c_float))
        f.cell_val = cell_val.ctypes.data_as(C.POINTER(C.c_double))
        f.cell_nesc = cell_nesc.ctypes.data_as(C.POINTER(C.c_int32))
    lib().mxo_compute_dbl(C.byref(f))
    info = {k: getattr(f, k) for k in ("passes", "escalate", "sa_it0", "sa_order0", "steps", "bla_applies",
                                        "bla_skipped", "rebases", "ss_pixels", "ss_passes")}
    if sub > 1:
        info.update(cells=cells.reshape(H * sub, W * sub), cell_val=cell_val.reshape(H * sub, W * sub),
                    cell_nesc=cell_nesc.reshape(H * sub, W * sub))
    return val.reshape(H, W), nesc.reshape(H, W), out.reshape(H, W), info


def sa_coeffs(cre, cim, prec_bits, mxit, delta, escape_radius=1e8):
    """Series coefficients A_0..A_29 (complex) of the reference at (cre, cim), the last
    accepted iteration and the truncation order (calCoefficient / SACheckMagnitude)."""
    A = np.zeros(60, np.float64)
    order = C.c_int(0)
    it = lib().mxo_sa_coeffs(cre.encode(), cim.encode(), prec_bits, mxit, escape_radius, delta.real, delta.imag,
                             _p(A), C.byref(order))
    return A[0::2] + 1j * A[1::2], it, order.value


def sa_seed(A, order, delta, dc):
    """Horner evaluation of the series at pixel offsets dc (complex array)."""
    Ai = np.zeros(60, np.float64)
    Ai[0::2], Ai[1::2] = np.real(A), np.imag(A)
    dc = np.asarray(dc, np.complex128)
    dr, di = np.ascontiguousarray(dc.real), np.ascontiguousarray(dc.imag)
    out = np.zeros(2 * dc.size, np.float64)
    lib().mxo_sa_seed(_p(Ai), order, delta.real, delta.imag, dc.size, _p(dr), _p(di), _p(out))
    return out[0::2] + 1j * out[1::2]


def cdiv(a, b):
    a = np.ascontiguousarray(np.asarray(a, np.complex128))
    b = np.ascontiguousarray(np.asarray(b, np.complex128))
    out = np.zeros(a.size, np.complex128)
    lib().mxo_cdiv(a.size, _p(a), _p(b), _p(out))
    return out


def compute(cre, cim, scale, W, H, mxit, c_method=0, escape_radius=1e8, prec_bits=None, use_bla=True, bla_minskip=8,
            interior=True, threads=None):
    """Mandel::Compute's path selection (mandel_perturbation.cpp:1434, 1569-1576, 1336-1355):
    the plain-double path at 1e6 < scale <= 1e280 unless its reference escalates the frame
    to floatexp, the floatexp path above.  Returns (val, n_esc, out, info) with
    info["path"] 1 (double) or 2 (floatexp) and info["escalated"]."""
    from decimal import Decimal
    kw = dict(escape_radius=escape_radius, prec_bits=prec_bits, use_bla=use_bla, bla_minskip=bla_minskip,
              interior=interior, threads=threads)
    if Decimal(scale) <= Decimal("1e280"):
        val, nesc, out, info = compute_dbl(cre, cim, scale, W, H, mxit, c_method=c_method, **kw)
        if not info["escalate"]:
            info.update(path=1, escalated=0)
            return val, nesc, out, info
    val, nesc, out, info = compute_fe(cre, cim, scale, W, H, mxit, c_method=c_method, **kw)
    info.update(path=2, escalated=int(Decimal(scale) <= Decimal("1e280")))
    return val, nesc, out, info


def fe_op(op, am, ae, bm=None, be=None):
    am = np.ascontiguousarray(am, np.float64)
    ae = np.ascontiguousarray(ae, np.int64)
    n = am.size
    bm = np.ascontiguousarray(np.zeros(n) if bm is None else bm, np.float64)
    be = np.ascontiguousarray(np.zeros(n, np.int64) if be is None else be, np.int64)
    rm = np.zeros(n)
    re = np.zeros(n, np.int64)
    lib().mxo_fe_op(op, n, _p(am), _p(ae), _p(bm), _p(be), _p(rm), _p(re))
    return rm, re


def orbit(cre, cim, prec_bits, mxit, escape_radius=1e8, deriv=False, fx_limbs=0):
    """Reference orbit shadows: MPFR at prec_bits, or with fx_limbs > 0 the floatexp
    path's fixed-point orbit (fx_limbs 64-bit limbs) from the point held at prec_bits."""
    zf = np.zeros(2 * (mxit + 1))
    zrm = np.zeros(mxit + 1)
    zre = np.zeros(mxit + 1, np.int64)
    zim = np.zeros(mxit + 1)
    zie = np.zeros(mxit + 1, np.int64)
    dfr = np.zeros(mxit + 1)
    dfi = np.zeros(mxit + 1)
    r = lib().mxo_orbit_shadows(cre.encode(), cim.encode(), prec_bits, int(fx_limbs), mxit, float(np.float32(escape_radius)), int(deriv),
                                _p(zf), _p(zrm), _p(zre), _p(zim), _p(zie), _p(dfr), _p(dfi))
    n = r + 1
    return dict(ref_it=r, zf=zf[:2 * n].copy(), zre_m=zrm[:n].copy(), zre_e=zre[:n].copy(),
                zim_m=zim[:n].copy(), zim_e=zie[:n].copy(), dfr=dfr[:n].copy(), dfi=dfi[:n].copy())


def _limbs(v, L):
    a = np.zeros(L, np.uint64)
    for i in range(L):
        a[i] = (abs(v) >> (64 * i)) & 0xFFFFFFFFFFFFFFFF
    return a


def fixed_mul(L, x, y):
    """Fixed-point product of two signed integers standing for x * 2^(-64(L-1)),
    as the floatexp path's orbit computes it; returns the signed integer result."""
    a, b, r = _limbs(x, L), _limbs(y, L), np.zeros(L, np.uint64)
    sr = C.c_int(0)
    sgn = lambda v: (v > 0) - (v < 0)
    lib().mxo_fixed_mul(L, _p(a), sgn(x), _p(b), sgn(y), _p(r), C.byref(sr))
    return sr.value * sum(int(r[i]) << (64 * i) for i in range(L))


def fixed_to_fe(L, x):
    """FloatExp shadow (mantissa, exponent) of the fixed-point value x * 2^(-64(L-1))."""
    a = _limbs(x, L)
    m, e = C.c_double(0), C.c_int64(0)
    lib().mxo_fixed_to_fe(L, _p(a), (x > 0) - (x < 0), C.byref(m), C.byref(e))
    return m.value, e.value


def ss_flags(out, mxit, ede):
    """Adaptive supersampling's refined pixels (bool array) from base float values."""
    out = np.ascontiguousarray(out, np.float32)
    H, W = out.shape
    fl = np.zeros(W * H, np.uint8)
    lib().mxo_ss_flags(_p(out), W, H, mxit, int(ede), _p(fl))
    return fl.reshape(H, W).astype(bool)


def bla_tables(zf, reflen, dcmax, eps=2.0 ** -53, ede=False, dfr=None, dfi=None):
    tot = lib().mxo_bla_total(reflen)
    if tot == 0:
        return None
    r2 = np.zeros(tot)
    l = np.zeros(tot, np.int64)
    coef = np.zeros(4 * tot)
    der = np.zeros(4 * tot)
    rmax2 = C.c_double(0)
    off = np.zeros(64, np.int64)
    size = np.zeros(64, np.int64)
    zf = np.ascontiguousarray(zf, np.float64)
    dfr = np.ascontiguousarray(np.zeros(reflen) if dfr is None else dfr, np.float64)
    dfi = np.ascontiguousarray(np.zeros(reflen) if dfi is None else dfi, np.float64)
    nl = lib().mxo_bla_tables(_p(zf), _p(dfr), _p(dfi), reflen, dcmax, eps, int(ede), _p(r2), _p(l), _p(coef), _p(der),
                              C.cast(C.byref(rmax2), C.c_void_p), _p(off), _p(size))
    return dict(nlev=nl, r2=r2, l=l, coef=coef, der=der, rmax2=rmax2.value, off=off[:nl].copy(), size=size[:nl].copy())


def pixels_fe(orb, bla, dcr_m, dcr_e, dci_m, dci_e, mxit, esc2, dx_m, dx_e, ede=False, use_interior=True,
              bla_minlevel=3, threads=None, period=0):
    """Per-pixel rescaled loop over caller-supplied orbit/BLA arrays (period > 0: the
    orbit is one period of a nucleus, read modulo the period)."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return _p(a)

    pi = PixelsIn()
    pi.zf = arr(orb["zf"], np.float64)
    pi.zre_m, pi.zre_e = arr(orb["zre_m"], np.float64), arr(orb["zre_e"], np.int64)
    pi.zim_m, pi.zim_e = arr(orb["zim_m"], np.float64), arr(orb["zim_e"], np.int64)
    pi.mx_ref_it = period if period else len(orb["zre_m"]) - 1
    pi.period = period
    if bla is None:
        pi.nlev = 0
    else:
        pi.nlev = bla["nlev"]
        pi.lev_off, pi.lev_size = arr(bla["off"], np.int64), arr(bla["size"], np.int64)
        pi.r2, pi.l, pi.coef = arr(bla["r2"], np.float64), arr(bla["l"], np.int64), arr(bla["coef"], np.float64)
        pi.rmax2 = bla["rmax2"]
    pi.bla_minlevel, pi.esc2, pi.use_interior, pi.ede = bla_minlevel, esc2, int(use_interior), int(ede)
    pi.dx_m, pi.dx_e, pi.mxit = dx_m, dx_e, mxit
    pi.threads = threads or os.cpu_count()
    n = len(dcr_m)
    val = np.zeros(n)
    nesc = np.zeros(n, np.int32)
    st = np.zeros(7, np.int64)
    lib().mxo_pixels_fe(C.byref(pi), n, arr(dcr_m, np.float64), arr(dcr_e, np.int64), arr(dci_m, np.float64),
                        arr(dci_e, np.int64), _p(val), _p(nesc), _p(st))
    return val, nesc, dict(steps=int(st[0]), bla_applies=int(st[1]), bla_skipped=int(st[2]), rebases=int(st[3]),
                           rescales=int(st[4]), underflow_steps=int(st[5]), deriv_underflows=int(st[6]))


def pixels_dbl(zf, z_m3, bla, sa_coef, sa_it, sa_order, sa_delta, dcr, dci, mxit, esc2, use_interior=True,
               bla_minlevel=3, threads=None, ede=False, dfr=None, dfi=None, sa_coef_b=None, dxf=0.0):
    """The double loop (series seed included) over caller-supplied arrays."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return _p(a)

    pi = PixelsIn()
    pi.zf = arr(zf, np.float64)
    pi.mx_ref_it = len(z_m3) - 1
    if bla is None:
        pi.nlev = 0
    else:
        pi.nlev = bla["nlev"]
        pi.lev_off, pi.lev_size = arr(bla["off"], np.int64), arr(bla["size"], np.int64)
        pi.r2, pi.l, pi.coef = arr(bla["r2"], np.float64), arr(bla["l"], np.int64), arr(bla["coef"], np.float64)
        pi.rmax2 = bla["rmax2"]
    pi.bla_minlevel, pi.esc2, pi.use_interior, pi.ede = bla_minlevel, esc2, int(use_interior), int(ede)
    pi.mxit = mxit
    pi.threads = threads or os.cpu_count()
    n = len(dcr)
    val = np.zeros(n)
    nesc = np.zeros(n, np.int32)
    gp = np.zeros(n)
    st = np.zeros(4, np.int64)
    opt = lambda a: arr(a, np.float64) if a is not None else None
    der = bla.get("der") if (bla is not None and ede) else None
    lib().mxo_pixels_dbl(C.byref(pi), arr(z_m3, np.float64), opt(dfr), opt(dfi), opt(der), arr(sa_coef, np.float64),
                         opt(sa_coef_b), sa_it, sa_order, sa_delta.real, sa_delta.imag, dxf, n, arr(dcr, np.float64),
                         arr(dci, np.float64), _p(val), _p(nesc), _p(gp), _p(st))
    return val, nesc, gp, dict(steps=int(st[0]), bla_applies=int(st[1]), bla_skipped=int(st[2]), rebases=int(st[3]))


def dbl_setup(cre, cim, scale, W, H, mxit, escape_radius=1e8, prec_bits=None, ede=False):
    """First-reference set-up of a double frame: orbit, |X|^2/1e6, series, dc tables."""
    pb = prec_bits or precision_bits(scale)
    zf = np.zeros(2 * (mxit + 1))
    m3 = np.zeros(mxit + 1)
    A = np.zeros(60)
    delta = np.zeros(2)
    it, order = C.c_int(0), C.c_int(0)
    dcol = np.zeros(W)
    drow = np.zeros(H)
    dfr = np.zeros(mxit + 1)
    dfi = np.zeros(mxit + 1)
    B = np.zeros(60)
    ri = lib().mxo_dbl_setup(cre.encode(), cim.encode(), scale.encode(), W, H, mxit, escape_radius, pb, _p(zf), _p(m3),
                             _p(A), _p(delta), C.byref(it), C.byref(order), _p(dcol), _p(drow), int(ede), _p(dfr),
                             _p(dfi), _p(B))
    d = dict(ref_it=ri, zf=zf[:2 * (ri + 1)], z_m3=m3[:ri + 1], sa_coef=A, sa_delta=complex(delta[0], delta[1]),
             sa_it=it.value, sa_order=order.value, dcol_d=dcol, drow_d=drow)
    if ede:
        d.update(dfr=dfr[:ri + 1], dfi=dfi[:ri + 1], sa_coef_b=B)
    return d


def direct(cre, cim, scale, W, H, mxit, c_method=0, escape_radius=1e8, step=1, prec_bits=None, wp_bits=None,
           threads=None, track_d=True, skip_bulbs=False):
    """Brute-force MPFR frame (ComputeDirect); unsampled pixels are NaN / 0.  skip_bulbs:
    pixels inside the main cardioid or the period-2 disc are marked interior (-2, -2)
    without iterating; every escaping pixel is unchanged."""
    pb = prec_bits or precision_bits(scale)
    val = np.full(W * H, np.nan)
    nesc = np.zeros(W * H, np.int32)
    (lib().mxo_direct_skip_bulbs if skip_bulbs else lib().mxo_direct)(cre.encode(), cim.encode(), scale.encode(), W, H, pb, wp_bits or pb, mxit, c_method,
                     escape_radius, step, threads or os.cpu_count(), int(track_d), _p(val), _p(nesc))
    return val.reshape(H, W), nesc.reshape(H, W)


def direct_list(cre, cim, scale, W, H, mxit, idx, c_method=0, escape_radius=1e8, prec_bits=None, wp_bits=None,
                threads=None, track_d=True):
    """Brute-force MPFR samples idx (row-major) of the frame's grid -> (val, n_esc)."""
    pb = prec_bits or precision_bits(scale)
    idx = np.ascontiguousarray(idx, np.int32)
    val = np.zeros(len(idx))
    nesc = np.zeros(len(idx), np.int32)
    lib().mxo_direct_list(cre.encode(), cim.encode(), scale.encode(), W, H, pb, wp_bits or pb, mxit, c_method,
                          escape_radius, _p(idx), len(idx), threads or os.cpu_count(), int(track_d), _p(val), _p(nesc))
    return val, nesc


def step_underflow(X, dc, S, w):
    """feRescaledStepUnderflow on explicit FloatExp inputs: X, dc = ((m, e), (m, e)),
    S = (m, e), w = (wr, wi) -> (S', (wr', wi'))."""
    xm = np.array([X[0][0], X[1][0]]); xe = np.array([X[0][1], X[1][1]], np.int64)
    cm = np.array([dc[0][0], dc[1][0]]); ce = np.array([dc[0][1], dc[1][1]], np.int64)
    om, oe, ow = C.c_double(), C.c_int64(), np.zeros(2)
    lib().mxo_step_underflow(_p(xm), _p(xe), _p(cm), _p(ce), S[0], S[1], w[0], w[1], C.cast(C.byref(om), C.c_void_p),
                             C.cast(C.byref(oe), C.c_void_p), _p(ow))
    return (om.value, oe.value), (ow[0], ow[1])


def deriv_underflow(z, SJ, j):
    """feDerivativeStepUnderflow on explicit inputs: z = ((m, e), (m, e)), SJ = (m, e),
    j = (jr, ji) -> (SJ', (jr', ji'), 1/SJ' as double)."""
    zm = np.array([z[0][0], z[1][0]]); ze = np.array([z[0][1], z[1][1]], np.int64)
    om, oe, oj = C.c_double(), C.c_int64(), np.zeros(3)
    lib().mxo_deriv_underflow(_p(zm), _p(ze), SJ[0], SJ[1], j[0], j[1], C.cast(C.byref(om), C.c_void_p),
                              C.cast(C.byref(oe), C.c_void_p), _p(oj))
    return (om.value, oe.value), (oj[0], oj[1]), oj[2]


def shallow_points(cr, ci, mxit, escape_radius=1e8, ede=False, scalar=False):
    """The shallow path's per-sample loop on explicit c (double): floatPointCompute
    (scalar) or one lane of solveShallowSimdList -> (unrounded value, n_esc)."""
    cr = np.ascontiguousarray(cr, np.float64)
    ci = np.ascontiguousarray(ci, np.float64)
    val = np.zeros(len(cr))
    nesc = np.zeros(len(cr), np.int32)
    esc2 = float(np.float32(escape_radius) * np.float32(escape_radius))
    lib().mxo_shallow_points(_p(cr), _p(ci), len(cr), mxit, esc2, int(ede), int(scalar), _p(val), _p(nesc))
    return val, nesc


def shallow(cre, cim, scale, W, H, mxit, ede=False, sub=1, flagged=None, escape_radius=1e8, full_h=0, row_base=0,
            prec_bits=None, threads=None):
    """A shallow view (scale <= 1e6) as the engine's method 0 computes it -> (val, n_esc,
    (c0r, c0i, dx, dy)); val's float is the engine's output.  With `flagged` (row-major
    pixel indices) the adaptive pass: sub*sub - 1 cells per pixel, xi outer, yi inner."""
    lst = None if flagged is None else np.ascontiguousarray(flagged, np.int32)
    n = W * H if lst is None else len(lst) * (sub * sub - 1)
    val = np.zeros(n)
    nesc = np.zeros(n, np.int32)
    g4 = np.zeros(4)
    lib().mxo_shallow(cre.encode(), cim.encode(), scale.encode(), W, H, full_h, row_base, prec_bits or precision_bits(scale),
                      mxit, escape_radius, int(ede), sub, None if lst is None else _p(lst), 0 if lst is None else len(lst),
                      threads or os.cpu_count(), _p(g4), _p(val), _p(nesc))
    return val, nesc, tuple(g4)


def point(cre, cim, wp_bits, mxit, esc2=float(np.float32(1e8) * np.float32(1e8)), ede=False, track_d=True, dx="1"):
    n = C.c_int(0)
    v = lib().mxo_point(cre.encode(), cim.encode(), wp_bits, mxit, int(ede), int(track_d), esc2, dx.encode(),
                        C.cast(C.byref(n), C.c_void_p))
    return v, n.value


def nucleus(cre, cim, scale, W, H, maxp, prec_bits=None):
    """The oracle's nucleus search for a view: (period, nucleus as a decimal string pair,
    island size as a float log10), period 0 when there is none."""
    nd = 4000
    rb, ib = C.create_string_buffer(nd), C.create_string_buffer(nd)
    re_, ie = C.c_long(), C.c_long()
    sm, se = C.c_double(), C.c_int64()
    p = lib().mxo_nucleus(cre.encode(), cim.encode(), scale.encode(), W, H, maxp, prec_bits or precision_bits(scale),
                          rb, C.byref(re_), ib, C.byref(ie), nd, C.byref(sm), C.byref(se))
    if p <= 0:
        return 0, None, None

    def dec(buf, e):
        d = buf.value.decode()
        neg = d.startswith("-")
        d = d.lstrip("-")
        return ("-" if neg else "") + "0." + d + f"e{e.value}"

    import math
    return p, (dec(rb, re_), dec(ib, ie)), math.log10(abs(sm.value)) + se.value * math.log10(2.0)


def grid_info(cre, cim, scale, W, H, prec_bits=None, full_h=0, row_base=0):
    dxm, dym = C.c_double(), C.c_double()
    dxe, dye = C.c_int64(), C.c_int64()
    lib().mxo_grid_info(cre.encode(), cim.encode(), scale.encode(), W, H, full_h, row_base,
                        prec_bits or precision_bits(scale), C.cast(C.byref(dxm), C.c_void_p),
                        C.cast(C.byref(dxe), C.c_void_p), C.cast(C.byref(dym), C.c_void_p),
                        C.cast(C.byref(dye), C.c_void_p))
    return (dxm.value, dxe.value), (dym.value, dye.value)


def palette():
    """The oracle's restatement of the engine's palette, (3, 2048) float32."""
    pal = np.zeros(3 * 2048, np.float32)
    lib().mxo_palette(_p(pal))
    return pal.reshape(3, 2048)


def colorize(it, ss=1, ede=False, density=60.0):
    """render.cpp's colour pass on the CPU: (H ss, W ss) floats -> (H, W, 3) bytes."""
    it = np.ascontiguousarray(it, np.float32)
    Hs, Ws = it.shape
    W, H = Ws // ss, Hs // ss
    rgb = np.zeros((H, W, 3), np.uint8)
    lib().mxo_colorize(_p(it), W, H, ss, int(ede), C.c_float(density), _p(rgb))
    return rgb
