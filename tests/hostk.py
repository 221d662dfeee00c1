"""Test infrastructure: ctypes bindings for _build/libmxhost.so, the floatexp kernels'
per-sample code built for the host (tests/host_kernels.h).  Frames come from the
product's host preparation (mxhip.frame_prepare), so the reference, BLA table and dc
tables are the ones the GPU kernels read."""
import ctypes as C
import os

import numpy as np

import mxhip

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "..", "mandelbrot-explorer_amd", "_build", "libmxhost.so")
_LIB = None


class Opts(C.Structure):
    _fields_ = [("form", C.c_int32), ("ede", C.c_int32), ("interior", C.c_int32), ("int_rep", C.c_int32),
                ("bla", C.c_int32), ("bla_minlevel", C.c_int32), ("interior_eps2", C.c_double),
                ("interior_confirm", C.c_int32), ("lean", C.c_int32), ("chunk", C.c_int32), ("threads", C.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C mandelbrot-explorer_amd`")
        L = C.CDLL(LIB_PATH)
        args = [C.POINTER(mxhip.FrameView), C.c_int, C.c_int, C.c_int, C.POINTER(Opts), C.c_void_p, C.c_int,
                C.c_void_p, C.c_void_p, C.c_void_p]
        L.mxh_run_fe64.argtypes = args
        L.mxh_run_fe32.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def view(fv):
    """An mx_frame_view over the arrays of a frame_prepare() dict (kept alive by the dict)."""
    v = mxhip.FrameView()
    keep = {}

    def put(name, dt, ct):
        a = np.ascontiguousarray(fv[name], dtype=dt)
        keep[name] = a
        setattr(v, name, _ptr(a, ct))

    for name in ("zf", "zfe_m", "bla_r2", "bla_coef", "dcol_m", "drow_m"):
        put(name, np.float64, C.c_double)
    for name in ("zfe_e", "lev_off", "lev_size", "dcol_e", "drow_e"):
        put(name, np.int32, C.c_int32)
    v.len, v.nlev, v.bla_total = fv["len"], fv["nlev"], len(fv["bla_r2"])
    v.rmax2, v.dcmax, v.dx_m, v.dx_e, v.esc2 = fv["rmax2"], fv["dcmax"], fv["dx_m"], fv["dx_e"], fv["esc2"]
    v.bla_minlevel, v.period = fv["bla_minlevel"], fv["period"]
    fv["_keep"] = keep
    return v


def run(fv, form, idx=None, ede=False, interior=True, int_rep=None, bla=True, bla_minlevel=None, lean=True,
        chunk=32, threads=None, eps2=3e-9 * 3e-9, confirm=4):
    """Samples `idx` (row-major indices, default all) of a prepared floatexp frame through
    the host build of the form's kernel code -> (n_esc, value, stats[n, 4])."""
    W, H = fv["W"], fv["H"]
    idx = np.arange(W * H, dtype=np.int32) if idx is None else np.ascontiguousarray(idx, dtype=np.int32)
    n = len(idx)
    o = Opts()
    o.form, o.ede, o.interior = form, int(ede), int(interior)
    o.int_rep = int(interior if int_rep is None else int_rep)
    o.bla = int(bla and fv["nlev"] > 0)
    o.bla_minlevel = fv["bla_minlevel"] if bla_minlevel is None else bla_minlevel
    o.interior_eps2, o.interior_confirm = eps2, confirm
    o.lean, o.chunk, o.threads = int(lean), chunk, threads or os.cpu_count()
    nesc = np.zeros(n, np.int32)
    val = np.zeros(n)
    st = np.zeros((n, 4), np.uint64)
    v = view(fv)
    fn = lib().mxh_run_fe64 if form == mxhip.MX_FORM_FE64 else lib().mxh_run_fe32
    rc = fn(C.byref(v), W, H, fv["mxit"], C.byref(o), idx.ctypes.data, n, nesc.ctypes.data, val.ctypes.data,
            st.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"mxh_run rc={rc}")
    return nesc, val, st
