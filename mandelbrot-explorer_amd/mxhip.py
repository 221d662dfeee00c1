"""ctypes binding of libmxhip.so (include/mx.h) for the Python side: tests, bench.py
and smoke().  The library must be built (make -C mandelbrot-explorer_amd); there is no
fallback, every call goes to the HIP library."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmxhip.so")

MX_OK, MX_E_INVALID, MX_E_NODEVICE, MX_E_HIP, MX_E_CANCELLED, MX_E_BUSY, MX_E_NOMEM, MX_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6, -7
MX_FORM_FE64, MX_FORM_FE32, MX_FORM_DOUBLE, MX_FORM_SC64 = 0, 1, 2, 3
MX_EXTERIOR_DIST_EST = 2
MX_OPT_VIEW_CENTER_REF, MX_OPT_FLOATEXP, MX_OPT_NO_PERIODIC, MX_OPT_PERIODIC = 1, 2, 4, 8
ERRORS = {-1: "invalid", -2: "no device", -3: "hip error", -4: "cancelled", -5: "busy", -6: "out of memory", -7: "unsupported"}


class MxError(RuntimeError):
    def __init__(self, rc, msg=""):
        super().__init__(f"mx error {rc} ({ERRORS.get(rc, '?')}): {msg}")
        self.rc = rc


class RefDesc(C.Structure):
    _fields_ = [("len", C.c_int32), ("zf", C.c_void_p), ("zfe_m", C.c_void_p), ("zfe_e", C.c_void_p),
                ("nlev", C.c_int32), ("lev_off", C.c_void_p), ("lev_size", C.c_void_p),
                ("bla_r2", C.c_void_p), ("bla_coef", C.c_void_p), ("rmax2", C.c_double),
                ("z_m3", C.c_void_p), ("sa_it", C.c_int32), ("sa_order", C.c_int32), ("sa_coef", C.c_void_p),
                ("sa_delta_re", C.c_double), ("sa_delta_im", C.c_double), ("df", C.c_void_p), ("bla_der", C.c_void_p),
                ("sa_coef_b", C.c_void_p), ("period", C.c_int32)]


class Tile(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("stride", C.c_int32), ("max_iter", C.c_int32),
                ("esc2", C.c_double), ("form", C.c_int32), ("ede", C.c_int32), ("interior", C.c_int32),
                ("int_rep", C.c_int32), ("bla", C.c_int32), ("bla_minlevel", C.c_int32),
                ("interior_eps2", C.c_double), ("interior_confirm", C.c_int32),
                ("dcol_m", C.c_void_p), ("dcol_e", C.c_void_p), ("drow_m", C.c_void_p), ("drow_e", C.c_void_p),
                ("dx_m", C.c_double), ("dx_e", C.c_int32), ("skip_sample", C.c_int32),
                ("dcol_d", C.c_void_p), ("drow_d", C.c_void_p), ("list", C.c_void_p), ("list_len", C.c_int32),
                ("chunk", C.c_int32), ("grid_blocks", C.c_int32), ("n_samples", C.c_int32),
                ("sample_dc_m", C.c_void_p), ("sample_dc_e", C.c_void_p), ("quorum", C.c_int32),
                ("lds_stage", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("center_re", C.c_char_p), ("center_im", C.c_char_p), ("scale", C.c_char_p),
                ("width", C.c_int32), ("height", C.c_int32), ("full_height", C.c_int32), ("row_base", C.c_int32),
                ("sub", C.c_int32), ("max_iterations", C.c_int32), ("coloring_method", C.c_int32),
                ("escape_radius", C.c_float), ("precision_bits", C.c_int32), ("form", C.c_int32),
                ("use_bla", C.c_int32), ("bla_minskip", C.c_int32), ("interior", C.c_int32),
                ("iterations", C.c_void_p), ("values", C.c_void_p), ("n_esc", C.c_void_p), ("options", C.c_int32)]


class FrameInfo(C.Structure):
    _fields_ = [("path", C.c_int32), ("ref_it", C.c_int32), ("ref_escaped", C.c_int32), ("ref_virtual", C.c_int32),
                ("ref_x", C.c_double), ("ref_y", C.c_double), ("interior_used", C.c_int32),
                ("cutoff_sensitive", C.c_int32), ("precision_bits", C.c_int32), ("bla_levels", C.c_int32),
                ("ref_seconds", C.c_double), ("bla_seconds", C.c_double), ("upload_seconds", C.c_double),
                ("probe_ms", C.c_double), ("kernel_ms", C.c_double), ("total_seconds", C.c_double),
                ("steps", C.c_uint64), ("bla_applies", C.c_uint64), ("bla_skipped", C.c_uint64),
                ("rebases", C.c_uint64), ("nominal_iterations", C.c_uint64), ("passes", C.c_int32),
                ("sa_it", C.c_int32), ("escalated", C.c_int32), ("ss_pixels", C.c_int32), ("ss_ms", C.c_double),
                ("period", C.c_int32), ("subpixel", C.c_int32), ("gmp_fallback", C.c_int32)]


class FrameView(C.Structure):
    _fields_ = [("len", C.c_int32), ("zf", C.POINTER(C.c_double)), ("zfe_m", C.POINTER(C.c_double)),
                ("zfe_e", C.POINTER(C.c_int32)), ("nlev", C.c_int32), ("lev_off", C.POINTER(C.c_int32)),
                ("lev_size", C.POINTER(C.c_int32)), ("bla_total", C.c_int64), ("bla_r2", C.POINTER(C.c_double)),
                ("bla_coef", C.POINTER(C.c_double)), ("rmax2", C.c_double), ("dcmax", C.c_double),
                ("dcol_m", C.POINTER(C.c_double)), ("dcol_e", C.POINTER(C.c_int32)),
                ("drow_m", C.POINTER(C.c_double)), ("drow_e", C.POINTER(C.c_int32)),
                ("dx_m", C.c_double), ("dx_e", C.c_int32), ("dy_m", C.c_double), ("dy_e", C.c_int32),
                ("esc2", C.c_double), ("bla_minlevel", C.c_int32), ("ref_it", C.c_int32),
                ("ref_escaped", C.c_int32), ("ref_virtual", C.c_int32), ("ref_x", C.c_double), ("ref_y", C.c_double),
                ("ref_pixel_n", C.c_int32), ("ref_pixel_value", C.c_double), ("cutoff_sensitive", C.c_int32),
                ("precision_bits", C.c_int32), ("ref_seconds", C.c_double), ("bla_seconds", C.c_double),
                ("path", C.c_int32), ("z_m3", C.POINTER(C.c_double)), ("sa_it", C.c_int32), ("sa_order", C.c_int32),
                ("sa_coef", C.POINTER(C.c_double)), ("sa_delta_re", C.c_double), ("sa_delta_im", C.c_double),
                ("dcol_d", C.POINTER(C.c_double)), ("drow_d", C.POINTER(C.c_double)),
                ("dfr", C.POINTER(C.c_double)), ("dfi", C.POINTER(C.c_double)), ("bla_der", C.POINTER(C.c_double)),
                ("sa_coef_b", C.POINTER(C.c_double)), ("escalated", C.c_int32), ("period", C.c_int32),
                ("subpixel", C.c_int32)]


# every symbol include/mx.h declares
EXPORTS = ["mx_version", "mx_device_count", "mx_ctx_create", "mx_ctx_destroy", "mx_last_error", "mx_cancel",
           "mx_reset_cancellation", "mx_ref_upload", "mx_ref_release", "mx_plan_create", "mx_plan_launch",
           "mx_plan_wait", "mx_plan_read", "mx_plan_unresolved", "mx_plan_destroy", "mx_tile_render", "mx_compute",
           "mx_frame_prepare", "mx_frame_free", "mx_batch_create", "mx_batch_launch", "mx_batch_wait",
           "mx_batch_destroy", "mx_ctx_cu_count", "mx_graph_create", "mx_graph_launch", "mx_graph_wait",
           "mx_graph_destroy", "mx_palette", "mx_colorize", "mx_plan_colorize", "mx_tile_render_noalloc", "mx_copy_probe"]

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C mandelbrot-explorer_amd` (no fallback path exists)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mx_version.restype = C.c_int
        L.mx_device_count.restype = C.c_int
        L.mx_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.mx_ctx_destroy.argtypes = [vp]
        L.mx_last_error.argtypes = [vp]
        L.mx_last_error.restype = C.c_char_p
        L.mx_cancel.argtypes = [vp]
        L.mx_reset_cancellation.argtypes = [vp]
        L.mx_ref_upload.argtypes = [vp, C.POINTER(RefDesc), C.POINTER(vp)]
        L.mx_ref_release.argtypes = [vp]
        L.mx_plan_create.argtypes = [vp, vp, C.POINTER(Tile), C.POINTER(vp)]
        L.mx_plan_launch.argtypes = [vp]
        L.mx_plan_wait.argtypes = [vp, C.POINTER(C.c_double)]
        L.mx_plan_read.argtypes = [vp, vp, vp, vp, vp]
        L.mx_plan_destroy.argtypes = [vp]
        L.mx_plan_unresolved.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(C.c_int32)]
        L.mx_tile_render.argtypes = [vp, vp, C.POINTER(Tile), vp, vp, vp, vp]
        L.mx_tile_render_noalloc.argtypes = [vp, vp, C.POINTER(Tile), vp, vp, vp, vp]
        L.mx_copy_probe.argtypes = [vp, C.c_size_t, vp]
        L.mx_ctx_cu_count.argtypes = [vp]
        L.mx_graph_create.argtypes = [vp, C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.mx_graph_launch.argtypes = [vp]
        L.mx_graph_wait.argtypes = [vp, C.POINTER(C.c_double)]
        L.mx_graph_destroy.argtypes = [vp]
        L.mx_batch_create.argtypes = [vp, C.POINTER(vp), C.c_int32, C.POINTER(vp)]
        L.mx_batch_launch.argtypes = [vp]
        L.mx_batch_wait.argtypes = [vp, C.POINTER(C.c_double)]
        L.mx_batch_destroy.argtypes = [vp]
        L.mx_compute.argtypes = [vp, C.POINTER(Request), C.POINTER(FrameInfo)]
        L.mx_frame_prepare.argtypes = [C.POINTER(Request), C.POINTER(vp), C.POINTER(FrameView)]
        L.mx_frame_free.argtypes = [vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_request(cre, cim, scale, W, H, mxit, c_method=0, escape_radius=0.0, prec_bits=0, form=MX_FORM_FE64,
                 use_bla=-1, bla_minskip=0, interior=-1, full_h=0, row_base=0, options=0, sub=1):
    r = Request()
    r.center_re, r.center_im, r.scale = cre.encode(), cim.encode(), scale.encode()
    r.width, r.height, r.full_height, r.row_base, r.sub = W, H, full_h, row_base, sub
    r.max_iterations, r.coloring_method, r.escape_radius = mxit, c_method, escape_radius
    r.precision_bits, r.form, r.use_bla, r.bla_minskip, r.interior = prec_bits, form, use_bla, bla_minskip, interior
    r.options = options
    return r


def _arr(ptr, n, dt):
    if n == 0 or not ptr:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)


def palette():
    """The engine's palette, (3, 2048) float32 (host only)."""
    pal = np.zeros(3 * 2048, np.float32)
    rc = lib().mx_palette(_p(pal))
    if rc != MX_OK:
        raise MxError(rc, "mx_palette")
    return pal.reshape(3, 2048)


def frame_prepare(cre, cim, scale, W, H, mxit, **kw):
    """Host-only preparation (grid, GMP reference orbit, BLA, dc tables) -> dict."""
    req = make_request(cre, cim, scale, W, H, mxit, **kw)
    fr = C.c_void_p()
    v = FrameView()
    rc = lib().mx_frame_prepare(C.byref(req), C.byref(fr), C.byref(v))
    if rc != MX_OK:
        raise MxError(rc, "frame_prepare")
    try:
        n = v.len
        tot = v.bla_total
        d = dict(len=n, zf=_arr(v.zf, 2 * n, np.float64), zfe_m=_arr(v.zfe_m, 2 * n, np.float64),
                 zfe_e=_arr(v.zfe_e, 2 * n, np.int32), nlev=v.nlev, lev_off=_arr(v.lev_off, v.nlev, np.int32),
                 lev_size=_arr(v.lev_size, v.nlev, np.int32), bla_r2=_arr(v.bla_r2, tot, np.float64),
                 bla_coef=_arr(v.bla_coef, 4 * tot, np.float64), rmax2=v.rmax2, dcmax=v.dcmax,
                 dcol_m=_arr(v.dcol_m, W, np.float64), dcol_e=_arr(v.dcol_e, W, np.int32),
                 drow_m=_arr(v.drow_m, H, np.float64), drow_e=_arr(v.drow_e, H, np.int32),
                 dx_m=v.dx_m, dx_e=v.dx_e, dy_m=v.dy_m, dy_e=v.dy_e, esc2=v.esc2, bla_minlevel=v.bla_minlevel,
                 ref_it=v.ref_it, ref_escaped=v.ref_escaped, ref_virtual=v.ref_virtual, ref_x=v.ref_x,
                 ref_y=v.ref_y, ref_pixel_n=v.ref_pixel_n, ref_pixel_value=v.ref_pixel_value,
                 cutoff_sensitive=v.cutoff_sensitive, precision_bits=v.precision_bits,
                 ref_seconds=v.ref_seconds, bla_seconds=v.bla_seconds, W=W, H=H, mxit=mxit, path=v.path,
                 escalated=v.escalated, period=v.period, subpixel=v.subpixel)
        if v.path == 1:
            d.update(z_m3=_arr(v.z_m3, n, np.float64), sa_it=v.sa_it, sa_order=v.sa_order,
                     sa_coef=_arr(v.sa_coef, 60, np.float64), sa_delta=complex(v.sa_delta_re, v.sa_delta_im),
                     dcol_d=_arr(v.dcol_d, W, np.float64), drow_d=_arr(v.drow_d, H, np.float64))
            if v.dfr:
                d.update(dfr=_arr(v.dfr, n, np.float64), dfi=_arr(v.dfi, n, np.float64),
                         bla_der=_arr(v.bla_der, 4 * tot, np.float64), sa_coef_b=_arr(v.sa_coef_b, 60, np.float64))
    finally:
        lib().mx_frame_free(fr)
    return d


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().mx_ctx_create(device, C.byref(self.h))
        if rc != MX_OK:
            raise MxError(rc, "mx_ctx_create")
        self.cu_count = lib().mx_ctx_cu_count(self.h)

    def close(self):
        if self.h:
            lib().mx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def error(self):
        return lib().mx_last_error(self.h).decode()

    def check(self, rc, what):
        if rc != MX_OK:
            raise MxError(rc, f"{what}: {self.error()}")

    def cancel(self):
        lib().mx_cancel(self.h)

    def reset_cancellation(self):
        lib().mx_reset_cancellation(self.h)

    def colorize(self, it, ss=1, c_method=0, density=60.0):
        """RGB bytes (H, W, 3) of the (H ss) x (W ss) float buffer `it` (the engine's
        layout), coloured and downsampled on the device."""
        it = np.ascontiguousarray(it, np.float32)
        Hs, Ws = it.shape
        W, H = Ws // ss, Hs // ss
        assert W * ss == Ws and H * ss == Hs
        rgb = np.zeros((H, W, 3), np.uint8)
        self.check(lib().mx_colorize(self.h, _p(it), W, H, ss, c_method, C.c_float(density), _p(rgb)), "mx_colorize")
        return rgb

    def compute(self, cre, cim, scale, W, H, mxit, sub=1, **kw):
        """One frame; with sub > 1 the buffers are (H sub) x (W sub) cells (-10, 0 and
        INT_MIN where no sample was computed)."""
        req = make_request(cre, cim, scale, W, H, mxit, sub=sub, **kw)
        n = W * H * sub * sub
        it = np.full(n, -10.0, np.float32)
        val = np.zeros(n)
        nesc = np.full(n, np.iinfo(np.int32).min, np.int32)
        req.iterations, req.values, req.n_esc = _p(it), _p(val), _p(nesc)
        info = FrameInfo()
        rc = lib().mx_compute(self.h, C.byref(req), C.byref(info))
        self.check(rc, "mx_compute")
        inf = {k: getattr(info, k) for k, _ in FrameInfo._fields_}
        return it.reshape(H * sub, W * sub), val.reshape(H * sub, W * sub), nesc.reshape(H * sub, W * sub), inf

    def upload(self, fv):
        """Upload a reference from a frame_prepare() dict (or the same keys)."""
        keep = []

        def a(x, dt):
            x = np.ascontiguousarray(x, dt)
            keep.append(x)
            return _p(x)

        d = RefDesc()
        d.len = fv["len"]
        d.period = fv.get("period", 0)
        d.zf, d.zfe_m, d.zfe_e = a(fv["zf"], np.float64), a(fv["zfe_m"], np.float64), a(fv["zfe_e"], np.int32)
        d.nlev = fv["nlev"]
        if d.nlev:
            d.lev_off, d.lev_size = a(fv["lev_off"], np.int32), a(fv["lev_size"], np.int32)
            d.bla_r2, d.bla_coef = a(fv["bla_r2"], np.float64), a(fv["bla_coef"], np.float64)
        d.rmax2 = fv["rmax2"]
        if fv.get("path") == 1:
            d.z_m3 = a(fv["z_m3"], np.float64)
            d.sa_it, d.sa_order = fv["sa_it"], fv["sa_order"]
            d.sa_coef = a(fv["sa_coef"], np.float64)
            d.sa_delta_re, d.sa_delta_im = fv["sa_delta"].real, fv["sa_delta"].imag
            if "dfr" in fv:
                df = np.empty(2 * fv["len"])
                df[0::2], df[1::2] = fv["dfr"], fv["dfi"]
                d.df = a(df, np.float64)
                d.bla_der = a(fv["bla_der"], np.float64) if fv["nlev"] else None
                d.sa_coef_b = a(fv["sa_coef_b"], np.float64)
        ref = C.c_void_p()
        self.check(lib().mx_ref_upload(self.h, C.byref(d), C.byref(ref)), "mx_ref_upload")
        return Ref(ref)

    def tile(self, fv, stride=1, form=MX_FORM_FE64, ede=False, interior=True, int_rep=None, bla=True,
             bla_minlevel=None, mxit=None, samples=None, chunk=0, grid_blocks=0, quorum=0, lds_stage=0):
        t = Tile()
        t.chunk, t.grid_blocks, t.quorum, t.lds_stage = chunk, grid_blocks, quorum, lds_stage
        t.width, t.height, t.stride = fv["W"], fv["H"], stride
        t.max_iter = mxit or fv["mxit"]
        t.esc2, t.form, t.ede = fv["esc2"], form, int(ede)
        t.interior = int(interior)
        t.int_rep = int(interior if int_rep is None else int_rep)
        t.bla = int(bla and fv["nlev"] > 0)
        t.bla_minlevel = fv["bla_minlevel"] if bla_minlevel is None else bla_minlevel
        t.interior_eps2, t.interior_confirm = 3e-9 * 3e-9, 4
        if form == MX_FORM_DOUBLE:
            keep = [np.ascontiguousarray(fv[k], np.float64) for k in ("dcol_d", "drow_d")]
            t.dcol_d, t.drow_d = [_p(x) for x in keep]
        else:
            keep = [np.ascontiguousarray(fv[k], dt) for k, dt in
                    (("dcol_m", np.float64), ("dcol_e", np.int32), ("drow_m", np.float64), ("drow_e", np.int32))]
            t.dcol_m, t.dcol_e, t.drow_m, t.drow_e = [_p(x) for x in keep]
        if samples is not None:
            lst = np.ascontiguousarray(samples, np.int32)
            keep.append(lst)
            t.list, t.list_len = _p(lst), lst.size
        t.dx_m, t.dx_e = fv["dx_m"], fv["dx_e"]
        t.skip_sample = -1
        t._keep = keep
        return t

    def render(self, ref, tile):
        n = tile.width * tile.height
        nesc = np.zeros(n, np.int32)
        val = np.zeros(n)
        out = np.zeros(n, np.float32)
        st = np.zeros(4, np.uint64)
        rc = lib().mx_tile_render(self.h, ref.h, C.byref(tile), _p(nesc), _p(val), _p(out), _p(st))
        self.check(rc, "mx_tile_render")
        return nesc, val, out, st

    def copy_probe(self, nbytes):
        """One device copy kernel over nbytes (counter calibration); its time in ms."""
        ms = C.c_double()
        self.check(lib().mx_copy_probe(self.h, nbytes, C.cast(C.byref(ms), C.c_void_p)), "mx_copy_probe")
        return ms.value

    def render_noalloc(self, ref, tile):
        """mx_tile_render_noalloc: the same results, staged in the device's fixed scratch area."""
        n = tile.width * tile.height
        nesc = np.zeros(n, np.int32)
        val = np.zeros(n)
        out = np.zeros(n, np.float32)
        st = np.zeros(4, np.uint64)
        rc = lib().mx_tile_render_noalloc(self.h, ref.h, C.byref(tile), _p(nesc), _p(val), _p(out), _p(st))
        self.check(rc, "mx_tile_render_noalloc")
        return nesc, val, out, st

    def plan(self, ref, tile):
        p = C.c_void_p()
        self.check(lib().mx_plan_create(self.h, ref.h, C.byref(tile), C.byref(p)), "mx_plan_create")
        return Plan(self, p, tile)

    def graph(self, plans):
        """The plans' launches, in order, captured as one HIP graph."""
        arr = (C.c_void_p * len(plans))(*[p.h for p in plans])
        g = C.c_void_p()
        self.check(lib().mx_graph_create(self.h, arr, len(plans), C.byref(g)), "mx_graph_create")
        return Graph(self, g, plans)

    def batch(self, plans):
        """One persistent launch over several plans (run in the given order)."""
        arr = (C.c_void_p * len(plans))(*[p.h for p in plans])
        b = C.c_void_p()
        self.check(lib().mx_batch_create(self.h, arr, len(plans), C.byref(b)), "mx_batch_create")
        return Batch(self, b, plans)


class Ref:
    def __init__(self, h):
        self.h = h

    def release(self):
        if self.h:
            lib().mx_ref_release(self.h)
            self.h = C.c_void_p()


class Graph:
    def __init__(self, ctx, h, plans):
        self.ctx, self.h, self.plans = ctx, h, list(plans)

    def launch(self):
        self.ctx.check(lib().mx_graph_launch(self.h), "mx_graph_launch")

    def wait(self):
        ms = C.c_double()
        self.ctx.check(lib().mx_graph_wait(self.h, C.byref(ms)), "mx_graph_wait")
        return ms.value

    def destroy(self):
        if self.h:
            lib().mx_graph_destroy(self.h)
            self.h = C.c_void_p()


class Batch:
    def __init__(self, ctx, h, plans):
        self.ctx, self.h, self.plans = ctx, h, list(plans)

    def launch(self):
        self.ctx.check(lib().mx_batch_launch(self.h), "mx_batch_launch")

    def wait(self):
        ms = C.c_double()
        self.ctx.check(lib().mx_batch_wait(self.h, C.byref(ms)), "mx_batch_wait")
        return ms.value

    def destroy(self):
        if self.h:
            lib().mx_batch_destroy(self.h)
            self.h = C.c_void_p()


class Plan:
    def __init__(self, ctx, h, tile):
        self.ctx, self.h, self.tile = ctx, h, tile

    def launch(self):
        self.ctx.check(lib().mx_plan_launch(self.h), "mx_plan_launch")

    def wait(self):
        ms = C.c_double()
        self.ctx.check(lib().mx_plan_wait(self.h, C.byref(ms)), "mx_plan_wait")
        return ms.value

    def read(self):
        n = self.tile.width * self.tile.height
        nesc = np.zeros(n, np.int32)
        val = np.zeros(n)
        out = np.zeros(n, np.float32)
        st = np.zeros(4, np.uint64)
        self.ctx.check(lib().mx_plan_read(self.h, _p(nesc), _p(val), _p(out), _p(st)), "mx_plan_read")
        return nesc, val, out, st

    def unresolved(self):
        """(sample indices, glitch metrics) the last launch left unresolved (double form)."""
        cnt = C.c_int32(0)
        self.ctx.check(lib().mx_plan_unresolved(self.h, None, None, 0, C.byref(cnt)), "mx_plan_unresolved")
        idx = np.zeros(cnt.value, np.int32)
        gp = np.zeros(cnt.value)
        if cnt.value:
            self.ctx.check(lib().mx_plan_unresolved(self.h, _p(idx), _p(gp), cnt.value, C.byref(cnt)), "mx_plan_unresolved")
        return idx, gp

    def colorize(self, ss=1, c_method=0, density=60.0):
        """RGB bytes (H/ss, W/ss, 3) of the last launch's float output, coloured on the device."""
        W, H = self.tile.width // ss, self.tile.height // ss
        rgb = np.zeros((H, W, 3), np.uint8)
        self.ctx.check(lib().mx_plan_colorize(self.h, ss, c_method, C.c_float(density), _p(rgb)), "mx_plan_colorize")
        return rgb

    def destroy(self):
        if self.h:
            lib().mx_plan_destroy(self.h)
            self.h = C.c_void_p()
