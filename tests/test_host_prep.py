"""Host side of the product without a GPU: the C ABI library loads and exports every
declared symbol; frame preparation (grid, GMP reference orbit, BLA table, dc tables)
equals the oracle's restatement; requests outside the path are refused."""
import os
import re
import subprocess

import numpy as np
import pytest

import cases
import mxhip
import mxo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mx.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(mx_\w+)\(", hdr, re.M)))
    assert declared and set(declared) == set(mxhip.EXPORTS)
    lib = mxhip.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.mx_version() == 1


# fixed-point orbit widths on either side of the short-product range 16 <= L < 72
# (name: frame, precision bits): 1e1400 holds 4718 bits (76 limbs, full products); a
# 1e290 frame held at 700 bits has 12 limbs
FIXED_EDGE = {
    "exterior1400": (("0", "1", cases.pow10(1400), 24, 16, 4000, 0), 0),
    "deep290_700bit": ((cases.LOC["deep1_x"], cases.LOC["deep1_y"], cases.pow10(290), 25, 17, 20000, 0), 700),
}


@pytest.mark.parametrize("name", ["extref875", "minibrot875", "exterior1000", "subpixel737", "exterior1400",
                                  "deep290_700bit", "exterior1000_ede", "subpixel737_ede"])
def test_prepare_matches_oracle(name):
    """Floatexp-path preparation equals the oracle's: the fixed-point reference orbit's
    shadows bit for bit, reference choice, grid steps, BLA and dc tables, and the
    oracle's per-pixel loop run on the product's arrays reproduces the oracle frame."""
    (cre, cim, scale, W, H, mxit, cm), pb = FIXED_EDGE.get(name) or (cases.FRAMES[name], 0)
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm, prec_bits=pb)
    val, nesc, out, info = mxo.compute_fe(cre, cim, scale, W, H, mxit, c_method=cm, prec_bits=pb or None,
                                          want_orbit=True)
    limbs = (max(53, pb or mxo.precision_bits(scale)) + 127) // 64 + 1
    assert (limbs < 16 or limbs >= 72) == (name in FIXED_EDGE)
    assert fv["path"] == 2 and fv["escalated"] == 0
    for k in ("zf", "zfe_m", "zfe_e"):
        assert np.array_equal(info["orbit"][k], fv[k]), k
    assert fv["ref_it"] == info["ref_it"] and fv["ref_virtual"] == info["ref_virtual"]
    assert (fv["period"], fv["subpixel"]) == (info["period"], info["subpixel"])
    assert (fv["ref_x"], fv["ref_y"]) == (info["ref_x"], info["ref_y"])
    # the oracle's orbit at the same point: rebuild it from the grid the oracle uses
    (dxm, dxe), (dym, dye) = mxo.grid_info(cre, cim, scale, W, H)
    assert (fv["dx_m"], fv["dx_e"], fv["dy_m"], fv["dy_e"]) == (dxm, dxe, dym, dye)
    # BLA tables built by the product equal the oracle's build on the same orbit, bit for bit;
    # a sub-pixel nucleus in view turns BLA off (mandel_perturbation.cpp:1624-1629)
    b = mxo.bla_tables(fv["zf"], fv["ref_it"], fv["dcmax"])
    if fv["subpixel"]:
        assert fv["nlev"] == 0 and fv["rmax2"] == 0.0
    else:
        assert b["nlev"] == fv["nlev"]
        assert np.array_equal(b["off"], fv["lev_off"]) and np.array_equal(b["size"], fv["lev_size"])
        assert np.array_equal(b["r2"].view(np.int64), fv["bla_r2"].view(np.int64))
        assert np.array_equal(b["coef"].view(np.int64), fv["bla_coef"].view(np.int64))
        assert b["rmax2"] == fv["rmax2"]
    # dc tables: fe_mul_d(dxfe, j - ref_x) (mandel_perturbation.cpp:2160-2163)
    am, ae = mxo.fe_op(1, np.full(W, dxm), np.full(W, dxe, np.int64), np.arange(W) - fv["ref_x"], None)
    assert np.array_equal(am, fv["dcol_m"]) and np.array_equal(ae, fv["dcol_e"])
    # the oracle's per-pixel loop on the product's orbit reproduces the oracle's own frame
    orb = dict(zf=fv["zf"], zre_m=fv["zfe_m"][0::2], zre_e=fv["zfe_e"][0::2], zim_m=fv["zfe_m"][1::2], zim_e=fv["zfe_e"][1::2])
    bla = None if fv["subpixel"] else dict(nlev=b["nlev"], off=b["off"], size=b["size"], r2=b["r2"], l=b["l"],
                                           coef=b["coef"], rmax2=b["rmax2"])
    jj, ii = np.meshgrid(np.arange(W), np.arange(H))
    v2, n2, _ = mxo.pixels_fe(orb, bla, fv["dcol_m"][jj.ravel()], fv["dcol_e"][jj.ravel()], fv["drow_m"][ii.ravel()],
                              fv["drow_e"][ii.ravel()], mxit, fv["esc2"], fv["dx_m"], fv["dx_e"], ede=bool(cm & 2),
                              use_interior=bool(info["interior_used"]), bla_minlevel=fv["bla_minlevel"],
                              period=fv["period"])
    mask = np.ones((H, W), bool)
    if not fv["ref_virtual"]:
        mask[H // 2, W // 2] = False
        assert fv["ref_pixel_n"] == nesc[H // 2, W // 2]
        assert fv["ref_pixel_value"] == val[H // 2, W // 2]
    assert np.array_equal(n2.reshape(H, W)[mask], nesc[mask])
    assert np.array_equal(v2.reshape(H, W)[mask], val[mask])


def test_orbit_shadows_equal_mpfr_truncation():
    """The product's fixed-point orbit of the exact point c = i (all values dyadic) has
    the shadows of the oracle's MPFR orbit truncated to 53 bits, entry by entry.  For an odd frame the centre pixel (H/2, W/2) is the view
    centre, c0 + dx*(W-1)/2, up to the truncation residue of the grid arithmetic
    (|residue| < 2^-(precision-64)); entries that are exactly 0 in the orbit of the
    exact centre c = i carry that residue instead and are compared as zeros."""
    cre, cim, scale = "0", "1", cases.pow10(300)
    W, H, mxit = 65, 33, 5000
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit)
    o = mxo.orbit(cre, cim, fv["precision_bits"], mxit)
    assert o["ref_it"] == fv["ref_it"] == mxit
    assert np.array_equal(o["zf"], fv["zf"])
    for comp, (om, oe) in ((0, ("zre_m", "zre_e")), (1, ("zim_m", "zim_e"))):
        pm, pe = fv["zfe_m"][comp::2], fv["zfe_e"][comp::2]
        residue = pe < -(fv["precision_bits"] - 80)
        assert (o[om][residue] == 0).all()
        assert np.array_equal(o[om][~residue], pm[~residue]) and np.array_equal(o[oe][~residue], pe[~residue])


@pytest.mark.parametrize("kw,rc", [(dict(W=1), mxhip.MX_E_INVALID), (dict(mxit=0), mxhip.MX_E_INVALID),
                                   (dict(scale="1000000"), mxhip.MX_E_UNSUPPORTED),
                                   (dict(c_method=8), mxhip.MX_E_UNSUPPORTED), (dict(row_base=5), mxhip.MX_E_INVALID),
                                   (dict(scale="-4" + "0" * 300), mxhip.MX_E_INVALID)])
def test_invalid_requests_are_refused(kw, rc):
    args = dict(cre="0", cim="1", scale=cases.pow10(300), W=16, H=8, mxit=1000, c_method=0, row_base=0)
    args.update(kw)
    with pytest.raises(mxhip.MxError) as e:
        mxhip.frame_prepare(args["cre"], args["cim"], args["scale"], args["W"], args["H"], args["mxit"],
                            c_method=args["c_method"], row_base=args["row_base"])
    assert e.value.rc == rc


def test_backend_factory_declines_without_a_device():
    """createHipComputeBackend returns nullptr and says why when no device is usable, so
    the engine keeps its CPU backend (compute_backend.cpp:281-286)."""
    drv = os.path.join(ROOT, "mandelbrot-explorer_amd", "_build", "backend_driver")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([drv, "nodevice"], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0, r.stdout + r.stderr


DBL_FRAMES = {
    # 640x360 at scale 4e10, bailout 65536, 128-bit precision
    "c1": ("-0.743643887037151", "0.131825904205330", "40000000000", 640, 360, 8192, 65536.0, 128),
    # verify.cpp:11876-11881 and 11864-11870
    "point31": ("-0.749139567333446841955467474699747367338762518832278501811",
                "0.040823298514634751035521346975478853963578400940553676068", "7354177" + "0" * 25, 48, 32, 500000, 0.0, 0),
    "slowpoint": ("-1.1758621450236620370", "-0.2447677973532398022", "782679600", 48, 32, 500000, 0.0, 0),
}


@pytest.mark.parametrize("c_method", [0, mxhip.MX_EXTERIOR_DIST_EST])
@pytest.mark.parametrize("name", sorted(DBL_FRAMES))
def test_double_prepare_matches_oracle(name, c_method):
    """Plain-double path host preparation: GMP orbit, |X|^2/1e6, series coefficients
    (calCoefficient), half extent delta and the truncated dc tables equal the
    oracle's, bit for bit; under EDE also the reference derivative, the derivative
    series and the BLA derivative couplings."""
    cre, cim, scale, W, H, mxit, R, pb = DBL_FRAMES[name]
    ede = c_method == mxhip.MX_EXTERIOR_DIST_EST
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb, c_method=c_method)
    o = mxo.dbl_setup(cre, cim, scale, W, H, mxit, escape_radius=R or 1e8, prec_bits=pb or None, ede=ede)
    assert fv["path"] == 1 and fv["ref_virtual"] == 0 and (fv["ref_x"], fv["ref_y"]) == (W // 2, H // 2)
    assert fv["ref_it"] == o["ref_it"] and (fv["sa_it"], fv["sa_order"]) == (o["sa_it"], o["sa_order"])
    keys = ["zf", "z_m3", "sa_coef", "dcol_d", "drow_d"] + (["dfr", "dfi", "sa_coef_b"] if ede else [])
    for k in keys:
        assert np.array_equal(fv[k].view(np.int64), o[k].view(np.int64)), k
    assert fv["sa_delta"] == o["sa_delta"]
    b = mxo.bla_tables(fv["zf"], fv["ref_it"], fv["dcmax"], ede=ede, dfr=fv.get("dfr"), dfi=fv.get("dfi"))
    assert b["nlev"] == fv["nlev"] and np.array_equal(b["r2"], fv["bla_r2"]) and np.array_equal(b["coef"], fv["bla_coef"])
    if ede:
        assert np.array_equal(b["der"].view(np.int64), fv["bla_der"].view(np.int64))


def test_escalated_frame_prepares_the_floatexp_reference():
    """The near-zero escalation: the product hands the frame to the floatexp path with the
    oracle's reference (exact view centre), orbit and BLA table."""
    from test_oracle_dbl import ESCALATE
    cre, cim, scale, W, H, mxit, pb = ESCALATE
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, prec_bits=pb)
    val, nesc, out, info = mxo.compute(cre, cim, scale, W, H, mxit, prec_bits=pb)
    assert fv["path"] == 2 and fv["escalated"] == 1 and info["escalated"] == 1
    assert fv["ref_it"] == info["ref_it"] and fv["ref_virtual"] == 1 == info["ref_virtual"]
    assert (fv["ref_x"], fv["ref_y"]) == (info["ref_x"], info["ref_y"])


@pytest.mark.parametrize("sub,c_method,scale,rc", [
    (2, 1, cases.pow10(300), mxhip.MX_E_INVALID),          # even sub: past the engine's buffer
    (17, 1, cases.pow10(300), mxhip.MX_E_INVALID),
])
def test_supersampling_requests_refused(sub, c_method, scale, rc):
    with pytest.raises(mxhip.MxError) as e:
        mxhip.frame_prepare("0", "1", scale, 16, 8, 1000, c_method=c_method, sub=sub)
    assert e.value.rc == rc


def test_supersampling_prepare_accepted():
    fv = mxhip.frame_prepare("0", "1", cases.pow10(300), 16, 8, 1000, c_method=1, sub=3)
    assert fv["path"] == 2
    fv = mxhip.frame_prepare("0", "1", "40000000000", 16, 8, 1000, c_method=1, sub=3)
    assert fv["path"] == 1
    fv = mxhip.frame_prepare("0", "1", "40000000000", 16, 8, 1000, c_method=1, sub=3, options=mxhip.MX_OPT_FLOATEXP)
    assert fv["path"] == 2


REF_ENGINE = "/root/reference/src/engine"


@pytest.mark.skipif(not os.path.isdir(REF_ENGINE), reason="the reference tree is not present")
def test_plugin_abi_matches_the_engine_header(tmp_path):
    """The plugin built against this library's restated compute_backend.h and against
    the engine's own header sees the same ComputeRequest field offsets, the same info
    struct sizes and the same vtable slot for every IComputeBackend virtual
    (src/engine/compute_backend.h:21-113)."""
    out = {}
    for name, extra in (("restated", []), ("engine", ["-I", REF_ENGINE])):
        exe = str(tmp_path / name)
        subprocess.run(["g++", "-std=c++17", "-I/opt/conda/include", *extra, os.path.join(ROOT, "tests", "abi_probe.cpp"),
                        "-o", exe], check=True, capture_output=True)
        out[name] = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    assert out["restated"][0] == "header restated" and out["engine"][0] == "header engine"
    assert out["restated"][1:] == out["engine"][1:]
    assert "slot compute 57" in out["engine"]
