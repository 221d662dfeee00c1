"""Parity of the plain-double HIP kernel and its pass driver (through the C ABI) with
the oracle's restatement of the engine's double path (series seed, double BLA,
glitch re-referencing), and with the reference's own golden frames.  Escape counts
bit-exact; smooth values within 1e-9 relative (only log() differs: device math
library against the host's)."""
import numpy as np
import pytest

import mxhip
import mxo
from test_host_prep import DBL_FRAMES
from test_oracle_dbl import _stats
import cases

REL = 1e-9
pytestmark = pytest.mark.gpu

FRAMES = dict(DBL_FRAMES)
FRAMES["ticktock"] = (cases.LOC["ticktock_x"], cases.LOC["ticktock_y"], cases.pow10(141), 120, 80, 200000, 0.0, 0)
FRAMES["flake"] = (cases.LOC["flake_x"], cases.LOC["flake_y"], cases.pow10(157), 120, 80, 200000, 0.0, 0)


def _check(val_g, nesc_g, val_o, nesc_o, label):
    bad = np.argwhere(nesc_g != nesc_o)
    assert bad.size == 0, f"{label}: {len(bad)} escape counts differ, first {bad[:5].tolist()} gpu={nesc_g[tuple(bad[0])]} cpu={nesc_o[tuple(bad[0])]}"
    ext = nesc_o > 0
    d = np.abs(val_g[ext] - val_o[ext]) / np.maximum(np.abs(val_o[ext]), 1e-300)
    assert (d <= REL).all(), f"{label}: max rel value diff {d.max()}"
    assert (val_g[(nesc_o == -1) | (nesc_o == -2)] == -2.0).all()


@pytest.mark.parametrize("name,c_method", [("c1", 0), ("point31", 0), ("slowpoint", 0), ("ticktock", 0), ("flake", 0),
                                           ("c1", 2), ("point31", 2), ("ticktock", 2)])
def test_double_frame_matches_oracle(ctx, name, c_method):
    cre, cim, scale, W, H, mxit, R, pb = FRAMES[name]
    print(f"[dbl parity] {name} {W}x{H} mxit={mxit} c_method={c_method}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb, c_method=c_method)
    val_o, nesc_o, out_o, info_o = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=R or 1e8, prec_bits=pb or None,
                                                   c_method=c_method)
    assert info["path"] == 1 and info["passes"] == info_o["passes"] and info["sa_it"] == info_o["sa_it0"]
    _check(val, nesc, val_o, nesc_o, name)
    assert np.array_equal(it < -1.5, out_o < -1.5) and not (it == -10).any()
    for k in ("steps", "bla_applies", "bla_skipped", "rebases"):
        assert info[k] == info_o[k], k


# the period-106 minibrot of the ticktock view at scale 2e9: the 8x8-strided probe hits
# its interior
FRAMES["mini106"] = ("-1.4799461355795131264511335866615963781", "0.00090132482780272173726842509958825902",
                     "2000000000", 64, 48, 20000, 0.0, 0)


@pytest.mark.parametrize("name,interior_used", [("c1", 0), ("mini106", 1)])
def test_first_pass_beside_the_probe(ctx, name, interior_used):
    """The first pass starts on a second stream beside the interior-gate probe, with
    interior detection off: kept when the probe finds no interior sample (c1), stopped
    and discarded when it finds one (mini106), after which the pass runs again with
    detection on.  Both frames equal the oracle."""
    cre, cim, scale, W, H, mxit, R, pb = FRAMES[name]
    print(f"[dbl beside probe] {name} {W}x{H} mxit={mxit}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb)
    val_o, nesc_o, out_o, info_o = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=R or 1e8, prec_bits=pb or None)
    assert info["path"] == 1 and info["interior_used"] == interior_used and info["passes"] == info_o["passes"]
    _check(val, nesc, val_o, nesc_o, name)
    for k in ("steps", "bla_applies", "bla_skipped", "rebases"):
        assert info[k] == info_o[k], k
    # the context is reusable after a discarded side pass
    it2, val2, nesc2, _ = ctx.compute(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb)
    assert np.array_equal(nesc2, nesc) and np.array_equal(val2, val)


@pytest.mark.parametrize("name,golden,tol", [("point31", "point_7e31.f32", (12000.0, 200.0, 4000.0)),
                                             ("slowpoint", "slowpoint_7e8.f32", (90000.0, 300.0, 3000.0))])
def test_double_frame_matches_reference_golden(ctx, name, golden, tol):
    """verify.cpp:11860-11881: class exact, max/mean/p99 smooth difference bounded."""
    cre, cim, scale, W, H, mxit, R, pb = FRAMES[name]
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit)
    mis, mx, mean, p99 = _stats(it, cases.golden(golden, H, W))
    assert mis == 0 and mx <= tol[0] and mean <= tol[1] and p99 <= tol[2], (mis, mx, mean, p99)


@pytest.mark.parametrize("ede", [False, True])
def test_double_kernel_on_product_arrays(ctx, ede):
    """Kernel alone: one launch over the frame against the product's own reference,
    series and dc tables, and the oracle's loop on the same arrays, unresolved
    samples (glitches) and their metric included."""
    cre, cim, scale, W, H, mxit, R, pb = FRAMES["point31"]
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=2 if ede else 0)
    ref = ctx.upload(fv)
    try:
        t = ctx.tile(fv, form=mxhip.MX_FORM_DOUBLE, interior=True, ede=ede)
        plan = ctx.plan(ref, t)
        try:
            plan.launch()
            plan.wait()
            nesc, val, out, st = plan.read()
            uidx, ugp = plan.unresolved()
        finally:
            plan.destroy()
    finally:
        ref.release()
    b = mxo.bla_tables(fv["zf"], fv["ref_it"], fv["dcmax"], ede=ede, dfr=fv.get("dfr"), dfi=fv.get("dfi"))
    bla = dict(nlev=b["nlev"], off=b["off"], size=b["size"], r2=b["r2"], l=b["l"], coef=b["coef"], der=b["der"],
               rmax2=b["rmax2"])
    jj, ii = np.meshgrid(np.arange(W), np.arange(H))
    dxf = float(np.ldexp(fv["dx_m"], fv["dx_e"]))
    v2, n2, gp2, st2 = mxo.pixels_dbl(fv["zf"], fv["z_m3"], bla, fv["sa_coef"], fv["sa_it"], fv["sa_order"], fv["sa_delta"],
                                      fv["dcol_d"][jj.ravel()], fv["drow_d"][ii.ravel()], mxit, fv["esc2"],
                                      bla_minlevel=fv["bla_minlevel"], ede=ede, dfr=fv.get("dfr"), dfi=fv.get("dfi"),
                                      sa_coef_b=fv.get("sa_coef_b"), dxf=dxf)
    _check(val, nesc, v2, n2, f"point31 tile ede={ede}")
    un = np.nonzero(gp2 >= 0)[0]
    assert un.size > 0
    order = np.argsort(uidx)
    assert np.array_equal(uidx[order], un) and np.array_equal(ugp[order], gp2[un])
    assert int(st[0]) == st2["steps"] and int(st[1]) == st2["bla_applies"] and int(st[3]) == st2["rebases"]


def test_double_sample_list_equals_grid(ctx):
    """A launch over a list of samples computes exactly what the grid launch computes
    for those samples (the later-pass mode)."""
    cre, cim, scale, W, H, mxit, R, pb = FRAMES["c1"]
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb)
    ref = ctx.upload(fv)
    try:
        n_all, v_all, _, _ = ctx.render(ref, ctx.tile(fv, form=mxhip.MX_FORM_DOUBLE))
        rng = np.random.default_rng(11)
        sel = np.sort(rng.choice(W * H, 5000, replace=False)).astype(np.int32)
        n_l, v_l, _, _ = ctx.render(ref, ctx.tile(fv, form=mxhip.MX_FORM_DOUBLE, samples=sel))
    finally:
        ref.release()
    assert np.array_equal(n_l[sel], n_all[sel]) and np.array_equal(v_l[sel], v_all[sel])
    rest = np.setdiff1d(np.arange(W * H), sel)
    assert (n_l[rest] == np.iinfo(np.int32).min).all()


def test_escalated_frame_matches_oracle(ctx):
    """A 1e200 frame whose double-path reference passes too close to zero: the floatexp
    path from the exact view centre, bit-exact against the oracle's."""
    from test_oracle_dbl import ESCALATE
    cre, cim, scale, W, H, mxit, pb = ESCALATE
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, prec_bits=pb)
    val_o, nesc_o, out_o, info_o = mxo.compute(cre, cim, scale, W, H, mxit, prec_bits=pb)
    assert info["path"] == 2 and info["escalated"] == 1 and info_o["escalated"] == 1
    assert info["ref_it"] == info_o["ref_it"] and info["ref_virtual"] == 1
    _check(val, nesc, val_o, nesc_o, "escalated")
    for k in ("steps", "bla_applies", "bla_skipped", "rebases"):
        assert info[k] == info_o[k], k


@pytest.mark.parametrize("name,ede", [("c1", False), ("ticktock", False), ("point31", True), ("c1", True)])
def test_orbit_in_lds_changes_nothing(ctx, name, ede):
    """The orbit staged in LDS (all of it for C1's 8,193 entries; the prefix for
    ticktock's longer orbit) gives the global-memory launch's results and counters bit
    for bit, glitched samples included."""
    cre, cim, scale, W, H, mxit, R, pb = FRAMES[name]
    print(f"[dbl lds] {name} ede={ede}", flush=True)
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb, c_method=2 if ede else 0)
    ref = ctx.upload(fv)
    try:
        outs = []
        for stage in (-1, 1):
            plan = ctx.plan(ref, ctx.tile(fv, form=mxhip.MX_FORM_DOUBLE, interior=True, ede=ede, lds_stage=stage))
            try:
                plan.launch()
                ms = plan.wait()
                outs.append(plan.read() + plan.unresolved() + (ms,))
            finally:
                plan.destroy()
    finally:
        ref.release()
    (n0, v0, o0, s0, u0, g0, ms0), (n1, v1, o1, s1, u1, g1, ms1) = outs
    print(f"[dbl lds] {name}: global {ms0:.3f} ms, LDS {ms1:.3f} ms (orbit {fv['len']} entries)", flush=True)
    assert np.array_equal(n0, n1) and np.array_equal(v0.view(np.int64), v1.view(np.int64))
    assert np.array_equal(s0, s1)
    assert np.array_equal(np.sort(u0), np.sort(u1))


@pytest.mark.parametrize("name,c_method", [("ticktock140", 1), ("mini106", 3), ("point31", 1)])
def test_double_supersampled_frame_matches_oracle(ctx, name, c_method):
    """Adaptive supersampling on the double path (mandel_perturbation.cpp:1789-2005):
    the same cells refined, every cell's escape count equal to the oracle's and its
    value within 1e-9; ticktock at 1e140 re-references at a sub-sample, the minibrot's
    edge is refined under EDE."""
    from test_oracle_dbl_ss import MINI106, TICKTOCK140

    frames = {"ticktock140": TICKTOCK140 + (0.0, 0), "mini106": MINI106 + (0.0, 0), "point31": FRAMES["point31"]}
    cre, cim, scale, W, H, mxit, R, pb = frames[name]
    print(f"[dbl ss] {name} {W}x{H} sub 3 c_method={c_method}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb, c_method=c_method, sub=3)
    _, _, _, oi = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=R or 1e8, prec_bits=pb or None,
                                  c_method=c_method, sub=3)
    assert info["path"] == 1 and info["ss_pixels"] == oi["ss_pixels"] > 0
    done = oi["cell_nesc"] != np.iinfo(np.int32).min
    assert np.array_equal(it != -10, done)
    _check(val[done], nesc[done], oi["cell_val"][done], oi["cell_nesc"][done], name)
    assert np.array_equal(it[done], oi["cells"][done]) or np.abs(it[done] - oi["cells"][done]).max() < 1e-3
