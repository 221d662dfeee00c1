"""The benchmark workloads on the GPU.

C2 (1920x1080 at width 1e-300, seed-7 boundary walk, max_iter 1e6, smooth) and C3
(5760x3240 samples at width 1e-1000, max_iter 2e6, exterior DE): each at a reduced size
against the oracle's whole frame, then at full size through size-independent checks --
random samples against the oracle's per-pixel loop on the same reference and tables,
and (C3) the committed MPFR samples; strips of the C2 view against the oracle's.
C5 (the 240-frame zoom to 1e-500 around the seed-42 walk centre): reduced frames
against the oracle, and a HIP graph of full-size frames replayed against the same
frames launched one by one.  Escape counts bit-exact throughout; values within 1e-9
relative (the device's log())."""
import json
import os

import numpy as np
import pytest

import mxhip
import mxo

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
WALK7 = json.load(open(os.path.join(GOLD, "walk_de_seed7.json")))
WALK42 = json.load(open(os.path.join(GOLD, "walk_de_seed42.json")))["500"]
REL = 1e-9
C2 = (WALK7["300"]["re"], WALK7["300"]["im"], "4" + "0" * 300, 1920, 1080, 1000000, 0)
C3 = (WALK7["1000"]["re"], WALK7["1000"]["im"], "4" + "0" * 1000, 5760, 3240, 2000000, 2)


def _check(val, nesc, oval, onesc, label):
    bad = np.argwhere(nesc != onesc)
    assert bad.size == 0, f"{label}: {len(bad)} escape counts differ, first {bad[:4].tolist()}"
    ext = onesc > 0
    rel = np.abs(val[ext] - oval[ext]) / np.maximum(np.abs(oval[ext]), 1e-300)
    assert rel.max(initial=0.0) <= REL, f"{label}: value rel diff {rel.max()}"


def _oracle_samples(fv, idx, mxit, ede):
    """The oracle's per-pixel loop on the frame's own orbit, BLA table and dc tables."""
    W = fv["W"]
    orb = dict(zf=fv["zf"], zre_m=fv["zfe_m"][0::2], zre_e=fv["zfe_e"][0::2], zim_m=fv["zfe_m"][1::2],
               zim_e=fv["zfe_e"][1::2])
    bla = None
    if fv["nlev"]:
        b = mxo.bla_tables(fv["zf"], fv["len"] - 1, fv["dcmax"])
        bla = dict(nlev=b["nlev"], off=b["off"], size=b["size"], r2=b["r2"], l=b["l"], coef=b["coef"], rmax2=b["rmax2"])
    i, j = idx // W, idx % W
    v, n, _ = mxo.pixels_fe(orb, bla, fv["dcol_m"][j], fv["dcol_e"][j], fv["drow_m"][i], fv["drow_e"][i], mxit,
                            fv["esc2"], fv["dx_m"], fv["dx_e"], ede=ede, use_interior=True,
                            bla_minlevel=fv["bla_minlevel"], period=fv.get("period", 0))
    return v, n


def _full_frame(ctx, cfg, form=mxhip.MX_FORM_FE64):
    cre, cim, scale, W, H, mxit, cm = cfg
    fv = mxhip.frame_prepare(cre, cim, scale, W, H, mxit, c_method=cm)
    ref = ctx.upload(fv)
    try:
        t = ctx.tile(fv, form=form, ede=cm == 2, interior=True)
        center = (H // 2) * W + W // 2
        if not fv["ref_virtual"]:
            t.skip_sample = center
        nesc, val, out, st = ctx.render(ref, t)
    finally:
        ref.release()
    return fv, nesc, val


@pytest.mark.parametrize("name,cfg,w,h", [("c2", C2, 96, 54), ("c3", C3, 72, 40)])
def test_reduced_frame_matches_oracle(ctx, name, cfg, w, h):
    cre, cim, scale, W, H, mxit, cm = cfg
    print(f"[workload] {name} reduced {w}x{h}", flush=True)
    it, val, nesc, info = ctx.compute(cre, cim, scale, w, h, mxit, c_method=cm, form=mxhip.MX_FORM_FE64)
    oval, onesc, _, oinfo = mxo.compute_fe(cre, cim, scale, w, h, mxit, c_method=cm)
    assert info["ref_it"] == oinfo["ref_it"] and info["period"] == oinfo["period"]
    _check(val, nesc, oval, onesc, name)


@pytest.mark.parametrize("name,cfg", [("c2", C2), ("c3", C3)])
def test_full_frame_samples_match_oracle(ctx, name, cfg):
    cre, cim, scale, W, H, mxit, cm = cfg
    print(f"[workload] {name} full {W}x{H}", flush=True)
    fv, nesc, val = _full_frame(ctx, cfg)
    rng = np.random.default_rng(3)
    center = (H // 2) * W + W // 2
    idx = rng.choice(W * H, 513, replace=False)
    idx = idx[idx != center][:512]
    v, n = _oracle_samples(fv, idx, mxit, cm == 2)
    _check(val[idx], nesc[idx], v, n, f"{name} samples")
    if name == "c3":
        # the committed MPFR samples of this frame: the plain form's agreement
        d = json.load(open(os.path.join(GOLD, "mpfr_c3.json")))
        mi, mn, mv = np.asarray(d["idx"]), np.asarray(d["n_esc"]), np.asarray(d["value"])
        same = np.where(mn > 0, nesc[mi] == mn, nesc[mi] < 0).mean()
        print(f"[workload] c3 plain form vs MPFR: {same:.4f} identical over {mi.size} samples", flush=True)
        assert d["check_n_esc_identical"]
        assert same >= 0.99


def test_strips_match_oracle(ctx):
    """Strips (full_height, row_base) of a reduced C2 view: each strip is its own frame
    of the taller grid, with its own centre-pixel reference (Mandel::Compute with
    full_h and row_base, mandel_perturbation.cpp:1396-1424), against the oracle's strip."""
    cre, cim, scale, W, H, mxit, cm = C2
    w, h = 96, 60
    for rb, rh in ((0, 20), (20, 24), (44, 16)):
        print(f"[workload] c2 strip rows {rb}..{rb + rh}", flush=True)
        its, vals, nescs, _ = ctx.compute(cre, cim, scale, w, rh, mxit, c_method=cm, full_h=h, row_base=rb)
        ov, on, _, _ = mxo.compute_fe(cre, cim, scale, w, rh, mxit, c_method=cm, full_h=h, row_base=rb)
        _check(vals, nescs, ov, on, f"strip {rb}")


def _c5_frame(k):
    from decimal import Decimal, getcontext
    getcontext().prec = 540
    end = Decimal(10) ** (-500) / 4
    sc = format(((Decimal(1) / end) ** (Decimal(k) / Decimal(239))).to_integral_value(), "f")
    return sc, int(round(1000 + (5_000_000 - 1000) * k / 239))


@pytest.mark.parametrize("k", [3, 100, 239])
def test_zoom_frame_reduced_matches_oracle(ctx, k):
    scale, mxit = _c5_frame(k)
    opts = mxhip.MX_OPT_VIEW_CENTER_REF | mxhip.MX_OPT_FLOATEXP
    print(f"[workload] c5 frame {k} reduced", flush=True)
    it, val, nesc, info = ctx.compute(WALK42["re"], WALK42["im"], scale, 64, 36, mxit, options=opts)
    oval, onesc, _, oinfo = mxo.compute_fe(WALK42["re"], WALK42["im"], scale, 64, 36, mxit, view_center=True)
    assert info["ref_virtual"] == oinfo["ref_virtual"] == 1
    _check(val, nesc, oval, onesc, f"c5 frame {k}")


def test_zoom_graph_equals_frame_by_frame(ctx):
    """Frames 120..135 of the zoom at full size, each against a reference at the exact
    view centre, captured as one HIP graph: every replay equals launching the frames
    one by one."""
    opts = mxhip.MX_OPT_VIEW_CENTER_REF | mxhip.MX_OPT_FLOATEXP
    refs, plans = [], []
    try:
        for k in range(120, 136, 3):
            scale, mxit = _c5_frame(k)
            fv = mxhip.frame_prepare(WALK42["re"], WALK42["im"], scale, 1280, 720, mxit, options=opts)
            ref = ctx.upload(fv)
            refs.append(ref)
            plans.append(ctx.plan(ref, ctx.tile(fv, form=mxhip.MX_FORM_FE64, interior=True)))
        print(f"[workload] c5 graph of {len(plans)} frames", flush=True)
        alone = []
        for p in plans:
            p.launch()
            p.wait()
            alone.append(p.read())
        g = ctx.graph(plans)
        for rep in range(3):
            g.launch()
            g.wait()
            for p, a in zip(plans, alone):
                nesc, val, out, st = p.read()
                assert np.array_equal(nesc, a[0]) and np.array_equal(val.view(np.int64), a[1].view(np.int64)), rep
            # eager work between replays
            plans[0].launch()
            plans[0].wait()
        g.destroy()
    finally:
        for p in plans:
            p.destroy()
        for r in refs:
            r.release()


C4 = os.path.join(GOLD, "c4_requests.json")


@pytest.mark.skipif(not os.path.exists(C4), reason="tests/golden/c4_requests.json not generated")
def test_batched_tiles_match_oracle(ctx):
    """C4: the 96 requests' tiles (64..1024 square, max_iter 1e4..1e7) over the 12
    walk locations, each location's reference, BLA table and dc tables built once on the
    floatexp path with the centre pixel's full reference; one persistent launch of all
    tiles (largest work first) equals the tiles launched one by one, and a 12x12 window of
    each location's first request equals the oracle's per-pixel loop on the same
    reference."""
    import autotune

    spec = json.load(open(C4))
    # a request file written before every walk finished lists the absent locations
    missing = spec.get("missing", [])
    have = {l["seed"] for l in spec["locations"]}
    assert len(have) == 12 - len(missing) and not have & set(missing)
    assert len(spec["requests"]) == 8 * len(have) and all(q["location"] in have for q in spec["requests"])
    opts = mxhip.MX_OPT_FLOATEXP | mxhip.MX_OPT_NO_PERIODIC
    locs, refs = {}, []
    plans, subs = [], []
    try:
        for l in spec["locations"]:
            mx = max(q["max_iter"] for q in spec["requests"] if q["location"] == l["seed"])
            print(f"[c4] location {l['seed']} depth {l['depth']} max_iter {mx}", flush=True)
            fv = mxhip.frame_prepare(l["re"], l["im"], "4" + "0" * l["depth"], 1024, 1024, mx, options=opts)
            assert fv["path"] == 2
            locs[l["seed"]] = (fv, ctx.upload(fv))
            refs.append(locs[l["seed"]][1])
        for q in spec["requests"]:
            fv, ref = locs[q["location"]]
            sub = autotune.subframe(fv, q["x0"], q["y0"], q["side"], q["side"])
            t = ctx.tile(sub, form=mxhip.MX_FORM_FE64, interior=True, mxit=q["max_iter"])
            t.skip_sample = sub["skip"]
            plans.append(ctx.plan(ref, t))
            subs.append(sub)
        order = sorted(range(len(plans)), key=lambda k: -spec["requests"][k]["side"] ** 2 * spec["requests"][k]["max_iter"])
        b = ctx.batch([plans[k] for k in order])
        print("[c4] batch launch", flush=True)
        b.launch()
        b.wait()
        batched = [p.read()[:2] for p in plans]
        b.destroy()
        for k, p in enumerate(plans):
            p.launch()
            p.wait()
            nesc, val, _, _ = p.read()
            assert np.array_equal(nesc, batched[k][0]) and np.array_equal(val, batched[k][1]), k
        seen = set()
        for k, q in enumerate(spec["requests"]):
            if q["location"] in seen:
                continue
            seen.add(q["location"])
            fv = locs[q["location"]][0]
            side = q["side"]
            w = 12
            x0, y0 = (side - w) // 2, (side - w) // 3
            idx = np.array([(y0 + i) * side + x0 + j for i in range(w) for j in range(w)])
            idx = idx[idx != subs[k]["skip"]]
            ov, on = _oracle_samples(dict(subs[k], W=side), idx, q["max_iter"], False)
            _check(batched[k][1][idx], batched[k][0][idx], ov, on, f"c4 request {k}")
    finally:
        for p in plans:
            p.destroy()
        for r in refs:
            r.release()
