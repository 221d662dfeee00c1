"""Benchmark of the per-pixel perturbation path on one MI355X.

Workload (the headline configuration): a 1920x1080 frame with 3x3 supersampling
rendered as a 5760x3240 sample grid (uniform SS as a bigger frame, render.cpp:145,
186-189), view width 1e-1000 (scale 4e1000) around the seed-7 boundary-walk centre
(tests/golden/walk_de_seed7.json, tests/golden/make_walk.py --rule de), max_iter 2e6,
BLA on, exterior distance estimate output (carried dz/dc).  The reference orbit, BLA
table and dc tables are built on the host (GMP) and uploaded before any timed region;
a step is one main pass of the kernel over every sample (the engine's delta-loop
phase, mandel_perturbation.cpp:1733-1736).  The forms the metric names run on the same
resident inputs:
  fe64  plain form: IEEE FP64, the engine's operation order (escape counts identical
        to the engine's CPU path) -- `value`;
  fe32  mixed form: FP32-mantissa floatexp delta, FP64 radii / rebase / logarithm;
  sc64  the mixed form's algorithm with an FP64 delta.
The mixed forms' figures, their ratio to the plain form and their agreement with it
(whole frame) and with per-pixel MPFR iteration (committed samples) are reported
beside `value`.

Metric: nominal giga pixel-iterations per second (sum of escape counts, max_iter for
interior samples: verify.cpp:261-273) and the frame time.

Reported beside the headline (each leg can be skipped): the end-to-end frame through
mx_compute, the plain-double path's 640x360 frame at scale 4e10 with the CPU oracle on the
same frame, 96 batched uneven tiles over 12 boundary-walk locations in one persistent
launch (tests/golden/c4_requests.json), and a 240-frame 1280x720 zoom to 1e-500 replayed
as one HIP graph.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("mandelbrot-explorer_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

# thread binding of the CPU baseline's OpenMP runtime (read when it first starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

import autotune  # noqa: E402
import mxhip  # noqa: E402

# MI355X published vector peaks.  Neither form issues packed FP32 (v_pk_*) or MFMA:
#   FP64 vector 78.6 TFLOP/s counts an FMA as 2; the plain form has no FMA -> 39.3 T ops/s
#   FP32 vector 157.3 TFLOP/s (FMA as 2).  The mixed form's FMAs are unpacked, yet its
#   main pass issues 1.373e12 VALU wave instructions in 1.456 s (rocprofv3 SQ_INSTS_VALU
#   on the 5760x3240 frame): 6.0e13 lane instructions/s, above the 3.9e13 a 64-lane-per-
#   clock CU would allow, so the unpacked rate is not half the published one and the
#   published figure is the roof.
PEAK = {"fe64": 78.6e12 / 2, "fe32": 157.3e12, "sc64": 78.6e12}
# Algorithmic FP add/mul per executed step, counted from the plain kernel (mx_fe64.hip
# trip()): fused step 14, reconstruction + |z|^2 7, |w|^2 3, |S w|^2 5 -> 29; the
# carried derivative adds zm (4), J' (9) and |J|^2 (3) -> 16.  A BLA apply: the linear
# map 14, landing check 7, re-rescale 5, new d 2 -> 28; J -> A J + B adds 20.  The same
# algorithmic counts are used for every form.
F_STEP, F_STEP_DE = 29, 16
F_APPLY, F_APPLY_DE = 28, 20
FORMS = {"fe64": mxhip.MX_FORM_FE64, "fe32": mxhip.MX_FORM_FE32, "sc64": mxhip.MX_FORM_SC64}


WALK = os.path.join(ROOT, "tests", "golden", "walk_de_seed7.json")
WORKLOADS = {
    # name: (walk depth, W, H, max_iter, coloring, label)
    "c3": ("1000", 5760, 3240, 2000000, mxhip.MX_EXTERIOR_DIST_EST,
           "1920x1080 3x3 SS (5760x3240 samples) @ width 1e-1000, seed-7 walk centre, mxit 2e6, BLA, EDE"),
    "c3-1080": ("1000", 1920, 1080, 2000000, mxhip.MX_EXTERIOR_DIST_EST,
                "1920x1080 (no SS) @ width 1e-1000, seed-7 walk centre, mxit 2e6, BLA, EDE"),
    "c2": ("300", 1920, 1080, 1000000, 0, "1920x1080 @ width 1e-300, seed-7 walk centre, mxit 1e6, BLA, smooth"),
}


def load_workload(name):
    if name not in WORKLOADS:
        raise SystemExit(f"unknown workload {name}")
    depth, W, H, mxit, cm, label = WORKLOADS[name]
    c = json.load(open(WALK))[depth]
    # view width 10^-depth: scale 4 / width (mandel_perturbation.cpp:1403-1416)
    return dict(name=name, cre=c["re"], cim=c["im"], scale="4" + "0" * int(depth), W=W, H=H, mxit=mxit, c_method=cm,
                depth=int(depth),
                label=label + " (boundary walk, distance-estimate rule: tests/golden/make_walk.py --rule de)")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_form(ctx, ref, fv, form, ede, interior, center, steps, warmup, knobs=None):
    k = knobs or {}
    tile = ctx.tile(fv, stride=1, ede=ede, interior=interior, form=FORMS[form], bla_minlevel=k.get("bla_minlevel"),
                    chunk=k.get("chunk", 0), grid_blocks=k.get("blocks_per_cu", 0) * ctx.cu_count,
                    quorum=k.get("quorum", 0))
    tile.skip_sample = -1 if fv["ref_virtual"] else center
    plan = ctx.plan(ref, tile)
    try:
        for k in range(warmup):
            plan.launch()
            log(f"{form} warmup {k}: {plan.wait():.1f} ms")
        kms = []
        t0 = time.perf_counter()
        for k in range(steps):
            plan.launch()
            kms.append(plan.wait())
            log(f"{form} step {k}: {kms[-1]:.1f} ms")
        wall = time.perf_counter() - t0
        nesc, val, out, st = plan.read()
    finally:
        plan.destroy()
    if not fv["ref_virtual"]:
        nesc[center] = fv["ref_pixel_n"]
        val[center] = fv["ref_pixel_value"]
    return dict(nesc=nesc, val=val, kernel_ms=float(np.mean(kms)), ms_per_step=wall / steps * 1e3,
                stats=[int(x) for x in st])


def summary(form, r, mxit, ede):
    nominal = int(np.where(r["nesc"] > 0, r["nesc"], mxit).astype(np.int64).sum())
    steps, applies, skipped, rebases = r["stats"]
    flops = steps * (F_STEP + (F_STEP_DE if ede else 0)) + applies * (F_APPLY + (F_APPLY_DE if ede else 0))
    ach = flops / (r["kernel_ms"] / 1e3)
    return dict(ms_per_step=r["ms_per_step"], kernel_ms=r["kernel_ms"],
                gpix_iter_s=nominal / (r["ms_per_step"] / 1e3) / 1e9, nominal_pixel_iterations=nominal,
                executed_steps=steps, bla_applies=applies, bla_skipped_iterations=skipped, rebases=rebases,
                achieved_tflops=ach / 1e12, peak_tflops=PEAK[form] / 1e12, valu_frac=ach / PEAK[form])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--form", default="fe64", choices=list(FORMS), help="form reported as value")
    ap.add_argument("--compare", default="fe32,sc64", help="comma list of forms measured beside it ('' for none)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--mpfr-samples", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-double", action="store_true", help="skip the plain-double path (640x360 @ scale 4e10) leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end mx_compute frame")
    ap.add_argument("--no-c4", dest="c4", action="store_false",
                    help="skip the batched uneven-tiles leg (96 tiles, 12 references; run when its request file exists)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (chunk / grid size / BLA cutoff)")
    ap.add_argument("--knobs", default="auto", help="JSON file of launch knobs per form: read instead of tuning when "
                    "it exists, written after tuning otherwise; 'auto' (default): profiles/knobs_<workload>.json, "
                    "the autotuner's committed result that the profiles under profiles/ were taken with")
    ap.add_argument("--copy-calibration", action="store_true", help="only copy 1 GiB device to device with the "
                    "library's copy kernel and exit (the reading that scales rocprofv3's FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--no-c5", dest="c5", action="store_false", help="skip the zoom-sequence leg (hipGraph of 1280x720 frames)")
    ap.add_argument("--c5-frames", type=int, default=240, help="first N frames of the zoom sequence")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: this build runs the path on one GPU only "
              f"(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        with mxhip.Context(0) as ctx:
            copy_calibration(ctx)
        return
    wl = load_workload(args.workload)
    if args.knobs == "auto":
        args.knobs = os.path.join(ROOT, "profiles", f"knobs_{args.workload}.json")
    ede = bool(wl["c_method"] & 2)
    W, H, mxit = wl["W"], wl["H"], wl["mxit"]
    t0 = time.time()
    fv = mxhip.frame_prepare(wl["cre"], wl["cim"], wl["scale"], W, H, mxit, c_method=wl["c_method"])
    prep_s = time.time() - t0
    log(f"prepared {wl['label']}: ref_it={fv['ref_it']} levels={fv['nlev']} prec={fv['precision_bits']} "
        f"ref {fv['ref_seconds']:.2f}s bla {fv['bla_seconds']:.2f}s")
    center = (H // 2) * W + (W // 2)
    compare = [f for f in args.compare.split(",") if f and f != args.form]
    res = {}
    with mxhip.Context(0) as ctx:
        ref = ctx.upload(fv)
        try:
            # interior gate probe on the 8x8-strided grid (mandel_perturbation.cpp:1668-1700)
            pt = ctx.tile(fv, stride=8, ede=ede, interior=True, form=FORMS[args.form])
            pplan = ctx.plan(ref, pt)
            pplan.launch()
            probe_ms = pplan.wait()
            pn, _, _, _ = pplan.read()
            pplan.destroy()
            pn = pn.reshape(H, W)[::8, ::8]
            interior = bool(((pn < 0) & (pn != np.iinfo(np.int32).min)).any())
            log(f"probe {probe_ms:.1f} ms, interior detection {'on' if interior else 'off'}")
            tuned = {}
            if args.knobs and os.path.exists(args.knobs):
                tuned = json.load(open(args.knobs))
                log(f"knobs from {args.knobs}: {tuned}")
            elif not args.no_tune:
                for f in [args.form] + compare:
                    # the whole frame when one launch is cheap (the probe pass stands in)
                    whole = probe_ms * 64 < 2000.0
                    r = autotune.tune(ctx, ref, fv, FORMS[f], ede=ede, interior=interior, log=log,
                                      target_samples=None if whole else 2_000_000)
                    tuned[f] = {k: r[k] for k in ("chunk", "blocks_per_cu", "quorum", "bla_minlevel", "window", "ms")}
                    tuned[f]["trials"] = len(r["trials"])
                    log(f"autotune {f}: {tuned[f]}")
                if args.knobs:
                    json.dump(tuned, open(args.knobs, "w"), indent=1)
            res[args.form] = run_form(ctx, ref, fv, args.form, ede, interior, center, args.steps, args.warmup,
                                      tuned.get(args.form))
            for f in compare:
                res[f] = run_form(ctx, ref, fv, f, ede, interior, center, 1, 0, tuned.get(f))
        finally:
            ref.release()
        e2e_ms = None
        if not args.no_e2e:
            # one whole frame through mx_compute: reference orbit, BLA, upload, probe,
            # kernels and read-back, with the EDE reference pixel's GMP point computed on
            # a host thread beside the kernels
            t1 = time.time()
            ctx.compute(wl["cre"], wl["cim"], wl["scale"], W, H, mxit, c_method=wl["c_method"], form=FORMS[args.form])
            e2e_ms = (time.time() - t1) * 1e3
            log(f"end-to-end frame (mx_compute): {e2e_ms:.0f} ms")
        double = None if args.no_double else double_leg(ctx, args.no_cpu)
        batched = c4_leg(ctx) if args.c4 else None
        zoom = c5_leg(ctx, args.c5_frames) if args.c5 else None
    main_r = res[args.form]
    forms = {f: summary(f, r, mxit, ede) for f, r in res.items()}
    s = forms[args.form]
    accuracy = {}
    for f in compare:
        a, b = res[args.form], res[f]
        ext = (a["nesc"] > 0) & (b["nesc"] > 0)
        if ede:
            err = np.abs(a["val"][ext] - b["val"][ext]) / np.maximum(np.abs(b["val"][ext]), 1e-300)
        else:
            err = np.abs(a["val"][ext] - b["val"][ext])
        accuracy[f"{args.form}_vs_{f}"] = {"n_esc_identical_frac": float(np.mean(a["nesc"] == b["nesc"])),
                                           "class_identical_frac": float(np.mean((a["nesc"] > 0) == (b["nesc"] > 0))),
                                           "value_err_median": float(np.median(err)) if err.size else 0.0,
                                           "value_err_max": float(err.max()) if err.size else 0.0,
                                           "value_err": "relative" if ede else "absolute smooth iterations"}
        forms[f"speedup_{args.form}_vs_{f}"] = forms[f]["ms_per_step"] / s["ms_per_step"]
    fixture = mpfr_fixture(wl)
    if fixture is not None:
        accuracy["vs_mpfr"] = vs_mpfr(wl, res, fixture)
    elif args.mpfr_samples > 0:
        accuracy["vs_mpfr"] = mpfr_samples(wl, fv, res, args.mpfr_samples, center)
    cpu = None
    if not args.no_cpu:
        cpu, par = cpu_baseline(fv, res.get("fe64"), mxit, ede, interior, args.cpu_seconds, center)
        accuracy["fe64_vs_cpu_oracle_sample"] = par
    steps, applies, skipped, rebases = main_r["stats"]
    kname = "mx_fe64_kernel" if args.form == "fe64" else "mx_fe32_kernel"
    line = {
        "metric": f"Gpixel-iter/s + frame ms, 1080p @1e-{wl['depth']}: FP32-floatexp vs FP64 on GPU",
        "value": s["gpix_iter_s"], "unit": "Gpixel-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": s["ms_per_step"], "frame_ms": s["ms_per_step"] + probe_ms,
        "frame_ms_with_host_prepare": s["ms_per_step"] + probe_ms + prep_s * 1e3,
        "frame_ms_end_to_end": e2e_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.form == "fe32" else "f64",
        "data": "synthetic (seeded boundary-walk view)",
        "config": {"workload": wl["label"], "samples": W * H, "max_iter": mxit, "form": args.form,
                   "depth": f"1e-{wl['depth']}", "ede": ede, "interior_detection": interior,
                   "ref_it": fv["ref_it"], "precision_bits": fv["precision_bits"], "host_prepare_s": prep_s,
                   "host_reference_orbit_s": fv["ref_seconds"], "probe_ms": probe_ms},
        "roofline": {"bound": "valu-" + args.form, "achieved": s["achieved_tflops"], "peak": s["peak_tflops"],
                     "unit": "TFLOP/s", "frac": s["valu_frac"], "traffic": traffic_bytes(kname, wl),
                     "kernel": kname,
                     "kernel_ms": s["kernel_ms"], "executed_steps": steps, "bla_applies": applies,
                     "bla_skipped_iterations": skipped, "rebases": rebases,
                     "flops_per_step": F_STEP + (F_STEP_DE if ede else 0),
                     "flops_per_apply": F_APPLY + (F_APPLY_DE if ede else 0)},
        "forms": forms,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
        "double_path": double,
    }
    if batched is not None:
        line["batched_tiles"] = batched
    if tuned:
        line["autotune"] = tuned
    if zoom is not None:
        line["zoom_sequence"] = zoom
    print(json.dumps(line))


def copy_calibration(ctx, nbytes=1 << 30):
    """A 1 GiB device-to-device copy by the library's plain copy kernel (mx_copy_probe):
    it reads and writes nbytes each, so its rocprofv3 FETCH_SIZE / WRITE_SIZE give the
    factors that turn those counters into bytes."""
    ms = ctx.copy_probe(nbytes)
    log(f"copy calibration: {nbytes} bytes copied by mx_copy_kernel in {ms:.3f} ms ({2 * nbytes / ms / 1e6:.0f} GB/s)")
    return ms


def traffic_bytes(kernel, wl):
    """HBM bytes per launch of the main pass, from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs recorded in profiles/traffic.json (each counter scaled by what it reads for a
    copy kernel of known size); None when this workload has not been profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    ent = json.load(open(path)).get("entries", {}).get(f"{kernel}|{wl['name']}|1e-{wl['depth']}")
    return None if ent is None else ent["traffic_bytes"]


# The plain-double path's configuration (scale 1e6..1e280): 640x360 at scale 4e10
# around (-0.743643887037151, 0.131825904205330), mxit 8192, bailout 65536, 128 bits.
C1 = ("-0.743643887037151", "0.131825904205330", "40000000000", 640, 360, 8192, 65536.0, 128)


def double_leg(ctx, no_cpu, reps=5):
    """Whole frames through mx_compute (host reference + series, probe, passes) on the
    GPU, and the oracle's restatement of the same frame on the host cores."""
    cre, cim, scale, W, H, mxit, R, pb = C1
    ctx.compute(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb)  # warm-up
    tot, ker, pas, prb = [], [], [], []
    for _ in range(reps):
        it, val, nesc, info = ctx.compute(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb)
        tot.append(info["total_seconds"])
        ker.append(info["kernel_ms"] + info["probe_ms"])
        pas.append(info["kernel_ms"])
        prb.append(info["probe_ms"])
    nominal = int(info["nominal_iterations"])
    out = {"workload": "640x360 @ scale 4e10 (-0.743643887037151, 0.131825904205330), mxit 8192, bailout 65536, 128-bit",
           # kernel_ms: the interior-gate probe plus the passes; the first pass runs beside
           # the probe on a second stream, so the frame waits for the longer of the two
           "frame_ms": 1e3 * float(np.median(tot)), "kernel_ms": float(np.median(ker)),
           "probe_ms": float(np.median(prb)), "pass_ms": float(np.median(pas)),
           "gpix_iter_s_frame": nominal / float(np.median(tot)) / 1e9,
           "gpix_iter_s_kernel": nominal / (float(np.median(ker)) / 1e3) / 1e9,
           "nominal_pixel_iterations": nominal, "passes": info["passes"], "sa_it": info["sa_it"],
           "executed_steps": int(info["steps"]), "bla_applies": int(info["bla_applies"]), "rebases": int(info["rebases"]),
           # the frame's longest dependent chain: a sample that runs from the series skip to
           # max_iter, one step after the other (its steps cannot overlap)
           "longest_chain_steps": mxit - 1 - int(info["sa_it"]),
           "ns_per_chain_step": float(np.median(pas)) * 1e6 / (mxit - 1 - int(info["sa_it"]))}
    log(f"double path: frame {out['frame_ms']:.1f} ms, probe {out['probe_ms']:.2f} ms beside pass {out['pass_ms']:.2f} ms")
    if not no_cpu:
        import mxo
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1
        t0 = time.perf_counter()
        _, n_o, _, _ = mxo.compute_dbl(cre, cim, scale, W, H, mxit, escape_radius=R, prec_bits=pb, threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_oracle_frame_ms"] = dt * 1e3
        out["cpu_oracle_threads"] = threads
        out["n_esc_identical_to_oracle"] = bool(np.array_equal(n_o, nesc))
        log(f"double path: cpu oracle frame {dt * 1e3:.1f} ms on {threads} threads")
    return out


def c4_leg(ctx, forms=("fe64", "fe32"), reps=1):
    """The uneven tiles of tests/golden/c4_requests.json (96 over 12 boundary-walk
    locations; fewer when the file lists walks still missing):
    each location's reference, BLA table and dc tables are built once and resident;
    the tiles run as one persistent launch per form (largest work first), and the
    chosen form also tile by tile, one launch each, for the batching gain."""
    path = os.path.join(ROOT, "tests", "golden", "c4_requests.json")
    if not os.path.exists(path):
        log("c4: tests/golden/c4_requests.json missing (run tests/golden/make_c4.py)")
        return None
    spec = json.load(open(path))
    grid = 1024
    t0 = time.time()
    locs = {}
    # every location on the floatexp path with the centre pixel's full reference (the
    # engine's MANDEL_FE=1 MANDEL_PERIODIC=0)
    opts = mxhip.MX_OPT_FLOATEXP | mxhip.MX_OPT_NO_PERIODIC
    for l in spec["locations"]:
        mx = max(q["max_iter"] for q in spec["requests"] if q["location"] == l["seed"])
        fv = mxhip.frame_prepare(l["re"], l["im"], "4" + "0" * l["depth"], grid, grid, mx, options=opts)
        locs[l["seed"]] = (fv, ctx.upload(fv))
    prep_s = time.time() - t0
    log(f"c4: {len(locs)} references prepared in {prep_s:.1f} s")
    depths = sorted(l["depth"] for l in spec["locations"])
    out = {"workload": f"{len(spec['requests'])} square tiles (64..1024) of {len(locs)} seeded walk views (widths 1e-{depths[0]}.."
                       f"1e-{depths[-1]}, a 1024x1024 grid each), max_iter 1e4..1e7, smooth, BLA, interior "
                       f"detection (tests/golden/c4_requests.json)",
           "requests": len(spec["requests"]), "samples": int(sum(q["side"] ** 2 for q in spec["requests"])),
           "host_prepare_s": prep_s}
    res = {}
    try:
        for form in forms:
            plans, nominal_of = [], []
            for q in spec["requests"]:
                fv, ref = locs[q["location"]]
                sub = autotune.subframe(fv, q["x0"], q["y0"], q["side"], q["side"])
                t = ctx.tile(sub, form=FORMS[form], interior=True, mxit=q["max_iter"])
                t.skip_sample = sub["skip"]
                plans.append(ctx.plan(ref, t))
            order = sorted(range(len(plans)), key=lambda k: -spec["requests"][k]["side"] ** 2 * spec["requests"][k]["max_iter"])
            b = ctx.batch([plans[k] for k in order])
            b.launch()
            b.wait()  # warm-up
            t1 = time.perf_counter()
            kms = []
            for _ in range(reps):
                b.launch()
                kms.append(b.wait())
            wall = (time.perf_counter() - t1) / reps
            nominal, steps = 0, 0
            nescs = []
            for q, p in zip(spec["requests"], plans):
                nesc, _, _, st = p.read()
                nescs.append(nesc)
                nominal += int(np.where(nesc > 0, nesc, q["max_iter"]).astype(np.int64).sum())
                steps += int(st[0])
            r = {"batch_ms": wall * 1e3, "batch_kernel_ms": float(np.mean(kms)), "nominal_pixel_iterations": nominal,
                 "executed_steps": steps, "gpix_iter_s": nominal / wall / 1e9}
            log(f"c4 {form}: one launch {wall * 1e3:.0f} ms, {r['gpix_iter_s']:.0f} Gpixel-iter/s")
            if form == forms[0]:
                t1 = time.perf_counter()
                for k in order:
                    plans[k].launch()
                    plans[k].wait()
                seq = time.perf_counter() - t1
                r["tile_by_tile_ms"] = seq * 1e3
                r["batch_speedup_vs_tile_by_tile"] = seq / wall
                log(f"c4 {form}: tile by tile {seq * 1e3:.0f} ms")
            b.destroy()
            for p in plans:
                p.destroy()
            res[form] = (r, nescs)
    finally:
        for fv, ref in locs.values():
            ref.release()
    for form, (r, _) in res.items():
        out[form] = r
    if len(forms) > 1 and forms[1] in res:
        a, b_ = res[forms[0]][1], res[forms[1]][1]
        same = sum(int((x == y).sum()) for x, y in zip(a, b_))
        out[f"{forms[0]}_vs_{forms[1]}"] = {"n_esc_identical_frac": same / out["samples"],
                                           "speedup": res[forms[1]][0]["batch_ms"] / res[forms[0]][0]["batch_ms"]}
    return out


def c5_frames(n=240, width_end_exp=500, mxit0=1000, mxit1=5_000_000):
    """Frame k of the zoom: view width 4 (1e-500/4)^(k/(n-1)) (scale = 4/width, as a
    decimal integer) and max_iter 1e3 + (5e6 - 1e3) k/(n-1), linear in k."""
    from decimal import Decimal, getcontext
    getcontext().prec = width_end_exp + 40
    end = Decimal(10) ** (-width_end_exp) / 4
    out = []
    for k in range(n):
        t = Decimal(k) / Decimal(n - 1)
        scale = (Decimal(1) / end) ** t  # 4 / (4 end^t)
        out.append(dict(k=k, scale=format(scale.to_integral_value(), "f"), scale_d=scale,
                        mxit=int(round(mxit0 + (mxit1 - mxit0) * k / (n - 1)))))
    return out


def c5_leg(ctx, nframes=240, forms=("fe64", "fe32")):
    """Zoom animation: 1280x720 frames from width 4 to 1e-500 around the seed-42 walk
    centre, captured as one HIP graph per form.  Frames at scale <= 1e6 (the engine's
    shallow, non-perturbation path) are left out; every other frame runs in the
    floatexp forms against a reference at the exact view centre, recomputed only when
    the precision (in 64-bit limbs) grows and built to the largest max_iter of its
    frames, with one BLA table per reference (the shallowest frame's, valid for the
    deeper ones)."""
    seed = 42
    walk = json.load(open(os.path.join(ROOT, "tests", "golden", "walk_de_seed42.json")))
    cre, cim = walk["500"]["re"], walk["500"]["im"]
    W, H = 1280, 720
    opts = mxhip.MX_OPT_VIEW_CENTER_REF | mxhip.MX_OPT_FLOATEXP
    allf = c5_frames()[:nframes]
    # frames at scale <= 1e6 take the engine's shallow direct path (method 0): one
    # mx_compute each, every sample iterating its own c in double on the device
    shallow = {"frames": 0, "frame_ms": 0.0, "kernel_ms": 0.0, "nominal_pixel_iterations": 0}
    for f in allf:
        if f["scale_d"] > 1000000:
            continue
        ctx.compute(cre, cim, f["scale"], W, H, f["mxit"])  # warm-up
        t1 = time.perf_counter()
        _, _, _, info = ctx.compute(cre, cim, f["scale"], W, H, f["mxit"])
        shallow["frame_ms"] += (time.perf_counter() - t1) * 1e3
        shallow["kernel_ms"] += info["kernel_ms"]
        shallow["nominal_pixel_iterations"] += int(info["nominal_iterations"])
        shallow["frames"] += 1
    log(f"c5: {shallow['frames']} shallow frames, {shallow['frame_ms']:.1f} ms ({shallow['kernel_ms']:.2f} ms kernels)")
    frames = [f for f in allf if f["scale_d"] > 1000000]
    for f in frames:
        f["prec"] = int(len(f["scale"]) * 3.321928094887362) + 64
        f["limbs"] = (f["prec"] + 127) // 64
    groups = {}
    for f in frames:
        groups.setdefault(f["limbs"], []).append(f)
    t0 = time.time()
    refs, plans_by_form = [], {fm: [] for fm in forms}
    nominal_mxit = []
    try:
        for limbs, fs in sorted(groups.items()):
            prec = max(f["prec"] for f in fs)
            mx = max(f["mxit"] for f in fs)
            gv = mxhip.frame_prepare(cre, cim, fs[0]["scale"], W, H, mx, prec_bits=prec, options=opts)
            ref = ctx.upload(gv)
            refs.append(ref)
            log(f"c5: reference {len(refs)}/{len(groups)} ({len(fs)} frames, {prec} bits, ref_it {gv['ref_it']})")
            for f in fs:
                fv = mxhip.frame_prepare(cre, cim, f["scale"], W, H, 1, prec_bits=prec, options=opts, use_bla=0)
                fv = dict(fv, nlev=gv["nlev"], bla_minlevel=gv["bla_minlevel"])
                for fm in forms:
                    t = ctx.tile(fv, form=FORMS[fm], interior=True, mxit=f["mxit"])
                    plans_by_form[fm].append(ctx.plan(ref, t))
                nominal_mxit.append(f["mxit"])
        prep_s = time.time() - t0
        log(f"c5: {len(frames)} frames, {len(groups)} references, host {prep_s:.1f} s")
        out = {"workload": f"{len(allf)} frames 1280x720, width 4 -> 1e-500, seed-{seed} boundary-walk centre "
                           f"(tests/golden/walk_de_seed42.json), max_iter 1e3 -> 5e6 linear, smooth, BLA, interior "
                           f"detection; {len(frames)} perturbation frames in one HIP graph per form, "
                           f"{shallow['frames']} shallow frames (scale <= 1e6) through mx_compute",
               "frames": len(allf), "graph_frames": len(frames), "shallow": shallow, "references": len(groups),
               "host_prepare_s": prep_s}
        for fm in forms:
            plans = plans_by_form[fm]
            g = ctx.graph(plans)
            g.launch()
            log(f"c5 {fm}: warm-up replay {g.wait():.0f} ms")
            t1 = time.perf_counter()
            g.launch()
            gms = g.wait()
            wall = time.perf_counter() - t1
            nominal = 0
            for p, mx in zip(plans, nominal_mxit):
                nesc = p.read()[0]
                nominal += int(np.where(nesc > 0, nesc, mx).astype(np.int64).sum())
            r = {"graph_ms": wall * 1e3, "graph_gpu_ms": gms, "ms_per_frame": wall * 1e3 / len(plans),
                 "nominal_pixel_iterations": nominal, "gpix_iter_s": nominal / wall / 1e9,
                 "sequence_ms": wall * 1e3 + shallow["frame_ms"],
                 "sequence_gpix_iter_s": (nominal + shallow["nominal_pixel_iterations"]) /
                 (wall + shallow["frame_ms"] / 1e3) / 1e9}
            if fm == forms[0]:
                log(f"c5 {fm}: graph replay {wall * 1e3:.0f} ms; frame by frame next")
                t1 = time.perf_counter()
                for p in plans:
                    p.launch()
                    p.wait()
                r["frame_by_frame_ms"] = (time.perf_counter() - t1) * 1e3
            g.destroy()
            log(f"c5 {fm}: graph {wall * 1e3:.0f} ms ({r['ms_per_frame']:.1f} ms/frame), {r['gpix_iter_s']:.0f} Gpixel-iter/s")
            out[fm] = r
        if len(forms) > 1:
            out[f"speedup_{forms[0]}_vs_{forms[1]}"] = out[forms[1]]["graph_ms"] / out[forms[0]]["graph_ms"]
        return out
    finally:
        for fm in forms:
            for p in plans_by_form[fm]:
                p.destroy()
        for r_ in refs:
            r_.release()


def mpfr_fixture(wl):
    """The committed per-sample MPFR reference of this workload
    (tests/golden/mpfr_<name>.json, tests/golden/make_mpfr_samples.py), or None."""
    path = os.path.join(ROOT, "tests", "golden", f"mpfr_{wl['name']}.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    same = (d["W"], d["H"], d["max_iter"], d["c_method"], d["depth"], d["walk"]) == \
        (wl["W"], wl["H"], wl["mxit"], wl["c_method"], wl["depth"], os.path.basename(WALK))
    return d if same else None


def vs_mpfr(wl, res, d):
    """Every form against the committed MPFR samples: the fraction at identical escape
    count (and class), and the value error (relative for distances, absolute smooth
    iterations otherwise) over the samples both call exterior."""
    idx = np.asarray(d["idx"], np.int64)
    mn = np.asarray(d["n_esc"], np.int32)
    mv = np.asarray(d["value"], np.float64)
    ede = bool(wl["c_method"] & 2)
    out = {"samples": int(idx.size), "precision_bits": d["precision_bits"],
           "fixture": f"tests/golden/mpfr_{wl['name']}.json",
           "fixture_check": {"precision_bits": d["check_precision_bits"], "samples": d["check_samples"],
                             "n_esc_identical": d["check_n_esc_identical"]},
           "mpfr_exterior_frac": float((mn > 0).mean())}
    for f, rr in res.items():
        n_f, v_f = rr["nesc"][idx], rr["val"][idx]
        both = (n_f > 0) & (mn > 0)
        err = np.abs(v_f[both] - mv[both])
        if ede:
            err = err / np.abs(mv[both])
        out[f] = {"n_esc_identical_frac": float(np.mean(np.where(mn > 0, n_f == mn, n_f < 0))),
                  "class_identical_frac": float(np.mean((n_f > 0) == (mn > 0))),
                  "value_err_max": float(err.max()) if err.size else 0.0,
                  "value_err_median": float(np.median(err)) if err.size else 0.0,
                  "value_err": "relative" if ede else "absolute smooth iterations"}
    return out


def mpfr_samples(wl, fv, res, n, center):
    """Per-pixel MPFR iteration (accuratePointCompute semantics, multiplier early-out
    off) on random samples, compared with every form's escape counts."""
    from concurrent.futures import ThreadPoolExecutor
    from decimal import Decimal, getcontext

    import mxo

    W, H, mxit = wl["W"], wl["H"], wl["mxit"]
    getcontext().prec = wl["depth"] + 60
    cx, cy, sc = Decimal(wl["cre"]), Decimal(wl["cim"]), Decimal(wl["scale"])
    dw = 2 / sc
    dh = dw / W * H
    dx = dw / (W - 1) * 2
    dy = dh / (H - 1) * 2
    rng = np.random.default_rng(11)
    pick = [int(p) for p in rng.choice(W * H, n + 1, replace=False) if p != center][:n]
    pts = []
    for p in pick:  # decimal contexts are per thread: build the strings here
        i, j = divmod(p, W)
        pts.append((format(cx - dw + dx * j, "f"), format(cy - dh + dy * i, "f")))
    prec = fv["precision_bits"] + 64
    t = time.time()
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        out = list(ex.map(lambda pt: mxo.point(pt[0], pt[1], prec, mxit, track_d=False), pts))
    log(f"mpfr {len(pts)} samples in {time.time() - t:.1f}s")
    mn = np.array([o[1] for o in out])
    mv = np.array([o[0] for o in out])
    r = {"samples": len(pts), "precision_bits": prec}
    for f, rr in res.items():
        n_f = rr["nesc"][pick]
        r[f] = {"n_esc_identical_frac": float(np.mean(n_f == mn)),
                "class_identical_frac": float(np.mean((n_f > 0) == (mn > 0)))}
        if not (wl["c_method"] & 2):
            both = (n_f > 0) & (mn > 0)
            r[f]["smooth_err_max"] = float(np.abs(rr["val"][pick][both] - mv[both]).max()) if both.any() else 0.0
    return r


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fv, fe64, mxit, ede, interior, seconds, center):
    """The CPU path this workload takes in the engine -- per-pixel pixelRescaled with
    BLA, scalar per pixel, OpenMP over pixels (mandel_perturbation.cpp:2141-2148,
    2203-2210: the AVX2 lane loops run only with BLA off) -- as the oracle restates it,
    built with clang -O2 -mavx2 -fopenmp -ffp-contract=off (oracle/libmxo_clang.so,
    LLVM's OpenMP runtime), on every host core this process may use
    (OMP_NUM_THREADS, else all), threads bound close to cores; timed on a random
    sample of the same frame's samples.  Its outputs double as a spot check of the
    plain form (bit-exact escape counts expected)."""
    import mxo

    build = "libmxo_clang.so" if os.path.exists(os.path.join(ROOT, "oracle", "libmxo_clang.so")) else "libmxo.so"
    mxo.use_build(build)

    rng = np.random.default_rng(7)
    W, H = fv["W"], fv["H"]
    idx_all = np.arange(W * H)
    idx_all = idx_all[idx_all != center]
    orb = dict(zf=fv["zf"], zre_m=fv["zfe_m"][0::2], zre_e=fv["zfe_e"][0::2], zim_m=fv["zfe_m"][1::2], zim_e=fv["zfe_e"][1::2])
    bla = None
    if fv["nlev"]:
        b = mxo.bla_tables(fv["zf"], fv["len"] - 1, fv["dcmax"])
        bla = dict(nlev=b["nlev"], off=b["off"], size=b["size"], r2=b["r2"], l=b["l"], coef=b["coef"], rmax2=b["rmax2"])
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1

    def run(ix):
        i, j = ix // W, ix % W
        t = time.perf_counter()
        v, n, _ = mxo.pixels_fe(orb, bla, fv["dcol_m"][j], fv["dcol_e"][j], fv["drow_m"][i], fv["drow_e"][i], mxit,
                                fv["esc2"], fv["dx_m"], fv["dx_e"], ede=ede, use_interior=interior,
                                bla_minlevel=fv["bla_minlevel"], threads=threads)
        return v, n, time.perf_counter() - t

    ix = rng.choice(idx_all, 512, replace=False)
    _, _, dt = run(ix)
    n = int(min(len(idx_all), max(512, 512 * seconds / max(dt, 1e-3))))
    ix = rng.choice(idx_all, n, replace=False)
    v, ne, dt = run(ix)
    log(f"cpu oracle: {n} samples in {dt:.1f}s on {threads} threads")
    nominal = int(np.where(ne > 0, ne, mxit).astype(np.int64).sum())
    cpu = {"value": nominal / dt / 1e9, "unit": "Gpixel-iter/s", "cores": threads, "kind": "port",
           "sample": f"{n} random samples of the same frame ({dt:.1f} s, OpenMP over samples)",
           "build": ("clang -O2 -mavx2 -fopenmp -ffp-contract=off (oracle/libmxo_clang.so)" if build != "libmxo.so"
                     else "gcc -O2 -fopenmp -ffp-contract=off (oracle/libmxo.so)"),
           "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}
    mxo.use_build("libmxo.so")
    par = None
    if fe64 is not None:
        same = float((ne == fe64["nesc"][ix]).mean())
        ext = (ne > 0) & (fe64["nesc"][ix] > 0)
        rel = float((np.abs(v[ext] - fe64["val"][ix][ext]) / np.abs(v[ext])).max()) if ext.any() else 0.0
        par = {"samples": n, "n_esc_identical_frac": same, "max_rel_value_diff": rel}
    return cpu, par


if __name__ == "__main__":
    main()
