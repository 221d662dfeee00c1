"""Benchmark of the per-pixel perturbation path on one MI355X.

Workload (the headline configuration): a 1920x1080 frame with 3x3 supersampling
rendered as a 5760x3240 sample grid (uniform SS as a bigger frame, render.cpp:145,
186-189), view width 1e-1000 (scale 4e1000) around the seed-7 boundary-walk centre
(tests/golden/walk_seed7.json), max_iter 2e6, BLA on, exterior distance estimate
output (carried dz/dc).  Inputs (reference orbit, BLA table, dc tables) are resident
in HBM before the timed region; a step is one main pass of the kernel over every
sample (the engine's delta-loop phase, mandel_perturbation.cpp:1733-1736).

Metric: nominal giga pixel-iterations per second (sum of escape counts, max_iter for
interior samples: verify.cpp:261-273) and the frame time.

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

import numpy as np  # noqa: E402

import mxhip  # noqa: E402

# MI355X published peaks
PEAK_FP64_VALU_OPS = 78.6e12 / 2  # vector FP64 78.6 TFLOP/s counts an FMA as 2; this kernel issues no FMA
PEAK_HBM = 8.0e12
# FP64 add/mul per executed step, counted from the kernel source (mx_fe64.hip trip()):
# fused step 14, reconstruction + |z|^2 7, |w|^2 3, |S w|^2 5 -> 29; the carried
# derivative adds zm (4), J' (9) and |J|^2 (3) -> 16.  A BLA apply: the linear map 14,
# landing check 7, re-rescale 5, new d 2 -> 28; with the derivative J -> A J + B adds 20.
F_STEP, F_STEP_DE = 29, 16
F_APPLY, F_APPLY_DE = 28, 20


def load_workload(name):
    path = os.path.join(ROOT, "tests", "golden", "walk_seed7.json")
    walk = json.load(open(path)) if os.path.exists(path) else {}
    if name in ("c3", "c2") and not [k for k in walk if not k.startswith("_")]:
        name = "misi1000"
    if name == "misi1000":
        # verify.cpp:11805's exact Misiurewicz point c = i at the same depth and size
        return dict(name=name, cre="0", cim="1", scale="4" + "0" * 1000, W=5760, H=3240, mxit=2000000,
                    c_method=mxhip.MX_EXTERIOR_DIST_EST, depth=1000,
                    label="1920x1080 3x3 SS (5760x3240 samples) @ width 1e-1000 around c=i (Misiurewicz), mxit 2e6, BLA, EDE")
    if name == "c3":
        depth = "1000"
        W, H, mxit, cm = 5760, 3240, 2000000, mxhip.MX_EXTERIOR_DIST_EST
        label = "1920x1080 3x3 SS (5760x3240 samples) @ width 1e-1000, seed-7 walk centre, mxit 2e6, BLA, EDE"
    elif name == "c2":
        depth = "300"
        W, H, mxit, cm = 1920, 1080, 1000000, 0
        label = "1920x1080 @ width 1e-300, seed-7 walk centre, mxit 1e6, BLA, smooth"
    else:
        raise SystemExit(f"unknown workload {name}")
    if depth not in walk:
        avail = sorted(int(k) for k in walk if not k.startswith("_"))
        depth = str(avail[-1])
        label = label.replace("1e-1000", f"1e-{depth}").replace("1e-300", f"1e-{depth}") + " (deeper walk centre not yet generated)"
    c = walk[depth]
    scale = "4" + "0" * int(depth)
    return dict(name=name, cre=c["re"], cim=c["im"], scale=scale, W=W, H=H, mxit=mxit, c_method=cm, depth=int(depth), label=label)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--form", default="fe64", choices=["fe64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: this build runs the path on one GPU only "
              f"(multi-GPU sharding is design-only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    wl = load_workload(args.workload)
    ede = bool(wl["c_method"] & 2)
    t0 = time.time()
    fv = mxhip.frame_prepare(wl["cre"], wl["cim"], wl["scale"], wl["W"], wl["H"], wl["mxit"], c_method=wl["c_method"])
    prep_s = time.time() - t0
    print(f"[bench] prepared {wl['label']}: ref_it={fv['ref_it']} levels={fv['nlev']} prec={fv['precision_bits']} "
          f"ref {fv['ref_seconds']:.2f}s bla {fv['bla_seconds']:.2f}s", file=sys.stderr, flush=True)
    W, H, mxit = wl["W"], wl["H"], wl["mxit"]
    with mxhip.Context(0) as ctx:
        ref = ctx.upload(fv)
        # interior gate probe on the 8x8-strided grid (mandel_perturbation.cpp:1668-1700)
        pt = ctx.tile(fv, stride=8, ede=ede, interior=True)
        pplan = ctx.plan(ref, pt)
        pplan.launch()
        probe_ms = pplan.wait()
        pn, _, _, _ = pplan.read()
        pplan.destroy()
        pn = pn.reshape(H, W)[::8, ::8]
        interior = bool(((pn < 0) & (pn != np.iinfo(np.int32).min)).any())
        tile = ctx.tile(fv, stride=1, ede=ede, interior=interior)
        center = (H // 2) * W + (W // 2)
        tile.skip_sample = -1 if fv["ref_virtual"] else center
        plan = ctx.plan(ref, tile)
        print(f"[bench] probe {probe_ms:.1f} ms, interior detection {'on' if interior else 'off'}", file=sys.stderr, flush=True)
        for k in range(args.warmup):
            plan.launch()
            ms = plan.wait()
            print(f"[bench] warmup {k}: {ms:.1f} ms", file=sys.stderr, flush=True)
        kms = []
        t_start = time.perf_counter()
        for k in range(args.steps):
            plan.launch()
            kms.append(plan.wait())
            print(f"[bench] step {k}: {kms[-1]:.1f} ms", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t_start
        nesc, val, out, st = plan.read()
        plan.destroy()
        ref.release()
    if not fv["ref_virtual"]:
        nesc[center] = fv["ref_pixel_n"]
        val[center] = fv["ref_pixel_value"]
    nominal = int(np.where(nesc > 0, nesc, mxit).astype(np.int64).sum())
    ms_step = wall / args.steps * 1e3
    kernel_ms = float(np.mean(kms))
    value = nominal / (ms_step / 1e3) / 1e9
    steps, applies, skipped, rebases = (int(x) for x in st)
    flops = steps * (F_STEP + (F_STEP_DE if ede else 0)) + applies * (F_APPLY + (F_APPLY_DE if ede else 0))
    achieved = flops / (kernel_ms / 1e3)
    roof = {"bound": "valu-fp64", "achieved": achieved / 1e12, "peak": PEAK_FP64_VALU_OPS / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_VALU_OPS, "traffic": None,
            "kernel": "mx_fe64_kernel", "kernel_ms": kernel_ms, "executed_steps": steps, "bla_applies": applies,
            "bla_skipped_iterations": skipped, "rebases": rebases, "flops_per_step": F_STEP + (F_STEP_DE if ede else 0),
            "flops_per_apply": F_APPLY + (F_APPLY_DE if ede else 0)}
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(fv, nesc, val, mxit, ede, interior, args.cpu_seconds, center)
    interior_px = int((nesc < 0).sum())
    line = {
        "metric": "Gpixel-iter/s + frame ms, 1080p @1e-1000: FP32-floatexp vs FP64 on GPU",
        "value": value, "unit": "Gpixel-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "frame_ms": ms_step + probe_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded boundary-walk view)",
        "config": {"workload": wl["label"], "samples": W * H, "max_iter": mxit, "form": args.form,
                   "depth": f"1e-{wl['depth']}", "ede": ede, "interior_detection": interior,
                   "ref_it": fv["ref_it"], "precision_bits": fv["precision_bits"], "host_prepare_s": prep_s,
                   "probe_ms": probe_ms, "nominal_pixel_iterations": nominal, "interior_samples": interior_px,
                   "kernel_ms": kernel_ms},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity_sample": parity,
    }
    print(json.dumps(line))


def cpu_baseline(fv, nesc_g, val_g, mxit, ede, interior, seconds, center):
    """The oracle (the engine's per-pixel loop restated in C, OpenMP over samples) on
    the host cores, on a random sample of the same frame's samples, timed; its
    outputs double as a parity spot check of the GPU run."""
    import mxo

    rng = np.random.default_rng(7)
    W, H = fv["W"], fv["H"]
    idx_all = np.arange(W * H)
    idx_all = idx_all[idx_all != center]
    orb = dict(zf=fv["zf"], zre_m=fv["zfe_m"][0::2], zre_e=fv["zfe_e"][0::2], zim_m=fv["zfe_m"][1::2], zim_e=fv["zfe_e"][1::2])
    bla = None
    if fv["nlev"]:
        b = mxo.bla_tables(fv["zf"], fv["len"], fv["dcmax"])
        bla = dict(nlev=b["nlev"], off=b["off"], size=b["size"], r2=b["r2"], l=b["l"], coef=b["coef"], rmax2=b["rmax2"])
    threads = min(os.cpu_count() or 1, 16)

    def run(ix):
        i, j = ix // W, ix % W
        t = time.perf_counter()
        v, n, _ = mxo.pixels_fe(orb, bla, fv["dcol_m"][j], fv["dcol_e"][j], fv["drow_m"][i], fv["drow_e"][i], mxit,
                                fv["esc2"], fv["dx_m"], fv["dx_e"], ede=ede, use_interior=interior,
                                bla_minlevel=fv["bla_minlevel"], threads=threads)
        return v, n, time.perf_counter() - t

    ix = rng.choice(idx_all, 2048, replace=False)
    _, _, dt = run(ix)
    n = int(min(len(idx_all), max(2048, 2048 * seconds / max(dt, 1e-3))))
    ix = rng.choice(idx_all, n, replace=False)
    v, ne, dt = run(ix)
    nominal = int(np.where(ne > 0, ne, mxit).astype(np.int64).sum())
    same = float((ne == nesc_g[ix]).mean())
    ext = (ne > 0) & (nesc_g[ix] > 0)
    rel = float((np.abs(v[ext] - val_g[ix][ext]) / np.abs(v[ext])).max()) if ext.any() else 0.0
    cpu = {"value": nominal / dt / 1e9, "unit": "Gpixel-iter/s", "cores": threads, "kind": "port",
           "sample": f"{n} random samples of the same frame ({dt:.1f} s, OpenMP over samples)"}
    parity = {"samples": n, "n_esc_identical_frac": same, "max_rel_value_diff": rel}
    return cpu, parity


if __name__ == "__main__":
    main()
