"""Launch-parameter autotuner for the floatexp kernels.

Knobs, timed on the device on a window of the frame to be rendered:
  chunk        lean steps per trip between the full step's checks (mx_tile.chunk);
  grid_blocks  workgroups of 4 waves in the persistent grid, i.e. waves per CU;
  quorum       a wave ends a lean run once fewer than quorum/64 of its lanes still step;
  bla_minlevel the lowest BLA level tried (the BLA level cutoff), only with
               tune_cutoff and only for the mixed FP32 form.  Unlike the other two it
               changes results (another sequence of skips and roundings), so by
               default every form keeps the engine's cutoff (minimum skip 8,
               mandel_perturbation.cpp:1520-1530) and the forms stay comparable.
chunk, grid_blocks and quorum do not change any result.

The search is coordinate descent from the defaults (chunk 32, 8 blocks per CU, quorum
32, the frame's own cutoff): each knob in turn is swept with the others fixed and the
fastest value kept, over two rounds.  The sample is the whole frame when a launch is
cheap (the caller's choice), else the central window of the frame with about 2 M
samples (the frame's aspect ratio): enough work to fill the grid, a fraction of the
frame's cost, and neighbouring samples side by side as in the frame, so a wave's lanes
diverge as they do there (a strided subsample puts distant samples in one wave and
runs several times slower per sample).  Results can be kept in a JSON file keyed by
form, flags and CU count.
"""
import json
import os

import mxhip

CHUNKS = (8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512)
QUORUMS = (1, 8, 16, 24, 32, 40, 48, 56)
BLOCKS_PER_CU = (2, 3, 4, 5, 6, 8)
MINLEVELS = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10)


def window(fv, target_samples):
    """The central window of frame `fv` holding about target_samples samples."""
    W, H = fv["W"], fv["H"]
    f = min(1.0, (target_samples / float(W * H)) ** 0.5)
    w, h = max(8, int(W * f)) // 8 * 8 or W, max(8, int(H * f)) // 8 * 8 or H
    w, h = min(w, W), min(h, H)
    x0, y0 = (W - w) // 2, (H - h) // 2
    return subframe(fv, x0, y0, w, h)


def subframe(fv, x0, y0, w, h):
    """The w x h block of frame `fv` at column x0, row y0, with the index of the frame's
    reference pixel inside it ("skip", -1 when outside or virtual): that sample's value
    comes from the reference orbit, and its delta iteration (dc = 0) would run the
    whole orbit one step at a time."""
    sub = dict(fv, W=w, H=h, dcol_m=fv["dcol_m"][x0:x0 + w], dcol_e=fv["dcol_e"][x0:x0 + w],
               drow_m=fv["drow_m"][y0:y0 + h], drow_e=fv["drow_e"][y0:y0 + h], skip=-1)
    if not fv.get("ref_virtual", 1):
        ri, rj = int(fv["ref_y"]) - y0, int(fv["ref_x"]) - x0
        if 0 <= ri < h and 0 <= rj < w:
            sub["skip"] = ri * w + rj
    return sub


def _time(ctx, ref, fv, form, ede, interior, mxit, knobs, reps):
    t = ctx.tile(fv, form=form, ede=ede, interior=interior, mxit=mxit,
                 bla_minlevel=knobs["bla_minlevel"], chunk=knobs["chunk"],
                 grid_blocks=knobs["blocks_per_cu"] * ctx.cu_count, quorum=knobs.get("quorum", 0))
    t.skip_sample = fv["skip"]
    p = ctx.plan(ref, t)
    try:
        best = None
        for _ in range(reps):
            p.launch()
            ms = p.wait()
            best = ms if best is None else min(best, ms)
        return best
    finally:
        p.destroy()


def tune(ctx, ref, fv, form, ede=False, interior=True, mxit=None, target_samples=2_000_000, reps=1, log=None,
         tune_cutoff=False, passes=2):
    """Best knobs for rendering frame `fv` (a frame_prepare dict whose reference is
    uploaded as `ref`) in `form`; target_samples=None times the whole frame.  `passes`
    rounds of coordinate descent (the knobs interact: a longer lean run wants another
    quorum).  Returns dict(chunk, blocks_per_cu, quorum, bla_minlevel, window, ms,
    trials)."""
    sub = window(fv, target_samples) if target_samples else subframe(fv, 0, 0, fv["W"], fv["H"])
    knobs = dict(chunk=32, blocks_per_cu=8, quorum=32, bla_minlevel=fv["bla_minlevel"])
    trials = []

    def run(k):
        ms = _time(ctx, ref, sub, form, ede, interior, mxit, k, reps)
        trials.append(dict(k, ms=ms))
        if log:
            log(f"autotune {k} -> {ms:.1f} ms")
        return ms

    best = run(knobs)
    axes = [("quorum", QUORUMS), ("chunk", CHUNKS), ("blocks_per_cu", BLOCKS_PER_CU)]
    if tune_cutoff and form == mxhip.MX_FORM_FE32 and fv["nlev"] > 0:
        axes.append(("bla_minlevel", tuple(m for m in MINLEVELS if m < fv["nlev"])))
    for _ in range(passes):
        for name, values in axes:
            for v in values:
                if v == knobs[name]:
                    continue
                k = dict(knobs, **{name: v})
                if any(all(t[a] == k[a] for a in k) for t in trials):
                    continue
                ms = run(k)
                if ms < best:
                    best, knobs = ms, k
    return dict(knobs, window=[sub["W"], sub["H"]], ms=best, trials=trials)


def key(form, ede, interior, cu_count):
    return f"form={form}|ede={int(ede)}|interior={int(interior)}|cu={cu_count}"


def load(path, form, ede, interior, cu_count):
    """Cached knobs, or None."""
    if not os.path.exists(path):
        return None
    return json.load(open(path)).get(key(form, ede, interior, cu_count))


def store(path, form, ede, interior, cu_count, result):
    d = json.load(open(path)) if os.path.exists(path) else {}
    d[key(form, ede, interior, cu_count)] = {k: result[k] for k in ("chunk", "blocks_per_cu", "quorum", "bla_minlevel")}
    json.dump(d, open(path, "w"), indent=1)
