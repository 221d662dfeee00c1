"""Launch-parameter autotuner for the floatexp kernels.

Three knobs, timed on the device on a strided sample of the frame to be rendered:
  chunk        lean steps per trip between the full step's checks (mx_tile.chunk);
  grid_blocks  workgroups of 4 waves in the persistent grid, i.e. waves per CU;
  bla_minlevel the lowest BLA level tried (the BLA level cutoff), tuned only for the
               mixed FP32 form: it changes the rounding path, so the plain form keeps
               the engine's cutoff (minimum skip 8, mandel_perturbation.cpp:1520-1530).
chunk and grid_blocks do not change any result.

The search is coordinate descent from the defaults (chunk 24, 8 blocks per CU, the
frame's own cutoff): each knob in turn is swept with the others fixed and the fastest
value kept.  A sample is the frame's sample grid taken every `stride` rows and columns
(about 2 M samples: enough work to fill the grid, a fraction of the frame's cost).
Results can be kept in a JSON file keyed by form, flags and CU count.
"""
import json
import os

import mxhip

CHUNKS = (8, 16, 24, 32, 48, 64)
BLOCKS_PER_CU = (2, 3, 4, 5, 6, 8)
MINLEVELS = (1, 2, 3, 4, 5, 6)


def _time(ctx, ref, fv, form, ede, interior, stride, mxit, knobs, reps):
    t = ctx.tile(fv, stride=stride, form=form, ede=ede, interior=interior, mxit=mxit,
                 bla_minlevel=knobs["bla_minlevel"], chunk=knobs["chunk"],
                 grid_blocks=knobs["blocks_per_cu"] * ctx.cu_count)
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


def tune(ctx, ref, fv, form, ede=False, interior=True, mxit=None, target_samples=2_000_000, reps=1, log=None):
    """Best knobs for rendering frame `fv` (a frame_prepare dict whose reference is
    uploaded as `ref`) in `form`.  Returns dict(chunk, blocks_per_cu, bla_minlevel,
    stride, ms, trials)."""
    n = fv["W"] * fv["H"]
    stride = 1
    while n // ((stride + 1) ** 2) >= target_samples:
        stride += 1
    knobs = dict(chunk=24, blocks_per_cu=8, bla_minlevel=fv["bla_minlevel"])
    trials = []

    def run(k):
        ms = _time(ctx, ref, fv, form, ede, interior, stride, mxit, k, reps)
        trials.append(dict(k, ms=ms))
        if log:
            log(f"autotune {k} -> {ms:.1f} ms")
        return ms

    best = run(knobs)
    axes = [("chunk", CHUNKS), ("blocks_per_cu", BLOCKS_PER_CU)]
    if form == mxhip.MX_FORM_FE32 and fv["nlev"] > 0:
        axes.append(("bla_minlevel", tuple(m for m in MINLEVELS if m < fv["nlev"])))
    for name, values in axes:
        for v in values:
            if v == knobs[name]:
                continue
            k = dict(knobs, **{name: v})
            ms = run(k)
            if ms < best:
                best, knobs = ms, k
    return dict(knobs, stride=stride, ms=best, trials=trials)


def key(form, ede, interior, cu_count):
    return f"form={form}|ede={int(ede)}|interior={int(interior)}|cu={cu_count}"


def load(path, form, ede, interior, cu_count):
    """Cached knobs, or None."""
    if not os.path.exists(path):
        return None
    return json.load(open(path)).get(key(form, ede, interior, cu_count))


def store(path, form, ede, interior, cu_count, result):
    d = json.load(open(path)) if os.path.exists(path) else {}
    d[key(form, ede, interior, cu_count)] = {k: result[k] for k in ("chunk", "blocks_per_cu", "bla_minlevel")}
    json.dump(d, open(path, "w"), indent=1)
