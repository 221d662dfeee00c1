"""Seeded boundary walks: generate the deep view centres used by the benchmark
configurations and the large-frame tests; write tests/golden/walk_<rule>seed<S>.json
("walk_seed<S>.json" for the doubling rule).

Common to every rule: start c = -0.75 + 0i, view width w = 3; seed s drives
splitmix64; each step renders a 64x64 probe of the view (square, the engine's grid:
scale = 4/w) with the oracle (full reference from the centre pixel), picks one probe
pixel uniformly (splitmix64 mod k) among k candidates, recentres on it and halves w;
MPFR precision of each probe = ceil(-log2 w) + 64 bits.  Centres are recorded at every
requested depth; the decimal strings carry ceil(-log10 w) + 30 digits.  Probes at
scale <= 1e6 (brute-force MPFR) mark the pixels inside the main cardioid or the
period-2 disc as interior without iterating them: no rule ranks a non-escaping pixel by
anything but "did not escape", so the walks are the same as without the shortcut
(tests/test_oracle_walk_probe.py; C4's seed-6 location regenerated byte for byte).

--rule spec (the workload definition as written): probe max_iter 1024; every pixel is a
  candidate, with count = its escape iteration, or 1024 when it did not escape;
  ranked by (count desc, index asc); k = ceil(0.1 * 4096) = 410.
  Below widths of about 2^-1024 the pixels counted 1024 are those beyond the
  escape-time-1024 level curve, a smooth curve at that scale: the walk follows it and
  its 1e-1000 view is two flat regions (escape at 1023 / 1024), not the set's boundary.
--rule rising: the spec rule with probe max_iter 1024 + 4 * step, so that the level
  curve the walk follows stays within a vanishing fraction of a view width of the set
  (the escape time near the boundary grows by about one iteration per halving):
  the views stay boundary views at every depth.
--rule de (walk_de_seed<S>.json; the benchmark views): the probe renders the exterior
  distance estimate; candidates are the escaping probe pixels ranked by their distance
  estimate in pixels, closest to the set first (index order among equals),
  k = ceil(0.1 * n_escaped); the probe max_iter (from 1024) doubles, up to 2^21, until
  at least 90 % of the probe pixels escape.  The views stay on the set's boundary at
  every depth without following nested minibrots.
--rule doubling (the first rule, walk_seed<S>.json): candidates are the escaping probe
  pixels; the probe mxit doubles whenever the highest escape count exceeds half of it
  (up to 2^21); k = ceil(0.1 * n_escaped); a probe with no escaping pixel at the cap
  steps back to the previous view and draws again.  It follows the highest escape
  counts into nested minibrots: by 1e-600 every probe pixel exceeds 2^21 iterations.
A resumed walk restores the generator by the recorded draw count.

Usage: python tests/golden/make_walk.py SEED DEPTH10 [DEPTH10 ...] [--rule R] [--resume DEPTH10] [--gpu] [--out DIR]
e.g.   python tests/golden/make_walk.py 7 300 1000 --rule rising
--resume continues from the centre recorded at that depth (width 3 / 2^steps).
--gpu renders the probes above scale 1e6 with the product's plain form on the GPU
(same escape counts as the oracle, so the same walk); --out writes the JSON there
instead of beside this script.
"""
import json
import math
import os
import sys
import time
from decimal import Decimal, getcontext

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import mxo  # noqa: E402

MASK = (1 << 64) - 1
MXIT_CAP = 1 << 21
THREADS = int(os.environ.get("WALK_THREADS", os.cpu_count() or 1))


class SplitMix64:
    def __init__(self, seed):
        self.x = seed & MASK

    def next(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & MASK
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)


class GpuProbe:
    """Probes above scale 1e6 rendered by the product's plain FP64 form on the GPU with
    the oracle's frame semantics (floatexp path at every depth, the exact view centre
    as reference at scale <= 1e280, the centre pixel with its fallback above): the
    plain form's escape counts equal the oracle's bit for bit, so the walk is the
    oracle-driven one, at GPU speed."""

    def __init__(self):
        sys.path.insert(0, os.path.join(HERE, "..", "..", "mandelbrot-explorer_amd"))
        import mxhip
        self.mx = mxhip
        self.ctx = mxhip.Context(0)

    def __call__(self, cre, cim, scale, W, H, mxit, prec_bits, scale_small):
        opts = self.mx.MX_OPT_FLOATEXP | (self.mx.MX_OPT_VIEW_CENTER_REF if scale_small else 0)
        _, val, nesc, _ = self.ctx.compute(cre, cim, scale, W, H, mxit, prec_bits=prec_bits, form=self.mx.MX_FORM_FE64,
                                           options=opts)
        return val, nesc


def probe_mxit(rule, step):
    return 1024 + 4 * step if rule == "rising" else 1024


def walk(seed, depths, log=print, on_depth=None, resume=None, mxit_cap=MXIT_CAP, gpu=None, rule="doubling"):
    depths = sorted(depths)
    rng = SplitMix64(seed)
    getcontext().prec = max(depths) + 80
    cx, cy = Decimal("-0.75"), Decimal(0)
    w = Decimal(3)
    mxit = 1024
    out = {}
    step = 0
    draws = 0
    if resume is not None:
        cx, cy = Decimal(resume["re"]), Decimal(resume["im"])
        step = int(resume["steps"])
        w = Decimal(3) / (Decimal(2) ** step)
        mxit = int(resume["probe_mxit"])
        # the generator's position: the recorded draw count (older records: one per step)
        draws = int(resume.get("draws", step))
        for _ in range(draws):
            rng.next()
    prev = None  # (cx, cy, w, cand, k) of the previous view, for a step back
    redraws = 0
    t0 = time.time()
    while depths:
        target = Decimal(10) ** (-depths[0])
        if w <= target:
            digits = depths[0] + 30
            out[str(depths[0])] = dict(re=format(cx.quantize(Decimal(10) ** -digits), "f"),
                                       im=format(cy.quantize(Decimal(10) ** -digits), "f"),
                                       width=format(w, ".6e"), steps=step, probe_mxit=mxit, draws=draws)
            log(f"seed {seed}: reached 1e-{depths[0]} after {step} steps ({time.time() - t0:.0f}s)")
            if on_depth:
                on_depth(out)
            depths.pop(0)
            continue
        scale = Decimal(4) / w
        # the engine reads the scale as a decimal integer magnitude; keep 30 significant digits
        sc_int = format(scale.to_integral_value(), "f") if scale >= 1 else None
        prec = int(math.ceil(-math.log2(float(w)) if w > Decimal("1e-300") else (-w.log10() * Decimal(math.log2(10))))) + 64
        prec = max(prec, 80)
        W = H = 64
        if rule != "doubling":
            mxit = probe_mxit(rule, step)
        while True:
            if sc_int is not None and scale > 1e6:
                if gpu is not None:
                    val, nesc = gpu(format(cx, "f"), format(cy, "f"), sc_int, W, H, mxit, prec, scale <= Decimal("1e280"))
                else:
                    val, nesc, _, _ = mxo.compute_fe(format(cx, "f"), format(cy, "f"), sc_int, W, H, mxit, prec_bits=prec,
                                                     threads=THREADS, periodic=rule == "doubling",
                                                     c_method=2 if rule == "de" else 0)
            else:
                # pixels inside the main cardioid or the period-2 disc cannot escape: the
                # probe marks them without iterating (every escaping pixel is unchanged)
                val, nesc = mxo.direct(format(cx, "f"), format(cy, "f"), format(scale, "f"), W, H, mxit, prec_bits=prec, wp_bits=prec, threads=THREADS,
                                       c_method=2 if rule == "de" else 0, skip_bulbs=True)
            flat = nesc.ravel()
            if rule == "de":
                # escaping pixels ranked by their distance estimate (pixels), closest first
                fv = val.ravel()
                cand = [(float(fv[k]), k) for k in range(W * H) if flat[k] > 0]
                if 10 * len(cand) >= 9 * W * H or mxit >= mxit_cap:
                    break
                mxit *= 2
                continue
            if rule != "doubling":
                # every pixel, the ones that did not escape counted as the probe's mxit
                cand = [(-(int(flat[k]) if flat[k] > 0 else mxit), k) for k in range(W * H)]
                break
            cand = [(-int(flat[k]), k) for k in range(W * H) if flat[k] > 0]
            top = -min(cand)[0] if cand else 0
            if (len(cand) >= 41 and top <= mxit // 2) or mxit >= mxit_cap:
                break
            mxit *= 2
        if not cand:
            if prev is None:
                raise RuntimeError("walk stalled: no escaping pixel")
            # step back: the previous view, another draw among its candidates
            redraws += 1
            if redraws > 64:
                raise RuntimeError("walk stalled: 64 redraws without an escaping probe")
            cx, cy, w, cand, k = prev
            scale = Decimal(4) / w
            step -= 1
            log(f"seed {seed}: step {step + 1} has no escaping pixel; redrawing at step {step}")
        else:
            redraws = 0
            cand.sort()
            k = int(math.ceil(0.1 * len(cand)))
        prev = (cx, cy, w, cand, k)
        pick = cand[rng.next() % k][1]
        draws += 1
        i, j = divmod(pick, W)
        dw = Decimal(2) / scale
        dx = dw / (W - 1) * 2
        dh = dw / W * H
        dy = dh / (H - 1) * 2
        cx = cx - dw + dx * j
        cy = cy - dh + dy * i
        w = w / 2
        step += 1
        if step % 50 == 0:
            log(f"seed {seed}: step {step} width {float(w.log10()):.1f} probe_mxit {mxit} ({time.time() - t0:.0f}s)")
    return out


def walk_file(seed, rule):
    return f"walk_seed{seed}.json" if rule == "doubling" else f"walk_{rule}_seed{seed}.json"


def option(args, name, default=None):
    """value of `name VALUE` in args (removed from args)"""
    if name not in args:
        return default
    i = args.index(name)
    v = args[i + 1]
    del args[i:i + 2]
    return v


def main():
    args = sys.argv[1:]
    gpu = GpuProbe() if "--gpu" in args else None
    if gpu is not None:
        args.remove("--gpu")
    resume_depth = option(args, "--resume")
    out_dir = option(args, "--out", HERE)
    rule = option(args, "--rule", "doubling")
    if rule not in ("doubling", "spec", "rising", "de"):
        raise SystemExit(f"unknown rule {rule}")
    seed = int(args[0])
    depths = [int(x) for x in args[1:]]
    path = os.path.join(out_dir, walk_file(seed, rule))
    resume = json.load(open(path))[resume_depth] if resume_depth else None

    def save(res):
        old = json.load(open(path)) if os.path.exists(path) else {}
        old.update(res)
        old["_definition"] = f"tests/golden/make_walk.py --rule {rule} (seeded boundary walk)"
        json.dump(old, open(path, "w"), indent=1)
        print("wrote", path, flush=True)

    save(walk(seed, depths, log=lambda s: print(s, flush=True), on_depth=save, resume=resume, gpu=gpu, rule=rule))


if __name__ == "__main__":
    main()
