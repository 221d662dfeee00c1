"""Per-sample MPFR escape counts and values for a benchmark frame: the reference
against which the forms' accuracy is stated (the workload asks for the fraction of
samples at identical escape count and the largest value error against MPFR).

For `n` samples drawn uniformly without replacement (numpy default_rng(seed), the
frame's reference pixel left out) the oracle iterates each sample's c, formed on the
engine's mpf grid (mandel_perturbation.cpp:1403-1424), at the frame's precision + 64
bits (accuratePointCompute, mandel_perturbation.cpp:719-894), with the exterior
distance estimate where the frame asks for it.  A subset is recomputed at twice that
precision; the file records whether its escape counts stayed the same.

Writes tests/golden/mpfr_<name>.json.
Usage: python tests/golden/make_mpfr_samples.py NAME WALK.json DEPTH W H MXIT CMETHOD N [SEED]
e.g.   python tests/golden/make_mpfr_samples.py c3 tests/golden/walk_de_seed7.json 1000 5760 3240 2000000 2 2000
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import mxo  # noqa: E402


def main():
    name, walk, depth, W, H, mxit, cm, n = sys.argv[1:9]
    seed = int(sys.argv[9]) if len(sys.argv) > 9 else 11
    W, H, mxit, cm, n = int(W), int(H), int(mxit), int(cm), int(n)
    c = json.load(open(walk))[depth]
    scale = "4" + "0" * int(depth)
    pb = mxo.precision_bits(scale)
    rng = np.random.default_rng(seed)
    centre = (H // 2) * W + W // 2
    idx = rng.choice(W * H, n + 1, replace=False)
    idx = np.sort(idx[idx != centre][:n]).astype(np.int32)
    threads = int(os.environ.get("MPFR_THREADS", os.cpu_count() or 1))
    t = time.time()
    val, nesc = mxo.direct_list(c["re"], c["im"], scale, W, H, mxit, idx, c_method=cm, prec_bits=pb, wp_bits=pb + 64,
                                threads=threads)
    print(f"{n} samples in {time.time() - t:.0f} s", flush=True)
    k = min(64, n)
    v2, n2 = mxo.direct_list(c["re"], c["im"], scale, W, H, mxit, idx[:k], c_method=cm, prec_bits=pb,
                             wp_bits=2 * pb + 128, threads=threads)
    out = dict(_definition="tests/golden/make_mpfr_samples.py (per-sample MPFR reference)", workload=name,
               walk=os.path.basename(walk), depth=int(depth), W=W, H=H, max_iter=mxit, c_method=cm, seed=seed,
               precision_bits=pb + 64, idx=idx.tolist(), n_esc=nesc.tolist(), value=val.tolist(),
               check_precision_bits=2 * pb + 128, check_samples=k,
               check_n_esc_identical=bool(np.array_equal(n2, nesc[:k])),
               check_value_max_rel=float(np.max(np.abs(v2 - val[:k]) / np.maximum(np.abs(val[:k]), 1e-300))))
    path = os.path.join(HERE, f"mpfr_{name}.json")
    json.dump(out, open(path, "w"))
    print("wrote", path, "check identical:", out["check_n_esc_identical"], "max rel", out["check_value_max_rel"])


if __name__ == "__main__":
    main()
