"""Batched uneven-tile workload: 96 render requests over 12 boundary-walk locations.

Definition (the workload spec's batch, with the choices it leaves open fixed):
  location k = 1..12 is the boundary walk of seed k (tests/golden/make_walk.py,
  --rule de) taken to depth 10^-e_k, e_k = round(E_LO + (E_HI - E_LO) * u) with u from
  splitmix64 seeded 4000 + k (first draw), E_LO = 50, E_HI = 1000;
  request r = 0..95 sits at location 1 + r % 12; its tile is square, side drawn from
  {64, 128, 256, 512, 1024} and its max_iter round(10^(4 + 3 u)), from splitmix64
  seeded 5000 + r (draws: side, max_iter, column x0, row y0);
  location k's view, width 10^-e_k around its centre, is a 1024 x 1024 sample grid;
  request r renders the side x side tile of that grid at column x0 = draw mod
  (1025 - side) and row y0 = draw mod (1025 - side), so the requests at one location
  share one reference orbit (built to the largest max_iter among them), one BLA
  table and one pair of dc tables; every location is rendered on the floatexp path
  with the centre pixel's full reference (the engine's MANDEL_FE=1 MANDEL_PERIODIC=0),
  smooth output, interior detection on.
Writes tests/golden/c4_seed<k>.json (one location centre each) and, once all twelve
exist, tests/golden/c4_requests.json.
Usage: python tests/golden/make_c4.py [--plan] [--seeds 1,2,...] [--partial] [--out DIR]
  --plan: print the draws only; --seeds: walk only these locations (the request file
  is written once every location's centre exists); --partial: write the request file
  from the locations that exist, with their requests only, the others listed under
  "missing"; --out: as make_walk.py
"""
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_walk import SplitMix64, option, walk  # noqa: E402

E_LO, E_HI = 50, 1000
SIDES = (64, 128, 256, 512, 1024)
NLOC, NREQ = 12, 96
GRID = 1024


def unit(rng):
    return (rng.next() >> 11) / float(1 << 53)


def draws():
    locs = []
    for k in range(1, NLOC + 1):
        rng = SplitMix64(4000 + k)
        locs.append(dict(seed=k, depth=int(round(E_LO + (E_HI - E_LO) * unit(rng)))))
    reqs = []
    for r in range(NREQ):
        rng = SplitMix64(5000 + r)
        side = SIDES[rng.next() % len(SIDES)]
        mxit = int(round(10 ** (4 + 3 * unit(rng))))
        x0 = rng.next() % (GRID + 1 - side)
        y0 = rng.next() % (GRID + 1 - side)
        reqs.append(dict(location=1 + r % NLOC, side=side, max_iter=mxit, x0=x0, y0=y0))
    return locs, reqs


def main():
    locs, reqs = draws()
    if "--plan" in sys.argv:
        for l in locs:
            print(l)
        print(sum(q["side"] ** 2 for q in reqs), "samples")
        return
    args = sys.argv[1:]
    seeds = option(args, "--seeds")
    seeds = None if seeds is None else {int(x) for x in seeds.split(",")}
    out_dir = option(args, "--out", HERE)
    out = {"_definition": "tests/golden/make_c4.py (seeded batch of uneven tiles)", "locations": [], "requests": reqs}
    for l in locs:
        path = os.path.join(out_dir, f"c4_seed{l['seed']}.json")
        if not os.path.exists(path) and (seeds is None or l["seed"] in seeds):
            res = walk(l["seed"], [l["depth"]], log=lambda s: print(s, flush=True), rule="de")
            c = res[str(l["depth"])]
            json.dump(dict(seed=l["seed"], rule="de", depth=l["depth"], re=c["re"], im=c["im"], width=c["width"],
                           steps=c["steps"], probe_mxit=c["probe_mxit"]), open(path, "w"), indent=1)
            print("location", l, "done", flush=True)
        if os.path.exists(path):
            c = json.load(open(path))
            out["locations"].append(dict(seed=l["seed"], depth=l["depth"], re=c["re"], im=c["im"]))
    if len(out["locations"]) == NLOC:
        json.dump(out, open(os.path.join(out_dir, "c4_requests.json"), "w"), indent=1)
        print("wrote c4_requests.json", flush=True)
    elif "--partial" in args and out["locations"]:
        # the locations whose walks are done, with their requests; the others listed
        have = {l["seed"] for l in out["locations"]}
        out["missing"] = sorted(l["seed"] for l in locs if l["seed"] not in have)
        out["requests"] = [q for q in reqs if q["location"] in have]
        json.dump(out, open(os.path.join(out_dir, "c4_requests.json"), "w"), indent=1)
        print(f"wrote c4_requests.json with {len(have)} of {NLOC} locations", flush=True)


if __name__ == "__main__":
    main()
