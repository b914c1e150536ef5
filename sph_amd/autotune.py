"""Autotuner of the gather kernels' launch shape.

The plain step's three gather kernels (viscosity + predict, density,
relaxation) stage each workgroup's neighbour cells in LDS up to a tile
capacity, and deal their workgroups to the 8 XCDs in runs of consecutive
tiles (longer runs share more candidate rows in one XCD's L2, shorter runs
spread dense regions of the fluid more evenly over the XCDs).  Both are
compile-time shapes: sph_amd.build compiles one library per shape into
sph_amd/variants/.  This tuner times every variant on a workload, each in a
fresh process (SPH_AMD_LIB names the variant), and records the results and
the fastest variant in configs/tuned.json; the variant recorded before stays
unless another beats it by more than --keep-within (0.5%, the spread of
repeated timings), so that noise does not flip the choice.  sph_amd.lib() loads the variant
recorded there for the default workload unless SPH_AMD_LIB names a library.
Every variant computes the same bit-exact step; only speed differs.

    python -m sph_amd.autotune [--workload dam_break_1m|batched_256] [--steps K]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

from . import TUNED

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
DEFAULT_WORKLOAD = "dam_break_1m"


def _solver(workload):
    import sph_amd as sa
    from sph_amd import scenes as sc
    if workload == "dam_break_1m":
        tun, box, parts = sc.dam_break(1024, 1024, (1200.0, 800.0), seed=12345)
        s = sa.Solver(tun, box, len(parts))
        s.upload(parts)
        return s, len(parts)
    if workload == "batched_256":
        scenes = sc.batched_scenes()
        parts, counts = sc.concat(scenes)
        s = sa.Solver.batch([(t, b, len(p)) for t, b, p, _ in scenes])
        s.upload_batch(parts, counts)
        return s, int(counts.sum())
    raise ValueError(workload)


def _worker(workload, steps, warmup):
    """Time one variant (the library SPH_AMD_LIB names) on the workload."""
    s, n = _solver(workload)
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    el = time.perf_counter() - t0
    print(json.dumps({"ms_per_step": el * 1e3 / steps, "particle_steps_per_s": n * steps / el}))


def pick_best(results, incumbent=None, keep_within=0.005):
    """The fastest variant, unless the incumbent (the variant recorded before)
    is within keep_within (a fraction) of it: a challenger must beat the
    incumbent by more than the run-to-run spread of one timing to replace it."""
    best = min(results, key=lambda k: results[k]["ms_per_step"])
    if incumbent in results and (results[incumbent]["ms_per_step"]
                                 <= results[best]["ms_per_step"] * (1.0 + keep_within)):
        return incumbent
    return best


def tune(workload=DEFAULT_WORKLOAD, steps=100, warmup=20, timeout=300, incumbent=None,
         keep_within=0.005):
    from sph_amd import build as bd
    results = {}
    for cap, run in bd.VARIANTS:
        path = os.path.join(bd.VARIANT_DIR, bd.variant_name(cap, run))
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path}: build the variants first (python -m sph_amd.build)")
        env = dict(os.environ, SPH_AMD_LIB=path)
        r = subprocess.run([sys.executable, "-m", "sph_amd.autotune", "--worker", "--workload",
                            workload, "--steps", str(steps), "--warmup", str(warmup)],
                           cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"variant {bd.variant_name(cap, run)} failed: {r.stderr[-2000:]}")
        results[bd.variant_name(cap, run)] = json.loads(r.stdout.strip().splitlines()[-1])
        print(bd.variant_name(cap, run), results[bd.variant_name(cap, run)], flush=True)
    best = pick_best(results, incumbent, keep_within)
    return {"best": best, "steps": steps, "warmup": warmup, "results": results,
            "keep_within": keep_within}


def load_tuned():
    if not os.path.exists(TUNED):
        return {}
    with open(TUNED) as f:
        return json.load(f)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=["dam_break_1m", "batched_256"])
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--out", default=TUNED)
    ap.add_argument("--keep-within", type=float, default=0.005,
                    help="keep the recorded variant unless another is this fraction faster")
    args = ap.parse_args(argv)
    if args.worker:
        _worker(args.workload, args.steps, args.warmup)
        return
    prev = load_tuned()
    incumbent = prev.get("workloads", {}).get(args.workload, {}).get("best")
    res = tune(args.workload, args.steps, args.warmup, incumbent=incumbent,
               keep_within=args.keep_within)
    cfg = prev if args.out == TUNED else {}
    cfg.setdefault("default", DEFAULT_WORKLOAD)
    cfg["comment"] = ("written by python -m sph_amd.autotune on one MI355X: ms per graph-replayed "
                      "step of each library variant (LDS tile capacity t, tiles per XCD run x) "
                      "after the warm-up steps")
    cfg.setdefault("workloads", {})[args.workload] = res
    with open(args.out, "w") as f:
        json.dump(cfg, f, indent=1, sort_keys=True)
    print(json.dumps({"workload": args.workload, "best": res["best"]}))


if __name__ == "__main__":
    main()
