"""Writes tests/golden/c2_digests.json: digests of the 1,048,576-particle dam
break (1024 x 1024 lattice at 0.5h in a 1200 x 800 box, +-0.01h jitter from
seed 12345, the benchmark's workload) after 250, 500 and 1000 steps of the
oracle's gather form, the definition the device is checked against.

For each checkpoint and each field of the state in id order (x, y, v_x, v_y,
x_prev, y_prev, density, density_near) the file holds the SHA-256 of the
float32 bytes plus a few statistics for diagnosis, and the oracle's
counters (bucket_overflow, neighbor_overflow, key_out_of_range,
max_cell_count, max_neighbors).  The full states are 40 MB each; digests are
what the repository keeps.

Run in the build container (CPU only; about an hour):
    python tests/golden/make_c2_digests.py

With --cap 100 the oracle applies the reference's 100-slot buckets
(src/hash.c:153-164; the run first overflows a cell between steps 600 and
610) and writes tests/golden/c2_digests_cap100.json at steps 650 and 1000.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402

FIELDS = ("x", "y", "v_x", "v_y", "x_prev", "y_prev", "density", "density_near")
CHECKPOINTS = (250, 500, 1000)
SEED = 12345
OUT = os.path.join(HERE, "c2_digests.json")


def field_digests(state):
    """SHA-256 and statistics of each field of an id-ordered state."""
    out = {}
    for f in FIELDS:
        a = np.ascontiguousarray(state[f], dtype="<f4")
        d = a.astype(np.float64)
        out[f] = {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
                  "sum": float(d.sum()), "min": float(d.min()), "max": float(d.max())}
    return out


def c2_tunables():
    """rho0 = 10, h = 1, g = 9.8 (applied as -g), dt = 1/60, k = 0.004,
    k_near = 0.01, k_spring = 10, sigma = 0, beta = 0.3, no mover."""
    return po.make_tunables(rest_density=10.0, smoothing_radius=1.0, g=9.8, k=0.004, k_near=0.01,
                            k_spring=10.0, sigma=0.0, beta=0.3, time_step=1.0 / 60.0,
                            mover_type=0, mover_width=0.0, mover_height=0.0, active=1)


def c2_particles():
    """The library's host-side scene builder (SplitMix64 jitter; host code,
    no device needed)."""
    import sph_amd as sa
    return sa.scene_lattice(1024, 1024, 0.5, 0.01, seed=SEED)


def main():
    cap = int(sys.argv[sys.argv.index("--cap") + 1]) if "--cap" in sys.argv else 0
    out = os.path.join(HERE, f"c2_digests_cap{cap}.json") if cap else OUT
    checkpoints = (650, 1000) if cap else CHECKPOINTS
    tun = c2_tunables()
    box = po.make_box(1200, 800)
    parts = c2_particles()
    sim = po.GatherSim(box, tun, len(parts))
    sim.set_cap(cap)
    sim.set_state(parts)
    res = {"workload": "1024 x 1024 lattice at 0.5h, box 1200 x 800, jitter 0.01h, seed 12345; "
                       "g 9.8, dt 1/60, rho0 10, k 0.004, k_near 0.01, k_spring 10, sigma 0, "
                       "beta 0.3, no mover; gather form of the oracle (float32)",
           "seed": SEED, "bucket_cap": cap, "checkpoints": {}}
    done = 0
    t0 = time.time()
    for cp in checkpoints:
        while done < cp:
            k = min(10, cp - done)
            sim.step(k)
            done += k
            print(f"step {done}  {time.time() - t0:.0f} s  {sim.counters()}", flush=True)
        st = sim.get_state()
        res["checkpoints"][str(cp)] = {"fields": field_digests(st), "counters": sim.counters()}
        with open(out + ".tmp", "w") as f:
            json.dump(res, f, indent=1)
        os.replace(out + ".tmp", out)
        print("wrote checkpoint", cp, flush=True)


if __name__ == "__main__":
    main()
