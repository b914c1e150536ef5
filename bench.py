"""Benchmark of the TinySPH step on one MI355X.

Workload: the plain 2D dam break at N = 1,048,576 - a 1024 x 1024 lattice at
spacing 0.5h (h = 1) in a 1200 x 800 box, seeded +-0.01h jitter, v = 0,
g = 9.8, dt = 1/60, rho0 = 10, k = 0.004, k_near = 0.01, k_spring = 10,
sigma = 0, beta = 0.3, no mover.  A step is one full reference step
(src/fluid.c:273-341): gravity, viscosity, predict, hash + density,
relaxation, velocity update, hash.  The state stays in HBM; the timed region
replays the captured HIP graph of the fused step.  The value is the median of
--repeats timed runs (default 3), each of exactly --steps steps from the same
uploaded initial state after --warmup untimed steps.

Prints one JSON line: particle-steps/s, the HBM roofline of the dominant
kernel (HIP-event timed on the solver's stream), the whole-step algorithmic
traffic, and the reference's own C step (compiled from its sources, -O3
-ffast-math as its makefile) timed on one host core over a bounded sample.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import sph_amd as sa  # noqa: E402
from sph_amd import scenes as sc  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes per particle-step of the whole step, by phase (one
# compulsory read of each phase's inputs and one write of its outputs;
# neighbour re-reads through LDS / L2 not counted): 192 per particle plus 56
# per cell (2 sorts x (clear + count + scan read / write) + 3 gathers x
# (start + count)).
B_ALG = (192, 56)

# Algorithmic bytes per particle (P) and per cell (C) of each kernel of the
# fused step: every input array read once, every output written once.
KERNEL_BYTES = {
    # reads x, y, vx, vy, key (20); writes x', y', key', slot (16); count rmw per cell
    "visc_predict": (36, 8),
    "scan_reduce": (0, 4),       # read count
    "scan_apply": (0, 12),       # read count, write start, clear count
    "scatter": (24, 0),          # read key, slot, id (12); write src, id, sorted key (12)
    "reorder": (48, 0),          # read src, id, sorted key (12), 2 float2 fields (16); write 2 fields, id (20)
    "density": (20, 4),          # read x, y, key (12); write rho, rho_near (8); cell starts
    #                              (plus 16 B of candidate masks written for the relaxation)
    "relax": (52, 12),           # read x, y, xp, yp, rho, rhon, key (28); write x, y, vx, vy, key, slot (24)
    #                              (plus the 16 B of masks read back)
}
# The per-kernel timing behind `roofline` always covers the same state,
# whatever --steps / --warmup are: eager steps KERNEL_WINDOW + 1 ..
# KERNEL_WINDOW + KERNEL_STEPS after a fresh upload (the settling column,
# where the gathers are costliest early in the run).  The committed profiles
# are taken over the same window (--kernels-only).
KERNEL_WINDOW = 220
KERNEL_STEPS = 10
LAUNCHES_PER_STEP = 11
# Second line: the compressed second half of the run.
LATE_START, LATE_STEPS = 500, 100
L2_FILE = os.path.join(ROOT, "profiles", "l2_hit_c2_dam_break.txt")
SHADOW_EVERY = 8

# profiler names of the kernels (template instances of the fused step)
KERNEL_SYMBOLS = {"visc_predict": "k_visc_predict<true, true, true, false>",
                  "relax": "k_relax<true, false>", "density": "k_density<false>",
                  "reorder": "k_reorder<2>", "scatter": "k_scatter",
                  "scan_apply": "k_scan_apply", "scan_reduce": "k_scan_reduce"}
# launches of each kind per step
KERNEL_LAUNCHES = {"visc_predict": 1, "scan_reduce": 2, "scan_apply": 2, "scatter": 2,
                   "reorder": 2, "density": 1, "relax": 1}


def workload(n_side=1024, box=(1200.0, 800.0), seed=12345):
    tun = sa.make_tunables(rest_density=10.0, smoothing_radius=1.0, g=9.8, k=0.004, k_near=0.01,
                           k_spring=10.0, sigma=0.0, beta=0.3, time_step=1.0 / 60.0,
                           mover_type=0, mover_width=0.0, mover_height=0.0, active=1)
    parts = sa.scene_lattice(n_side, n_side, 0.5, 0.01, seed=seed)
    return tun, sa.make_box(*box), parts


def cpu_model():
    """The host CPU's model name (Linux), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(tun, box, parts, budget_s=30.0, min_steps=50):
    """The reference's own step (oracle/_ref, -O3 -ffast-math), 1 core, over
    at least min_steps steady steps (more if they fit in budget_s)."""
    from oracle import pyoracle as po
    fast = po.have_compiled_ref(fast=True)
    if fast:
        sim = po.CompiledRef(box, tun, len(parts), fast=True)
        sim.load(parts)
        step = sim.step
        kind = "reference"
    else:  # the oracle's restatement of the same scatter path
        sim = po.RefSim(box, tun, len(parts), real="fast_f32")
        p = np.ascontiguousarray(parts.view(po.P32))
        step = lambda k: sim.step(p, k)  # noqa: E731
        kind = "port"
    step(1)  # first step: allocations touched, lists empty
    t0 = time.perf_counter()
    step(1)
    per = time.perf_counter() - t0
    k = max(min_steps, min(200, int(budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    step(k)
    dt = time.perf_counter() - t0
    return {"value": len(parts) * k / dt, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "cpu": cpu_model(),
            "sample": f"{k} steady steps (after 2 warm-up) of the same N={len(parts)} dam break, "
                      f"{dt:.1f} s on one host core ({os.cpu_count()} visible)"}


def _cpu_scene_worker(args):
    """One batched scene stepped by the reference's own C step on one core."""
    index, budget_s = args
    from oracle import pyoracle as po
    t, b, parts, _ = sc.batched_scenes(only={index})[0]
    sim = po.CompiledRef(b, t, len(parts), fast=True)
    sim.load(np.ascontiguousarray(parts).view(po.P32))
    sim.step(1)
    k, t0 = 0, time.perf_counter()
    while True:
        sim.step(1)
        k += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            return len(parts), k, el


def cpu_baseline_batched(cores, budget_s=10.0):
    """The batched configuration on the host: one scene per core, the
    reference's own C step (oracle/_ref, -O3 -ffast-math), `cores` scenes
    spread over the 256 (every 256/cores-th) stepped concurrently for
    budget_s.  Runs in forked workers before this process touches the GPU."""
    import multiprocessing as mp
    from oracle import pyoracle as po
    if not po.have_compiled_ref(fast=True):
        return None
    idx = [i * 256 // cores for i in range(cores)]
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_scene_worker, [(i, budget_s) for i in idx])
    value = sum(n * k / el for n, k, el in res)
    parts = sum(n for n, _, _ in res)
    return {"value": value, "unit": "particle-steps/s", "cores": cores, "kind": "reference",
            "cpu": cpu_model(),
            "sample": f"{cores} of the 256 scenes ({parts} particles, every {256 // cores}th), one "
                      f"per core, {budget_s:.0f} s each of steady steps after one warm-up step; "
                      f"throughput summed over the concurrent workers"}


def timed_repeats(s, upload, steps, warmup, repeats):
    """Median wall time of `repeats` runs of exactly `steps` graph-replayed
    steps, each from the same uploaded initial state after `warmup` untimed
    steps, bracketed by stream synchronisation.  Returns (median, all)."""
    runs = []
    for _ in range(repeats):
        upload()
        s.reset_counters()
        s.step(warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        runs.append(time.perf_counter() - t0)
    return float(np.median(runs)), runs


def batched(steps, warmup, repeats=3):
    """The batched configuration: 256 scenes (2k-64k particles each, four
    parameter presets) stepped by the same launches through one graph."""
    scenes = sc.batched_scenes()
    parts, counts = sc.concat(scenes)
    s = sa.Solver.batch([(t, b, len(p)) for t, b, p, _ in scenes])
    el, runs = timed_repeats(s, lambda: s.upload_batch(parts, counts), steps, warmup, repeats)
    n = int(counts.sum())
    counters = s.counters()
    # the gaps between the step's 11 launches: graph-replayed step time against
    # the sum of the kernels' own durations (HIP events on each launch) over
    # the steps right after the timed ones
    kt = s.time_kernels(5)
    kernel_sum = sum(v * KERNEL_LAUNCHES.get(k, 1) for k, v in kt.items() if k in KERNEL_LAUNCHES)
    s.close()
    return {"value": n * steps / el, "unit": "particle-steps/s", "ms_per_step": el * 1e3 / steps,
            "steps": steps, "warmup": warmup, "scenes": len(scenes), "particles": n,
            "repeats_ms_per_step": [r * 1e3 / steps for r in runs],
            "kernels_ms": kt, "kernel_sum_ms": kernel_sum,
            "launch_gap_ms": el * 1e3 / steps - kernel_sum,
            "counters": counters}


def surrogate_leg(step_eq, nscenes=64, n=65536, tol=0.01, seed=11, sample_every=100):
    """Batched inference plus fallback: nscenes sloshing dam breaks of n
    particles.  The same scenes run step_eq plain steps (batched graph), and
    step_eq / 4 judged surrogate macro-steps (the plain solver re-takes any
    scene whose mean density is off its plain density trend by more than tol,
    and every scene on a fixed schedule).  Both runs are timed in chunks of
    sample_every plain-step equivalents; between chunks, outside the timers,
    each scene's mean density is sampled (a hash and density pass of the
    current state, which leaves the run unchanged): the drift of the
    surrogate run from the plain run, per scene, at the final state and of
    the time-averaged mean density.  A third run adds shadow judging (one
    scene in SHADOW_EVERY per macro-step also takes the macro-step plainly
    from the same start) and reports the density error of the macro-steps
    the trend judge accepted against that plain result: the per-step error
    the 1% bound is about."""
    from sph_amd import surrogate as su
    if not os.path.exists(su.WEIGHTS):
        return {"skipped": f"no trained weights at {os.path.relpath(su.WEIGHTS, ROOT)}"}
    # 4x wider than the training scenes (16,384 particles in 128 x 96) at the
    # same fluid depth
    scenes = sc.sloshing_dam_breaks(nscenes, n=n, box=(512.0, 96.0), seed=seed, widen=4)
    total = nscenes * n
    chunks = step_eq // sample_every
    per_chunk = sample_every // su.MACRO

    def density(s):
        s.phase("hash_density")
        return s.scene_density()[0].astype(np.float64)

    plain = su.make_batch(scenes)
    plain.step(1)
    t_plain, rho_p = 0.0, []
    for _ in range(chunks):
        plain.sync()
        t0 = time.perf_counter()
        plain.step(sample_every)
        plain.sync()
        t_plain += time.perf_counter() - t0
        rho_p.append(density(plain))
    c_plain = plain.counters()
    plain.close()

    sur = su.make_batch(scenes)
    sur.step(1)
    m = sa.Surrogate(seed=0)
    su.load_weights(m)
    m.macro_step(sur, tol)  # sets up the judge's buffers (untimed)
    t_sur, rho_s, flags = 0.0, [], []
    for _ in range(chunks):
        dt, f = su.run_judged(m, sur, per_chunk, tol)
        t_sur += dt
        flags.append(f)
        rho_s.append(density(sur))
    sur.close()
    # the same run with shadow judging: 1 scene in SHADOW_EVERY per
    # macro-step also takes the macro-step plainly from the same start,
    # which measures the error of the steps the trend judge accepts
    sh = su.make_batch(scenes)
    sh.step(1)
    m2 = sa.Surrogate(seed=0)
    su.load_weights(m2)
    m2.set_shadow(SHADOW_EVERY)
    t_sh, sh_flags = su.run_judged(m2, sh, chunks * per_chunk + 1, tol)
    shadow = m2.shadow_stats()
    shadow.update({"every": SHADOW_EVERY, "seconds": t_sh,
                   "value": total * (chunks * per_chunk + 1) * su.MACRO / t_sh,
                   "shadow_failure_rate": float((sh_flags == 3).mean()),
                   "fallback_rate": float((sh_flags != 0).mean()),
                   "definition": "per-scene relative error of the mean density (and the "
                                 "per-particle relative RMS density error) of macro-steps the "
                                 "trend judge accepted, against the plain macro-step from the "
                                 "same start; 1 scene in SHADOW_EVERY per macro-step, rotating"})
    sh.close()
    m2.close()
    m.close()
    rho_p, rho_s = np.array(rho_p), np.array(rho_s)
    err_final = np.abs(rho_s[-1] - rho_p[-1]) / rho_p[-1]
    err_avg = np.abs(rho_s.mean(0) - rho_p.mean(0)) / rho_p.mean(0)
    flags = np.concatenate(flags)
    steps_done = chunks * sample_every
    v_plain = total * steps_done / t_plain
    v_sur = total * steps_done / t_sur
    return {"scenes": nscenes, "particles": total, "plain_step_equivalents": steps_done,
            "plain": {"value": v_plain, "unit": "particle-steps/s", "seconds": t_plain},
            "surrogate": {"value": v_sur, "unit": "particle-steps/s", "seconds": t_sur,
                          "macro_steps": int(flags.shape[0]), "tol": tol},
            "speedup": v_sur / v_plain,
            "fallback_rate": float((flags != 0).mean()) if flags.size else 0.0,
            "judged_failure_rate": float((flags == 1).mean()) if flags.size else 0.0,
            "forced_plain_rate": float((flags == 2).mean()) if flags.size else 0.0,
            "shadow": shadow,
            "drift_vs_plain_run": {"final_max": float(err_final.max()),
                              "final_mean": float(err_final.mean()),
                              "time_avg_max": float(err_avg.max()),
                              "time_avg_mean": float(err_avg.mean()),
                              "definition": "per-scene mean density vs the plain run of the same "
                                            f"scenes: at the final state, and averaged over "
                                            f"{chunks} samples every {sample_every} steps"},
            "plain_counters": c_plain}


def calibration_copy(mbytes):
    """A device-to-device copy of known size through the HIP runtime (a blit
    kernel), run after the timed region: its FETCH_SIZE / WRITE_SIZE counter
    readings calibrate the counters' units against known bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = int(mbytes * 1e6)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def profile_summary(db, last=0, skip=0):
    """Text summary of a rocprofv3 results database (its default output):
    per kernel, the dispatch count and mean / total duration from the kernel
    trace, and the mean per dispatch of every collected counter (FETCH_SIZE
    and WRITE_SIZE are in KiB).  `last` keeps only the library's last N
    dispatches of the trace (the steady state; the runtime's own blit
    kernels are left out), after dropping its final `skip` dispatches."""
    import sqlite3
    con = sqlite3.connect(db)
    names = {r[0] for r in con.execute("select name from sqlite_master")}
    lines = []
    short = lambda k: re.sub(r"\(.*", "", k).replace("void ", "")  # noqa: E731
    if "kernels" in names:
        rows = con.execute("select name, dispatch_id, start, end, duration from kernels "
                           "order by start").fetchall()
        if last:  # the library's last N dispatches (runtime blit kernels left out)
            rows = [r for r in rows if not short(r[0]).startswith("__amd")]
            rows = rows[:len(rows) - skip][-last:]
        agg = {}
        for name, _, _, _, dur in rows:
            agg.setdefault(short(name), []).append(dur)
        if rows:
            lines.append(f"{'kernel':42s} {'calls':>6s} {'mean_us':>10s} {'total_ms':>10s}")
            for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
                lines.append(f"{k:42s} {len(v):6d} {sum(v) / len(v) / 1e3:10.2f} "
                             f"{sum(v) / 1e6:10.3f}")
            span = rows[-1][3] - rows[0][2]
            busy = sum(r[4] for r in rows)
            lines.append(f"span {span / 1e3:.1f} us, kernels busy {busy / 1e3:.1f} us "
                         f"({len(rows)} dispatches)")
    if "counters_collection" in names:
        rows = con.execute("select dispatch_id, kernel_name, counter_name, value "
                           "from counters_collection").fetchall()
        agg, disp = {}, {}
        for d, k, cn, v in rows:
            agg.setdefault(short(k), {}).setdefault(cn, 0.0)
            agg[short(k)][cn] += v
            disp.setdefault(short(k), set()).add(d)
        lines.append(f"{'kernel':42s} {'dispatches':>10s} counters (mean per dispatch)")
        for k, m in sorted(agg.items()):
            nd = len(disp[k])
            cs = ", ".join(f"{cn}={v / nd:.1f}" for cn, v in sorted(m.items()))
            lines.append(f"{k:42s} {nd:10d} {cs}")
    return "\n".join(lines)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_c2_dam_break.txt")


def _short(k):
    return re.sub(r"\(.*", "", k).replace("void ", "")


def _window_ids(rows, last):
    """Dispatch ids of the library's last `last` dispatches (the runtime's
    blit kernels left out), or of every dispatch when last is 0."""
    lib = sorted({d for d, k in rows if not _short(k).startswith("__amd")})
    return set(lib[-last:]) if last else {d for d, _ in rows}


def _counter_means(db, counter, last=0):
    """Per kernel: {dispatch id: value} of one counter (summed over its
    records), over the library's last `last` dispatches when given."""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, value from counters_collection "
                       "where counter_name = ?", (counter,)).fetchall()
    keep = _window_ids([(d, k) for d, k, _ in rows], last)
    out = {}
    for d, k, v in rows:
        if d in keep:
            m = out.setdefault(_short(k), {})
            m[d] = m.get(d, 0.0) + v
    return out


def counter_summary(db, last=0, simds=1024, xcds=8):
    """Mean per dispatch of every counter of a rocprofv3 --pmc run, per kernel
    (the last `last` dispatches when given), and the number of records each
    counter has per dispatch.  rocprofv3 reports a counter summed over the
    device: GRBM_GUI_ACTIVE over the `xcds` XCDs, so one XCD's busy cycles are
    GRBM_GUI_ACTIVE / xcds.  Derived where the counters are present:
    VALU issue = SQ_INSTS_VALU x 4 cycles (a wave64 instruction on a 16-lane
    SIMD) / (busy cycles x simds); resident waves per SIMD = SQ_WAVE_CYCLES x 4
    (it counts quad-cycles) / (busy cycles x simds)."""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value "
                       "from counters_collection").fetchall()
    keep = _window_ids([(r[0], r[1]) for r in rows], last)
    agg, disp, recs = {}, {}, {}
    for d, k, cn, v in rows:
        if d not in keep:
            continue
        name = re.sub(r"\(.*", "", k).replace("void ", "")
        agg.setdefault(name, {}).setdefault(cn, 0.0)
        agg[name][cn] += v
        disp.setdefault(name, set()).add(d)
        recs[(name, cn, d)] = recs.get((name, cn, d), 0) + 1
    lines = []
    for k in sorted(agg):
        n = len(disp[k])
        m = {cn: v / n for cn, v in agg[k].items()}
        nrec = max(c for (kk, _, _), c in recs.items() if kk == k)
        busy = m.get("GRBM_GUI_ACTIVE", 0.0) / xcds
        line = (f"{k}  dispatches={n}  records/dispatch/counter={nrec}  "
                + "  ".join(f"{cn}={v:.6g}" for cn, v in sorted(m.items())))
        if busy and "SQ_INSTS_VALU" in m:
            line += f"  | VALU issue {m['SQ_INSTS_VALU'] * 4 / (busy * simds):.3f}"
        if busy and "SQ_WAVE_CYCLES" in m:
            line += f"  waves/SIMD {m['SQ_WAVE_CYCLES'] * 4 / (busy * simds):.2f}"
        lines.append(line)
    return "\n".join(lines)


def traffic_summary(fetch_db, write_db, copy_mb, last=0):
    """HBM bytes per launch of every kernel from a FETCH_SIZE and a WRITE_SIZE
    counter run (KiB per dispatch), each scaled by the factor that makes the
    same counter read the known size of the bench's calibration copy
    (--calib-copy-mb: a device-to-device copy of copy_mb MB, which reads and
    writes copy_mb MB).  With `last`, only the library's last N dispatches of
    each run (the calibration copies are taken from the whole run)."""
    fetch = _counter_means(fetch_db, "FETCH_SIZE", last)
    write = _counter_means(write_db, "WRITE_SIZE", last)
    known_kib = copy_mb * 1e6 / 1024.0
    cal = {}
    # the calibration copies run after the window: take them from the whole run
    for name, tab in (("fetch", _counter_means(fetch_db, "FETCH_SIZE")),
                      ("write", _counter_means(write_db, "WRITE_SIZE"))):
        vals = [v for k, m in tab.items() if "copyBuffer" in k for v in m.values()
                if v > 0.5 * known_kib]
        if not vals:
            raise SystemExit(f"no calibration copy of {copy_mb} MB found in the {name} run")
        cal[name] = known_kib / (sum(vals) / len(vals))
    lines = [f"calibration: {copy_mb:g} MB device copy; scale FETCH_SIZE x{cal['fetch']:.4f}, "
             f"WRITE_SIZE x{cal['write']:.4f}",
             f"{'kernel':42s} {'dispatches':>10s} {'FETCH_KiB':>12s} {'WRITE_KiB':>12s} "
             f"{'HBM_MB_per_launch':>18s}"]
    for k in sorted(set(fetch) | set(write)):
        f = list(fetch.get(k, {}).values())
        w = list(write.get(k, {}).values())
        fm = sum(f) / len(f) if f else 0.0
        wm = sum(w) / len(w) if w else 0.0
        mb = (fm * cal["fetch"] + wm * cal["write"]) * 1024.0 / 1e6
        lines.append(f"{k:42s} {max(len(f), len(w)):10d} {fm:12.1f} {wm:12.1f} {mb:18.3f}")
    return "\n".join(lines)


def l2_summary(db, last=0):
    """L2 (TCC) hit rate per kernel from a rocprofv3 --pmc TCC_HIT_sum
    TCC_MISS_sum run: hits / (hits + misses) of the L2 requests, summed over
    the channels, per dispatch (the last `last` dispatches when given)."""
    hit = _counter_means(db, "TCC_HIT_sum", last)
    miss = _counter_means(db, "TCC_MISS_sum", last)
    lines = [f"{'kernel':42s} {'dispatches':>10s} {'TCC_HIT':>14s} {'TCC_MISS':>14s} {'L2_hit_rate':>12s}"]
    for k in sorted(set(hit) | set(miss)):
        h = list(hit.get(k, {}).values())
        m = list(miss.get(k, {}).values())
        hm = sum(h) / len(h) if h else 0.0
        mm = sum(m) / len(m) if m else 0.0
        rate = hm / (hm + mm) if hm + mm > 0 else float("nan")
        lines.append(f"{k:42s} {max(len(h), len(m)):10d} {hm:14.1f} {mm:14.1f} {rate:12.4f}")
    return "\n".join(lines)


def l2_hit_of(kernel_name):
    """L2 hit rate of a kernel from the committed summary (None when absent)."""
    if not os.path.exists(L2_FILE):
        return None
    with open(L2_FILE) as f:
        for line in f:
            parts = re.split(r"\s{2,}", line.strip())
            if parts and parts[0] == kernel_name and len(parts) >= 5:
                try:
                    return float(parts[-1])
                except ValueError:
                    return None
    return None


def kernel_window(s, upload, start=KERNEL_WINDOW):
    """Per-kernel mean durations (ms) over eager steps start + 1 ..
    start + KERNEL_STEPS after a fresh upload, each kernel timed by
    HIP events that its own launch records on the solver's stream."""
    upload()
    s.reset_counters()
    s.step(start)
    s.sync()
    return s.time_kernels(KERNEL_STEPS), s.counters()


def late_regime(s, upload):
    """Graph-replayed steps LATE_START + 1 .. LATE_START + LATE_STEPS: the
    compressed second half of the dam break."""
    upload()
    s.reset_counters()
    s.step(LATE_START)
    s.sync()
    t0 = time.perf_counter()
    s.step(LATE_STEPS)
    s.sync()
    el = time.perf_counter() - t0
    n = s.n
    return {"value": n * LATE_STEPS / el, "unit": "particle-steps/s",
            "ms_per_step": el * 1e3 / LATE_STEPS,
            "steps": f"{LATE_START + 1}-{LATE_START + LATE_STEPS} after a fresh upload",
            "counters": s.counters()}


def capped_line(s, upload, steps, warmup, repeats):
    """The same workload under the reference's 100-slot buckets
    (sph_set_bucket_cap): the timed steps as the headline times them, the
    per-kernel window, and graph-replayed steps 601-700, where the first
    cells pass 100 particles and the cap starts to drop particles."""
    s.set_bucket_cap(sa.REF_MAX_BUCKET)
    try:
        el, runs = timed_repeats(s, upload, steps, warmup, repeats)
        kt, _ = kernel_window(s, upload)
        upload()
        s.reset_counters()
        s.step(600)
        s.sync()
        t0 = time.perf_counter()
        s.step(100)
        s.sync()
        el_late = time.perf_counter() - t0
        counters = s.counters()
    finally:
        s.set_bucket_cap(0)
    n = s.n
    return {"bucket_cap": sa.REF_MAX_BUCKET, "value": n * steps / el, "unit": "particle-steps/s",
            "ms_per_step": el * 1e3 / steps, "repeats_ms_per_step": [r * 1e3 / steps for r in runs],
            "kernels_ms": kt,
            "steps_601_700": {"ms_per_step": el_late * 1e3 / 100, "value": n * 100 / el_late,
                              "counters": counters}}


def c1_line(cpu=True):
    """The N = 4096 configuration (64 x 64 lattice, box 60 x 40, 1000 steps):
    the GPU through the step graph, and (cpu) the reference's own C step
    (oracle/_ref, -O3 -ffast-math) on one host core over the same 1000
    steps."""
    tun, box, parts = workload(64, (60.0, 40.0), seed=20260101)
    n = len(parts)
    s = sa.Solver(tun, box, n)
    s.upload(parts)
    s.step(1)
    s.sync()
    t0 = time.perf_counter()
    s.step(999)
    s.sync()
    el = time.perf_counter() - t0
    s.close()
    out = {"particles": n, "steps": "2-1000 (graph)", "value": n * 999 / el,
           "unit": "particle-steps/s", "ms_per_step": el * 1e3 / 999}
    from oracle import pyoracle as po
    if cpu and po.have_compiled_ref(fast=True):
        sim = po.CompiledRef(box, tun, n, fast=True)
        sim.load(parts)
        sim.step(1)
        t0 = time.perf_counter()
        sim.step(999)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * 999 / el, "unit": "particle-steps/s", "cores": 1,
                               "kind": "reference", "cpu": cpu_model(),
                               "sample": "steps 2-1000 of the same scene, one host core"}
    return out


def drop_in_line(steps=42):
    """The reference loop (src/fluid.c:270-372) in the C driver linked
    against the reference-named entry points, on the 1024 x 1024 dam break
    (no jitter: the reference's own lattice) in the 1200 x 800 box: wall time
    per step of the loop with the state copied back to the host every step
    (all fields), every 4 steps (the reference packs coordinates every 4th
    substep), every 4 steps positions only, and never - all with the grid
    record's 100-slot buckets, as the reference sets them - and never with
    uncapped buckets (-u).  Runs the driver as a child process."""
    import subprocess
    exe = os.path.join(ROOT, "sph_amd", "sph_main")
    if not os.path.exists(exe):
        return {"skipped": "sph_amd/sph_main not built"}
    out = {}
    for name, args in (("sync_every_step", ["-s", "1"]), ("sync_every_4", ["-s", "4"]),
                       ("sync_every_4_positions", ["-s", "4", "-p"]), ("no_sync", ["-s", "0"]),
                       ("no_sync_uncapped", ["-s", "0", "-u"])):
        r = subprocess.run([exe, "-t", *args, "-b", "1200,800", "1024", "1024", str(steps)],
                           capture_output=True, text=True, timeout=300)
        m = re.search(r"([0-9.]+) ms/step, ([0-9.e+]+) particle-steps/s", r.stdout)
        out[name] = ({"ms_per_step": float(m.group(1)), "value": float(m.group(2))} if m
                     else {"error": (r.stderr or r.stdout)[-300:]})
    out["unit"] = "particle-steps/s"
    out["steps"] = f"3-{steps} of the loop"
    return out


def traffic_of(kernel_name):
    """Calibrated HBM bytes per launch of a kernel from the committed traffic
    summary (None when absent)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            # columns are separated by two or more spaces; kernel names hold single ones
            parts = re.split(r"\s{2,}", line.strip())
            if parts and parts[0] == kernel_name and len(parts) >= 5:
                try:
                    return float(parts[-1]) * 1e6
                except ValueError:
                    return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=3,
                    help="timed runs (each from the same initial state); value is the median")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload line")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate line")
    ap.add_argument("--only-surrogate", action="store_true",
                    help="run only the surrogate line (for profiling)")
    ap.add_argument("--surrogate-steps", type=int, default=2000,
                    help="plain-step equivalents of the surrogate line")
    ap.add_argument("--calib-copy-mb", type=float, default=0.0,
                    help="also run a device copy of this many MB (profiler calibration)")
    ap.add_argument("--profile-summary", metavar="DB",
                    help="print a summary of a rocprofv3 results database and exit")
    ap.add_argument("--skip", type=int, default=0,
                    help="with --profile-summary and --last N: drop the final M dispatches "
                         "before taking the last N")
    ap.add_argument("--last", type=int, default=0,
                    help="with --profile-summary: keep the last N trace dispatches")
    ap.add_argument("--counter-summary", metavar="DB",
                    help="print per-kernel counter means of a rocprofv3 --pmc run "
                         "(with --last N: its last N dispatches) and exit")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="print calibrated HBM traffic per kernel from two rocprofv3 --pmc "
                         "databases and exit (uses --calib-copy-mb; with --last N, the last "
                         "N dispatches)")
    ap.add_argument("--l2-summary", metavar="DB",
                    help="print per-kernel L2 hit rates of a rocprofv3 --pmc TCC_HIT_sum "
                         "TCC_MISS_sum run and exit (with --last N, the last N dispatches)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the per-kernel timing window (for the profiler) and print "
                         "its table")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the late-regime, bucket-cap, N=4096 and drop-in lines")
    ap.add_argument("--kernel-window", type=int, default=KERNEL_WINDOW,
                    help="with --kernels-only: the steps before the timed window (default: "
                         "the headline's window; 550 profiles the late regime)")
    ap.add_argument("--bucket-cap", type=int, default=0,
                    help="with --kernels-only: run under this bucket cap (the reference's: 100)")
    ap.add_argument("--only-c1", action="store_true",
                    help="run only the N = 4096 line (with --no-cpu: the GPU leg alone) and "
                         "print it")
    args = ap.parse_args()
    if args.profile_summary:
        print(profile_summary(args.profile_summary, args.last, args.skip))
        return
    if args.counter_summary:
        print(counter_summary(args.counter_summary, args.last))
        return
    if args.traffic_summary:
        print(traffic_summary(*args.traffic_summary, args.calib_copy_mb or 256.0, args.last))
        return
    if args.l2_summary:
        print(l2_summary(args.l2_summary, args.last))
        return
    if args.kernels_only:
        tun, box, parts = workload()
        s = sa.Solver(tun, box, len(parts))
        if args.bucket_cap:
            s.set_bucket_cap(args.bucket_cap)
        kt, cnt = kernel_window(s, lambda: s.upload(parts), args.kernel_window)
        if args.calib_copy_mb > 0:
            calibration_copy(args.calib_copy_mb)
        print(json.dumps({"window": f"eager steps {args.kernel_window + 1}-"
                                    f"{args.kernel_window + KERNEL_STEPS}"
                                    " after a fresh upload", "kernels_ms": kt, "counters": cnt}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one scene on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.only_c1:
        print(json.dumps({"c1": c1_line(cpu=not args.no_cpu)}))
        return

    # host-core baseline of the batched configuration, before any GPU use
    # (forked workers)
    cpu_batched = None
    if not args.no_cpu and not args.no_batched:
        cpu_batched = cpu_baseline_batched(min(16, os.cpu_count() or 1))
    if args.only_surrogate:
        print(json.dumps({"batched_surrogate": surrogate_leg(args.surrogate_steps)}))
        return
    tun, box, parts = workload()
    n = len(parts)
    s = sa.Solver(tun, box, n)
    el, runs = timed_repeats(s, lambda: s.upload(parts), args.steps, args.warmup, args.repeats)
    ms_per_step = el * 1e3 / args.steps
    value = n * args.steps / el
    counters = s.counters()

    # per-kernel HIP-event timing on the solver's stream
    # (kinds this bench does not know, e.g. from another library build, are
    # left out)
    late = None if args.no_extra else late_regime(s, lambda: s.upload(parts))
    capped = (None if args.no_extra else
              capped_line(s, lambda: s.upload(parts), args.steps, args.warmup, args.repeats))
    kt_all, window_counters = kernel_window(s, lambda: s.upload(parts))
    kt = {k: ms for k, ms in kt_all.items() if k in KERNEL_BYTES}
    sx, sy = s.cells()[2]
    c = sx * sy / n
    kernel_bw = {}
    for k, ms in kt.items():
        bp, bc = KERNEL_BYTES[k]
        bytes_launch = n * bp + sx * sy * bc
        kernel_bw[k] = {"ms": ms, "bytes": bytes_launch,
                        "GBs": bytes_launch / (ms * 1e-3) / 1e9 if ms > 0 else None}
    dom = max(kt, key=lambda k: kt[k] * KERNEL_LAUNCHES[k])
    b_step = B_ALG[0] + B_ALG[1] * c
    ach = kernel_bw[dom]["GBs"]
    out = {
        "metric": "particle-steps/s, 1 MI355X: plain vs batched+surrogate at <=1% density error",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "repeats_ms_per_step": [r * 1e3 / args.steps for r in runs],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain 2D dam break, N=1,048,576 (1024x1024 lattice at 0.5h, "
                               "box 1200x800, jitter 0.01h), fused step via HIP graph",
                   "particles": n, "cells": sx * sy,
                   "library": os.path.relpath(sa.LIB_PATH, ROOT)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS if ach else None,
                     "traffic": traffic_of(KERNEL_SYMBOLS.get(dom, dom)),
                     "algorithmic_bytes": kernel_bw[dom]["bytes"],
                     "kernel_ms": kt[dom],
                     "window": f"eager steps {KERNEL_WINDOW + 1}-{KERNEL_WINDOW + KERNEL_STEPS} "
                               "after a fresh upload (the profiles' window)",
                     "l2_hit_rate": l2_hit_of(KERNEL_SYMBOLS.get(dom, dom)),
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT)
                     if traffic_of(KERNEL_SYMBOLS.get(dom, dom)) is not None else None},
        "step_roofline": {"bytes_per_particle_step": b_step,
                          "achieved_GBs": b_step * value / 1e9,
                          "frac": b_step * value / 1e9 / HBM_PEAK_GBS},
        "kernels": kernel_bw,
        "counters": counters,
        "window_counters": window_counters,
    }
    if late:
        late["step_roofline_frac"] = b_step * late["value"] / 1e9 / HBM_PEAK_GBS
        out["late_regime"] = late
    if capped:
        out["bucket_cap_100"] = capped
    if not args.no_batched:
        out["batched"] = batched(max(10, args.steps // 2), max(2, args.warmup // 2),
                                 args.repeats)
        if cpu_batched:
            out["batched"]["cpu_baseline"] = cpu_batched
    if not args.no_extra:
        s.close()
        out["c1"] = c1_line()
        out["drop_in"] = drop_in_line()
    if not args.no_surrogate:
        out["batched_surrogate"] = surrogate_leg(args.surrogate_steps)
    if args.calib_copy_mb > 0:
        calibration_copy(args.calib_copy_mb)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tun, box, parts)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
