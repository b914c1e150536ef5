"""The oracle's reference-semantics restatement against the golden fixtures.

The fixtures were written by the reference's own hash.c / fluid.c compiled
from its sources (tests/golden/make_golden.py).  The restatement must match
them bit for bit: every field, every phase, every half-list length.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.common import (SNAP_FIELDS, golden, parts_from_snap, run_reference_loop,
                          tun_from_bytes)


def snap_of(p):
    return np.stack([p[f] for f in SNAP_FIELDS])


def assert_snap_equal(got, want, what):
    g = np.ascontiguousarray(got, np.float32).view(np.uint32)
    w = np.ascontiguousarray(want, np.float32).view(np.uint32)
    for k, f in enumerate(SNAP_FIELDS):
        bad = np.nonzero(g[k] != w[k])[0]
        assert len(bad) == 0, f"{what}: field {f} differs at {bad[:5]}"


def test_c1_dambreak_steps():
    d = golden("c1_dambreak.npz")
    p = parts_from_snap(d["input"])
    sim = po.RefSim(po.make_box(60, 40), tun_from_bytes(d["tunables"]), len(p))
    done = 0
    for s in (1, 2, 10, 100):
        sim.step(p, s - done)
        done = s
        assert_snap_equal(snap_of(p), d[f"step{s}"], f"step {s}")
        if s == 2:
            np.testing.assert_array_equal(sim.lists()[0], d["halflist_len_step2"])
    assert sim.counters() == dict(bucket_overflow=0, neighbor_overflow=0, key_out_of_range=0)


def test_c1_dambreak_step_1000():
    """The whole 1000-step horizon of the N=4096 configuration."""
    d = golden("c1_dambreak.npz")
    p = parts_from_snap(d["input"])
    sim = po.RefSim(po.make_box(60, 40), tun_from_bytes(d["tunables"]), len(p))
    sim.step(p, 1000)
    assert_snap_equal(snap_of(p), d["step1000"], "step 1000")
    assert sim.counters() == dict(bucket_overflow=0, neighbor_overflow=0, key_out_of_range=0)


@pytest.mark.parametrize("preset", ["x", "y", "a", "b"])
def test_reference_presets(preset):
    """The reference's demo scene (src/fluid.c:88-159) under its presets
    (src/controls.c:344-401), including super goo's k_spring = -30 and the
    impulse clamp of src/fluid.c:459 at sigma = 100, beta = 10."""
    d = golden("presets.npz")
    p = parts_from_snap(d["input"])
    box = po.make_box(*d["box"])
    sim = po.RefSim(box, tun_from_bytes(d[f"{preset}_tunables"]), len(p))
    sim.step(p, 2)
    assert_snap_equal(snap_of(p), d[f"{preset}_step2"], f"{preset} step 2")
    sim.step(p, 18)
    assert_snap_equal(snap_of(p), d[f"{preset}_step20"], f"{preset} step 20")


def test_reference_loop_with_per_frame_tunables():
    """The reference loop with the render rank's tunables arriving between
    predict_positions and hash_fluid on every frame's last substep
    (src/fluid.c:293-294): mover on the attract-mode path, presets taking
    turns."""
    d = golden("presets.npz")
    p = parts_from_snap(d["input"])
    box = po.make_box(*d["box"])
    frames = [tun_from_bytes(f) for f in d["frames_tunables"]]
    sim = po.RefSim(box, tun_from_bytes(d["frames_initial_tunables"]), len(p))
    run_reference_loop(lambda ph: sim.phase(p, ph), sim.set_tunables, 48, frames)
    assert_snap_equal(snap_of(p), d["frames_step48"], "frames step 48")
    run_reference_loop(lambda ph: sim.phase(p, ph), sim.set_tunables, 48, frames, first_step=48)
    assert_snap_equal(snap_of(p), d["frames_step96"], "frames step 96")


def test_phases64_every_phase_and_lists():
    d = golden("phases64.npz")
    p = parts_from_snap(d["input"])
    sim = po.RefSim(po.make_box(6, 6), tun_from_bytes(d["tunables"]), len(p))
    k = li = 0
    for s in range(8):
        for ph in range(7):
            sim.phase(p, ph)
            assert_snap_equal(snap_of(p), d["states"][k], f"step {s} phase {po.PHASES[ph]}")
            k += 1
            if ph in (3, 6):
                c, l = sim.lists(stride=64)
                np.testing.assert_array_equal(c, d["list_len"][li])
                for i in range(len(p)):
                    np.testing.assert_array_equal(l[i, :c[i]], d["lists"][li][i, :c[i]])
                li += 1


@pytest.mark.parametrize("mover", ["sphere", "rect"])
def test_movers(mover):
    d = golden("movers.npz")
    p = parts_from_snap(d["input"])
    sim = po.RefSim(po.make_box(12, 10), tun_from_bytes(d[mover + "_tunables"]), len(p))
    for s in range(6):
        sim.step(p, 1)
        assert_snap_equal(snap_of(p), d[mover][s], f"{mover} step {s}")


def test_caps_drop_like_reference():
    d = golden("caps.npz")
    p = parts_from_snap(d["input"])
    sim = po.RefSim(po.make_box(4, 4), tun_from_bytes(d["tunables"]), len(p))
    for s in (1, 2):
        sim.step(p, 1)
        assert_snap_equal(snap_of(p), d[f"step{s}"], f"caps step {s}")
        np.testing.assert_array_equal(sim.lists()[0], d[f"halflist_len_step{s}"])
    c = sim.counters()
    assert c["bucket_overflow"] > 0 and c["neighbor_overflow"] > 0


@pytest.mark.skipif(not po.have_compiled_ref(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_restatement_vs_compiled_reference_random(seed):
    """Random small scenes with random tunables and movers, 20 steps."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 400))
    box = po.make_box(float(rng.uniform(6, 14)), float(rng.uniform(6, 14)))
    t = po.make_tunables(rest_density=float(rng.uniform(2, 12)), smoothing_radius=1.0,
                         g=float(rng.uniform(0, 12)), k=float(rng.uniform(0.001, 0.3)),
                         k_near=float(rng.uniform(0.001, 1.0)), k_spring=float(rng.uniform(0, 10)),
                         sigma=float(rng.uniform(0, 1)), beta=float(rng.uniform(0, 1)),
                         time_step=1.0 / 60.0, mover_type=int(rng.integers(0, 2)),
                         mover_center_x=float(box.max_x / 2), mover_center_y=1.5,
                         mover_width=2.0, mover_height=1.5, active=1)
    p = np.zeros(n, po.P32)
    p["x"] = rng.uniform(0.1, box.max_x * 0.6, n).astype(np.float32)
    p["y"] = rng.uniform(0.1, box.max_y * 0.6, n).astype(np.float32)
    p["v_x"] = rng.uniform(-1, 1, n).astype(np.float32)
    p["id"] = np.arange(n)
    ref = po.CompiledRef(box, t, n)
    ref.load(p)
    sim = po.RefSim(box, t, n)
    q = p.copy()
    for s in range(20):
        ref.step(1)
        sim.step(q, 1)
        assert_snap_equal(snap_of(q), snap_of(ref.store()), f"seed {seed} step {s}")
