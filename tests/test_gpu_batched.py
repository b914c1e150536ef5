"""Batched scenes: many independent scenes stepped by the same launches.

Each scene must evolve exactly as it would alone: bit-exact against the
oracle's gather form run on that scene by itself.
"""
import numpy as np
import pytest

import sph_amd as sa
from sph_amd import scenes as sc
from oracle import pyoracle as po
from tests.common import assert_bitexact

pytestmark = pytest.mark.gpu


def oracle_run(tun, box, parts, steps):
    g = po.GatherSim(box, tun, len(parts))
    g.set_state(parts)
    g.step(steps)
    return g.get_state()


def split(out, counts):
    offs = np.concatenate([[0], np.cumsum(counts)])
    return [out[offs[i]:offs[i + 1]] for i in range(len(counts))]


def test_small_batch_bitexact_per_scene():
    rng = np.random.default_rng(4)
    scenes, caps = [], []
    for i in range(6):
        nx, ny = int(rng.integers(8, 40)), int(rng.integers(8, 40))
        box = sa.make_box(float(nx) * 1.3 + 3.0, float(ny) * 0.9 + 6.0)
        parts = sa.scene_lattice(nx, ny, 0.5, 0.01, seed=100 + i, x0=0.25, y0=0.25)
        if i == 2:
            parts = parts[:0].copy()  # an empty scene
        scenes.append((sc.preset_tunables(i % 4), box, parts, i % 4))
        caps.append(len(parts) + (37 if i % 2 else 0))  # spare capacity
    s = sa.Solver.batch([(t, b, c) for (t, b, _, _), c in zip(scenes, caps)])
    parts, counts = sc.concat(scenes)
    s.upload_batch(parts, counts)
    s.step(15)
    got = split(s.download(), counts)
    for i, (t, b, p, _) in enumerate(scenes):
        if len(p) == 0:
            assert len(got[i]) == 0
            continue
        assert_bitexact(got[i], oracle_run(t, b, p, 15), what=f"scene {i}")
    c = s.counters()
    assert c["bucket_overflow"] == 0 and c["neighbor_overflow"] == 0 and c["nonfinite"] == 0


def test_batched_config_256_scenes():
    """The batched configuration itself (256 scenes, 4.7M particles): a few
    steps, then three scenes of different sizes checked bit for bit."""
    scenes = sc.batched_scenes()
    s = sa.Solver.batch([(t, b, len(p)) for t, b, p, _ in scenes])
    parts, counts = sc.concat(scenes)
    s.upload_batch(parts, counts)
    assert s.nscenes == 256 and s.n == int(counts.sum())
    s.step(4)
    got = split(s.download(), counts)
    order = np.argsort(counts)
    for i in (order[0], order[len(order) // 2], order[-1]):
        t, b, p, _ = scenes[i]
        assert_bitexact(got[i], oracle_run(t, b, p, 4), what=f"scene {i} (n={len(p)})")
    for g in got:
        assert np.array_equal(np.sort(g["id"]), np.arange(len(g)))
    c = s.counters()
    assert c["nonfinite"] == 0 and c["bucket_overflow"] == 0


def test_scene_tunables_are_independent():
    """Changing one scene's tunables leaves the others untouched."""
    t0 = sc.preset_tunables(0)
    box = sa.make_box(20, 20)
    p = sa.scene_lattice(20, 20, 0.5, 0.01, seed=9, x0=0.25, y0=0.25)
    s = sa.Solver.batch([(t0, box, len(p)), (t0, box, len(p))])
    s.upload_batch(np.concatenate([p, p]), [len(p), len(p)])
    t1 = sc.preset_tunables(1)
    s.set_scene_tunables(1, t1)
    s.step(10)
    a, b = split(s.download(), [len(p), len(p)])
    assert_bitexact(a, oracle_run(t0, box, p, 10), what="scene 0")
    assert_bitexact(b, oracle_run(t1, box, p, 10), what="scene 1")


def test_frozen_scene_is_left_unchanged():
    """Frozen scenes pass through sph_step untouched; the others advance
    exactly as alone."""
    t = sc.preset_tunables(1)
    box = sa.make_box(24, 20)
    ps = [sa.scene_lattice(24, 16, 0.5, 0.01, seed=30 + i, x0=0.25, y0=0.25) for i in range(3)]
    s = sa.Solver.batch([(t, box, len(p)) for p in ps])
    s.upload_batch(np.concatenate(ps), [len(p) for p in ps])
    s.step(3)
    before = split(s.download(), [len(p) for p in ps])
    s.set_frozen([0, 1, 0])
    s.step(7)
    after = split(s.download(), [len(p) for p in ps])
    for f in ("x", "y", "v_x", "v_y"):
        assert np.array_equal(after[1][f], before[1][f]), f
    for i in (0, 2):
        assert_bitexact(after[i], oracle_run(t, box, ps[i], 10), what=f"scene {i}")
    s.set_frozen([0, 0, 0])
    s.step(2)
    got = split(s.download(), [len(p) for p in ps])
    assert_bitexact(got[0], oracle_run(t, box, ps[0], 12), what="scene 0 after thaw")


def test_sloshing_forcing_closed_form():
    """A lone particle far from the walls under sloshing: v_x after n steps is
    the sum of -A sin(omega k dt) dt over the steps taken (float tolerance)."""
    t = sc.preset_tunables(0)
    box = sa.make_box(200, 200)
    p = np.zeros(1, sa.PARTICLE)
    p["x"], p["y"] = 3.0, 150.0  # small |x|: fine float spacing for (x - x_prev) / dt
    s = sa.Solver.batch([(t, box, 1)])
    s.upload_batch(p, [1])
    A, T = 1.5, 2.0
    s.set_scene_forcing(0, A, T)
    dt = 1.0 / 60.0
    s.step(30)
    got = s.download()
    want = sum(-A * np.sin(2 * np.pi / T * k * dt) * dt for k in range(30))
    assert abs(got["v_x"][0] - want) < 1e-3, (got["v_x"][0], want)
    assert got["v_x"][0] < 0
