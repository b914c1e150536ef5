"""The reference's bucket cap on the device (sph_set_bucket_cap) against the
oracle's gather form with the same cap, bit for bit.

The cap leaves a cell's particles ranked cap or later (by id) out of every
neighbour set of that hash (src/hash.c:153-164).  The cases put the cut
everywhere the device's walk can see it: inside a row's three cells at cap 3
on the dam break, inside the second 32-candidate chunk of a cell on a dense
lattice staged in LDS (caps 33 and 40), and in the fourth chunk on the
overfull lattice that walks global memory (cap 100); the fixture of five
101-particle clusters drops one particle each.  The fused step runs both
captured and eager, and switching the cap on and off between captured steps
must take the other kernels.
"""
import numpy as np
import pytest

import sph_amd as sa
from oracle import pyoracle as po
from tests.common import assert_bitexact, dam_break_tunables, golden, parts_from_snap, tun_from_bytes

pytestmark = pytest.mark.gpu


def make_pair(parts, tun, box, cap):
    s = sa.Solver(sa.Tunables.from_buffer_copy(bytes(tun)), sa.AABB.from_buffer_copy(bytes(box)),
                  len(parts))
    s.upload(parts)
    s.set_bucket_cap(cap)
    g = po.GatherSim(box, tun, len(parts))
    g.set_cap(cap)
    g.set_state(parts.view(po.P32) if parts.dtype != po.P32 else parts)
    return s, g


def caps_case():
    d = golden("caps.npz")
    return parts_from_snap(d["input"]), tun_from_bytes(d["tunables"]), po.make_box(4, 4)


def c1_case():
    d = golden("c1_dambreak.npz")
    return parts_from_snap(d["input"]), tun_from_bytes(d["tunables"]), po.make_box(60, 40)


def test_caps_fixture_neighbor_sets_at_first_hash():
    parts, tun, box = caps_case()
    n = len(parts)
    s, g = make_pair(parts, tun, box, sa.REF_MAX_BUCKET)
    ref = po.RefSim(box, tun, n, max_bucket=100, max_neighbors=4096)
    rp = parts.copy()
    for ph, name in enumerate(po.PHASES[:4]):
        ref.phase(rp, ph)
        g.phase(ph)
        s.phase(name)
    full, half, ids = s.neighbors(stride=1024)
    gc, gl = g.neighbor_sets(stride=1024)
    np.testing.assert_array_equal(full, gc)
    counts, _ = ref.lists(stride=4096)
    np.testing.assert_array_equal(half, counts)
    for i in range(n):
        assert sorted(ids[i, :full[i]].tolist()) == sorted(gl[i, :gc[i]].tolist()), i
    assert np.sum(full == 0) == 5
    assert_bitexact(s.download(), g.get_state(), what="caps fixture, first hash")


@pytest.mark.parametrize("graph", [True, False])
def test_caps_fixture_steps(graph):
    parts, tun, box = caps_case()
    s, g = make_pair(parts, tun, box, sa.REF_MAX_BUCKET)
    for step in range(1, 11):
        s.step(1, graph=graph)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"caps fixture, step {step}")
    assert s.counters()["bucket_overflow"] == g.counters()["bucket_overflow"] > 0


@pytest.mark.parametrize("graph", [True, False])
def test_dam_break_cap_3(graph):
    parts, tun, box = c1_case()
    s, g = make_pair(parts, tun, box, 3)
    for upto in (1, 2, 5, 10, 20):
        while s.counters()["steps"] < upto:
            s.step(1, graph=graph)
            g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"cap 3, step {upto}")


@pytest.mark.parametrize("cap", [33, 40])
def test_dense_lattice_cut_in_second_chunk(cap):
    parts = sa.scene_lattice(128, 128, 0.15, 0.01, seed=7)
    s, g = make_pair(parts, dam_break_tunables(), po.make_box(60, 40), cap)
    for step in range(1, 4):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"dense, cap {cap}, step {step}")
    s.step(2)
    g.step(2)
    assert_bitexact(s.download(), g.get_state(), what=f"dense, cap {cap}, graph steps 4-5")


def test_overfull_lattice_cap_100():
    parts = sa.scene_lattice(48, 48, 0.06, 0.01, seed=5)
    s, g = make_pair(parts, dam_break_tunables(), po.make_box(60, 40), sa.REF_MAX_BUCKET)
    for step in range(1, 4):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"overfull, cap 100, step {step}")
    s.step(2)
    g.step(2)
    assert_bitexact(s.download(), g.get_state(), what="overfull, cap 100, graph steps 4-5")
    assert s.counters()["max_cell_count"] > 200


def test_switching_the_cap_between_captured_steps():
    parts, tun, box = c1_case()
    s, g = make_pair(parts, tun, box, 0)
    for k, cap in enumerate((0, 3, 0, 5, 5)):
        s.set_bucket_cap(cap)
        g.set_cap(cap)
        s.step(4)
        g.step(4)
        assert_bitexact(s.download(), g.get_state(), what=f"segment {k}, cap {cap}")


@pytest.fixture
def shim():
    sa.lib().sph_shim_release()
    yield
    sa.lib().sph_shim_release()


def test_reference_entry_points_apply_the_grid_cap(shim):
    """The reference-named entry points take the cap from the grid record
    (max_bucket_size, src/hash.c:137): initParticles builds a lattice at
    0.09h (about 120 particles per cell), and 12 steps of the reference's
    loop through them equal the capped oracle bit for bit."""
    import ctypes
    from tests.common import run_reference_loop
    from tests.test_gpu_reference_scenes import Shim
    t = dam_break_tunables()
    box = sa.make_box(8, 8)
    sh = Shim(sa.Tunables.from_buffer_copy(bytes(t)), box, 2000)
    ctypes.c_uint.from_buffer(sh.grid, 36).value = sa.REF_MAX_BUCKET  # max_bucket_size
    sh.init(sa.make_box(3.6, 3.65), 40, 0.09)  # 40 x 40 particles from the origin
    assert sh.n == 1600
    host = sh.parts[:sh.n].copy()
    sa.lib().sph_shim_set_sync_interval(0)
    g = po.GatherSim(box, t, sh.n)
    g.set_cap(sa.REF_MAX_BUCKET)
    g.set_state(host.view(po.P32))
    run_reference_loop(sh.phase, lambda _: None, 12, [])
    run_reference_loop(g.phase, lambda _: None, 12, [])
    assert g.counters()["bucket_overflow"] > 0
    assert_bitexact(sh.state(), g.get_state(), what="entry points with the grid's cap, 12 steps")


def test_cap_in_a_batch_of_scenes():
    """Every scene of a batched context takes the cap, each as it would alone;
    scene offsets in the sorted arrays do not shift any scene's ranks."""
    from sph_amd import scenes as sc
    rng = np.random.default_rng(8)
    scenes = []
    for i in range(4):
        nx, ny = int(rng.integers(12, 30)), int(rng.integers(12, 30))
        box = sa.make_box(float(nx) * 0.6 + 4.0, float(ny) * 0.5 + 5.0)
        parts = sa.scene_lattice(nx, ny, 0.25, 0.01, seed=300 + i, x0=0.25, y0=0.25)
        scenes.append((sc.preset_tunables(i % 4), box, parts, i % 4))
    s = sa.Solver.batch([(t, b, len(p)) for t, b, p, _ in scenes])
    parts, counts = sc.concat(scenes)
    s.upload_batch(parts, counts)
    s.set_bucket_cap(5)
    s.step(8)
    out = s.download()
    offs = np.concatenate([[0], np.cumsum(counts)])
    for i, (t, b, p, _) in enumerate(scenes):
        g = po.GatherSim(b, t, len(p))
        g.set_cap(5)
        g.set_state(p.view(po.P32))
        g.step(8)
        assert_bitexact(out[offs[i]:offs[i + 1]], g.get_state(), what=f"scene {i}, cap 5")


def test_cap_with_split_tunables():
    """sph_step_tunables2 (new tunables between prediction and the hash) under
    the cap: the same as the oracle's phases with the same switch."""
    from tests.common import FIRST_HALF, SECOND_HALF
    parts, tun, box = c1_case()
    s, g = make_pair(parts, tun, box, 3)
    t2 = type(tun).from_buffer_copy(bytes(tun))
    t2.k = tun.k * 2.0
    t2.g = tun.g * 0.5
    for step in range(6):
        s.step_tunables2(sa.Tunables.from_buffer_copy(bytes(tun)),
                         sa.Tunables.from_buffer_copy(bytes(t2)))
        for ph in FIRST_HALF:
            g.phase(ph)
        g.set_tunables(t2)
        for ph in SECOND_HALF:
            g.phase(ph)
        g.set_tunables(tun)
    assert_bitexact(s.download(), g.get_state(), what="split tunables, cap 3, 6 steps")


def test_entry_points_keep_the_cap_when_the_context_grows(shim):
    """Particles added beyond the context's room make the entry points create
    it again; the grid record's cap holds from the new context's first pass
    on - also for a viscosity pass run at once (a host read in the middle of
    the step), before the step's hash_fluid has seen the grid again."""
    import ctypes
    from tests.common import FIRST_HALF, SECOND_HALF, run_reference_loop
    from tests.test_gpu_reference_scenes import Shim
    t = dam_break_tunables()
    box = sa.make_box(8, 8)
    sh = Shim(sa.Tunables.from_buffer_copy(bytes(t)), box, 1600)
    ctypes.c_uint.from_buffer(sh.grid, 36).value = sa.REF_MAX_BUCKET
    sh.init(sa.make_box(3.6, 3.65), 40, 0.09)
    sa.lib().sph_shim_set_sync_interval(1)
    run_reference_loop(sh.phase, lambda _: None, 4, [])
    n0, extra = sh.n, 300
    sh.alloc(n0 + extra, keep=n0)
    rng = np.random.default_rng(5)
    new = sh.parts[n0:]
    new["x"] = rng.uniform(0.5, 3.0, extra).astype(np.float32)
    new["y"] = rng.uniform(0.5, 3.0, extra).astype(np.float32)
    new["x_prev"], new["y_prev"] = new["x"], new["y"]
    new["id"] = np.arange(n0, n0 + extra)
    sh.params.number_fluid_particles_local = n0 + extra
    sa.lib().sph_shim_mark_host_dirty()
    host = sh.parts.copy()
    g = po.GatherSim(box, t, n0 + extra)
    g.set_cap(sa.REF_MAX_BUCKET)
    g.set_state(host.view(po.P32))
    g.phase(po.PH_HASH)  # the lists of the changed state
    for step in range(3):
        for ph in FIRST_HALF:
            sh.phase(ph)
            g.phase(ph)
        mid = sh.state()  # runs the recorded calls now, phase by phase
        assert_bitexact(mid, g.get_state(), fields=("x", "y", "v_x", "v_y"),
                        what=f"grown context, step {step}, after predict")
        for ph in SECOND_HALF:
            sh.phase(ph)
            g.phase(ph)
    assert g.counters()["bucket_overflow"] > 0
    assert_bitexact(sh.state(), g.get_state(), what="grown context, 3 steps")
