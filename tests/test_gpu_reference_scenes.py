"""The reference's own scene and parameter sets on the device, tunables that
change between and within steps, and the reference-named entry points.

* The demo scene of src/fluid.c:88-159 under each preset of
  src/controls.c:344-401, bit-exact against the oracle's gather form.
  Preset y ("super goo": sigma = 100, beta = 10, k_spring = -30) drives the
  impulse clamp of src/fluid.c:459; the oracle counts the clamped impulses.
* Per-frame tunables through the context API: g, k, sigma and the sphere
  mover's centre (the reference's attract-mode path, src/renderer.c:493-531)
  change between replays of the captured step graph, with no re-capture.
* The reference loop's own timing of those changes - between
  predict_positions and hash_fluid on the last substep of each frame
  (src/fluid.c:293-294) - through sph_step_tunables2, through the C driver
  linked against the reference-named entry points, and through the same
  entry points called from here with the loop interrupted mid-step.
* sph_shim_mark_host_dirty: the host changes its particles (adds some, up
  to more than the context holds) and the run continues with viscosity on.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import sph_amd as sa
from oracle import pyoracle as po
from sph_amd import scenes as sc
from tests.common import FIRST_HALF, SECOND_HALF, assert_bitexact, run_reference_loop

pytestmark = pytest.mark.gpu


def pair(preset="x"):
    t, box, parts, info = sc.reference_scene(preset)
    s = sa.Solver(t, box, info["capacity"])
    s.upload(parts)
    g = po.GatherSim(box, t, len(parts))
    g.set_state(np.ascontiguousarray(parts).view(po.P32))
    return s, g, t, box, parts, info


@pytest.mark.parametrize("lanes", [0, 1])
@pytest.mark.parametrize("preset", ["x", "y", "a", "b"])
def test_reference_presets_bitexact(preset, lanes):
    """lanes 0: the size rule (four lanes per particle for this scene); 1: one."""
    s, g, t, box, parts, info = pair(preset)
    s.set_gather_lanes(lanes)
    done = 0
    for upto in (1, 2, 5, 20, 60):
        print(f"preset {preset}: steps to {upto}", flush=True)
        s.step(upto - done)
        g.step(upto - done)
        done = upto
        assert_bitexact(s.download(), g.get_state(),
                        what=f"preset {preset}, lanes {lanes}, step {upto}")
    c, oc = s.counters(), g.counters()
    assert c["nonfinite"] == 0 and c["bucket_overflow"] == oc["bucket_overflow"] == 0
    assert c["max_cell_count"] == oc["max_cell_count"]
    if preset == "y":
        assert oc["impulse_clamps"] > 0  # the clamp of src/fluid.c:459 was taken


def test_tunables_and_mover_change_between_graph_replays():
    """30 frames of 4 steps: before each frame new tunables (presets x, y,
    a, b taking turns, so g, k, k_near, k_spring, sigma, beta and rho0
    change, and the mover moves) are set, then the frame's steps replay the
    graph captured on the first one."""
    s, g, t, box, parts, info = pair("x")
    frames = sc.reference_frames(t, box, 30, frames_per_preset=5)
    s.step(4)
    g.step(4)
    for f, tf in enumerate(frames):
        s.set_tunables(tf)
        g.set_tunables(tf)
        s.step(4)
        g.step(4)
        if f % 5 == 4 or f == 0:
            print(f"frame {f}", flush=True)
            assert_bitexact(s.download(), g.get_state(), what=f"frame {f}")
    assert_bitexact(s.download(), g.get_state(), what="after 30 frames")


def test_split_step_tunables_match_mid_step_change():
    """sph_step_tunables2 on every frame's last substep against the oracle
    changing the tunables between predict and hash, as the reference loop
    does."""
    s, g, t, box, parts, info = pair("x")
    frames = sc.reference_frames(t, box, 16, frames_per_preset=4)
    cur = t
    for step in range(64):
        if step % 4 == 3:
            nxt = frames[step // 4]
            s.step_tunables2(cur, nxt)
            cur = nxt
        else:
            s.step(1)
    run_reference_loop(g.phase, g.set_tunables, 64, frames)
    assert_bitexact(s.download(), g.get_state(), what="64 steps, 16 mid-step changes")


def write_frame_file(path, t, box, info, frames, steps_per_frame=4):
    """The frame file sph_main reads (its header comment gives the layout)."""
    prm = sa.Param()
    prm.tunable_params = t
    prm.number_fluid_particles_global = info["n"]
    water = sa.make_box(*info["box"])
    with open(path, "wb") as f:
        f.write(b"SPHFRAME")
        f.write(np.int32(1).tobytes())
        f.write(bytes(box))
        f.write(bytes(water))
        f.write(bytes(prm))
        f.write(np.float32(info["spacing"]).tobytes())
        f.write(np.array([info["nx"], info["capacity"], steps_per_frame, len(frames)],
                         np.int32).tobytes())
        for tf in frames:
            f.write(bytes(tf))


@pytest.mark.parametrize("sync", [("1",), ("4", "-p")])
def test_reference_loop_driver_with_per_frame_tunables(tmp_path, sync):
    """The C driver runs the reference's loop (src/fluid.c:270-372) against
    the reference-named entry points: initParticles builds the demo scene,
    and every frame's tunables arrive mid-step from a file.  Its final state
    equals the oracle's through the same schedule, bit for bit."""
    t, box, parts, info = sc.reference_scene("x")
    frames = sc.reference_frames(t, box, 24)
    ff = tmp_path / "frames.bin"
    write_frame_file(ff, t, box, info, frames)
    out = tmp_path / "out.bin"
    exe = os.path.join(sa.PKG_DIR, "sph_main")
    cmd = [exe, "-s", *sync, "-f", str(ff), "96", str(out)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, sa.PARTICLE)
    assert len(got) == info["n"]
    g = po.GatherSim(box, t, len(parts))
    g.set_cap(sa.REF_MAX_BUCKET)  # the driver's grid record carries the reference's 100
    g.set_state(np.ascontiguousarray(parts).view(po.P32))
    run_reference_loop(g.phase, g.set_tunables, 96, frames)
    assert_bitexact(got, g.get_state(), what="driver, 96 steps")


# ---- the reference-named entry points called from here ----------------------

class Shim:
    """The reference's calling convention over ctypes: a particle array of
    `capacity` records and the pointer array into it (src/fluid.c:181-199)."""

    def __init__(self, t, box, capacity):
        self.L = sa.lib()
        vp = ctypes.c_void_p
        for name, args in (("initParticles", [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp,
                                              ctypes.c_int, ctypes.c_float, vp]),
                           ("apply_gravity", [vp, vp]), ("viscosity_impluses", [vp, vp, vp]),
                           ("predict_positions", [vp, vp, vp]),
                           ("identify_oob_particles", [vp, vp, vp, vp, vp]),
                           ("hash_fluid", [vp, vp, vp, ctypes.c_bool]),
                           ("hash_halo", [vp, vp, vp, ctypes.c_bool]),
                           ("double_density_relaxation", [vp, vp, vp]),
                           ("updateVelocities", [vp, vp, vp, vp])):
            f = getattr(self.L, name)
            f.argtypes = args
            f.restype = None
        self.params = sa.Param()
        self.params.tunable_params = t
        self.box = sa.AABB.from_buffer_copy(bytes(box))
        self.grid = (ctypes.c_char * 40)()
        self.alloc(capacity)

    def alloc(self, capacity, keep=0):
        old = getattr(self, "parts", None)
        self.parts = np.zeros(capacity, sa.PARTICLE)
        if keep:
            self.parts[:keep] = old[:keep]
        self.ptrs = (ctypes.c_void_p * capacity)()
        base = self.parts.ctypes.data
        for i in range(capacity):
            self.ptrs[i] = base + 52 * i
        self.capacity = capacity

    def init(self, water, nx, spacing):
        self.L.initParticles(self.ptrs, self.parts.ctypes.data, ctypes.byref(water), 0, nx, None,
                             self.capacity, ctypes.c_float(spacing), ctypes.byref(self.params))

    @property
    def n(self):
        return self.params.number_fluid_particles_local

    def phase(self, ph):
        L, p, pr, b = self.L, self.ptrs, ctypes.byref(self.params), ctypes.byref(self.box)
        if ph == po.PH_GRAVITY:
            L.apply_gravity(p, pr)
        elif ph == po.PH_VISCOSITY:
            L.viscosity_impluses(p, None, pr)
        elif ph == po.PH_PREDICT:
            L.predict_positions(p, b, pr)
            L.identify_oob_particles(p, self.parts.ctypes.data, None, b, pr)
        elif ph == po.PH_HASH_DENSITY:
            L.hash_fluid(p, self.grid, pr, True)
            L.hash_halo(p, self.grid, pr, True)
        elif ph == po.PH_RELAX:
            L.double_density_relaxation(p, None, pr)
        elif ph == po.PH_UPDATE_VELOCITIES:
            L.updateVelocities(p, None, b, pr)
        elif ph == po.PH_HASH:
            L.hash_fluid(p, self.grid, pr, False)
            L.hash_halo(p, self.grid, pr, False)

    def set_tunables(self, t):
        self.params.tunable_params = sa.Tunables.from_buffer_copy(bytes(t))

    def state(self):
        assert self.L.sph_shim_sync() == 0, sa.lib().sph_last_error()
        return self.parts[:self.n].copy()


@pytest.fixture
def shim():
    sa.lib().sph_shim_release()
    yield
    sa.lib().sph_shim_release()


def test_shim_steps_interrupted_mid_step(shim):
    """A host read in the middle of a step (sph_shim_sync after
    predict_positions) makes the entry points replay that step phase by
    phase; every other step runs fused.  Same result either way."""
    t, box, parts, info = sc.reference_scene("y")
    frames = sc.reference_frames(t, box, 10)
    sh = Shim(t, box, info["capacity"])
    sh.init(sa.make_box(*info["box"]), info["nx"], info["spacing"])
    assert sh.n == info["n"]
    L = sa.lib()
    L.sph_shim_set_sync_interval(0)
    g = po.GatherSim(box, t, sh.n)
    g.set_state(np.ascontiguousarray(parts).view(po.P32))
    frame = 0
    for step in range(40):
        for ph in FIRST_HALF:
            sh.phase(ph)
            g.phase(ph)
        if step % 7 == 3:
            mid = sh.state()  # mid-step read: the recorded calls run first
            assert_bitexact(mid, g.get_state(), fields=("x", "y", "v_x", "v_y"),
                            what=f"step {step} after predict")
        if step % 4 == 3:
            sh.set_tunables(frames[frame])
            g.set_tunables(frames[frame])
            frame += 1
        for ph in SECOND_HALF:
            sh.phase(ph)
            g.phase(ph)
    assert_bitexact(sh.state(), g.get_state(), what="40 steps")


def test_shim_mark_host_dirty_with_added_particles(shim):
    """The host adds particles twice: first within the room initParticles
    was given (twice the lattice, as src/fluid.c:156 sizes it), then beyond
    it, which makes the entry points grow the device context.  After each
    change the run continues with viscosity on (the state is re-hashed as
    the reference's lists are kept), bit-exact against the oracle given the
    same particles and a hash."""
    t, box, parts, info = sc.reference_scene("x")
    t.sigma = 2.0  # viscosity strong enough that skipping a pass would show
    sh = Shim(t, box, info["n"] * 2)
    sh.init(sa.make_box(*info["box"]), info["nx"], info["spacing"])
    sa.lib().sph_shim_set_sync_interval(1)
    loop = lambda fn, k: run_reference_loop(fn, lambda _: None, k, [])  # noqa: E731
    loop(sh.phase, 10)
    g = po.GatherSim(box, t, sh.n)
    g.set_state(np.ascontiguousarray(parts).view(po.P32))
    loop(g.phase, 10)
    assert_bitexact(sh.state(), g.get_state(), what="before any change")
    rng = np.random.default_rng(3)
    for extra in (300, 2 * info["n"]):
        n0 = sh.n
        if n0 + extra > sh.capacity:
            sh.alloc(n0 + extra, keep=n0)
        new = sh.parts[n0:n0 + extra]
        new["x"] = rng.uniform(1.0, 14.0, extra).astype(np.float32)
        new["y"] = rng.uniform(4.0, 8.0, extra).astype(np.float32)
        new["x_prev"], new["y_prev"] = new["x"], new["y"]
        new["v_x"] = rng.uniform(-1, 1, extra).astype(np.float32)
        new["id"] = np.arange(n0, n0 + extra)
        for i in range(n0, n0 + extra):  # initParticles left these pointers NULL
            sh.ptrs[i] = sh.parts.ctypes.data + 52 * i
        sh.params.number_fluid_particles_local = n0 + extra
        sa.lib().sph_shim_mark_host_dirty()
        host = sh.parts[:n0 + extra].copy()
        loop(sh.phase, 10)
        g = po.GatherSim(box, t, n0 + extra)
        g.set_state(host.view(po.P32))
        g.phase(po.PH_HASH)  # the lists of the changed state
        loop(g.phase, 10)
        assert_bitexact(sh.state(), g.get_state(), what=f"after adding {extra} particles")


@pytest.mark.parametrize("lanes", [1, 4])
def test_coincident_particles_nudge_owner(lanes):
    """Particles closer than 1e-6 take the reference's nudge of +1e-6 on the
    half-list owner only (src/fluid.c:583-586): pairs in one cell (lower id
    owns) and across a cell border (the particle in the backward cell owns).
    With one and with four lanes per particle (the nudge is a flag of the
    quad's pair terms, applied in order)."""
    t = sc.preset_tunables(0)
    box = sa.make_box(10, 10)
    p = np.zeros(6, sa.PARTICLE)
    p["x"] = [2.5, 2.5, 4.9999995, 5.0, 7.2, 7.2000003]
    p["y"] = [3.5, 3.5, 6.5, 6.5, 1.3, 1.3]
    p["x_prev"], p["y_prev"] = p["x"], p["y"]
    p["id"] = [4, 1, 0, 5, 2, 3]
    s = sa.Solver(t, box, len(p))
    s.upload(p)
    s.set_gather_lanes(lanes)
    g = po.GatherSim(box, t, len(p))
    g.set_state(p.view(po.P32))
    for step in range(3):
        s.step(1)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"lanes {lanes}, step {step}")
