"""Seeded scene generators of the benchmark configurations.

* ``dam_break(nx, ny, box, seed)`` - the plain dam break: an nx x ny lattice
  at spacing 0.5h in the lower-left of the box (src/geometry.c:29-67 order,
  id = index), +-0.01h seeded jitter, v = 0.
* ``batched_scenes(n, seed)`` - the batched configuration: particle counts
  log-uniform in [2048, 65536], box aspect U[1, 2], fluid fraction
  U[0.2, 0.5], one of four parameter presets (configs/presets.json) each;
  box area = N (0.5h)^2 / fraction, W / H = aspect, the fluid a block of the
  box's aspect in its lower-left corner.
* ``sloshing_dam_breaks(nscenes, n, box, seed)`` - the learned macro-step's
  training and inference scenes: n particles, the first n of an nx-wide
  lattice (nx seeded, so blocks range from tall to wide) in the lower-left
  of the box, preset 0, and a seeded horizontal box acceleration of
  amplitude U[0, 2] and period U[1, 4] s.  With ``widen`` the blocks (and
  the box) are that many times wider at the same depth, which keeps large
  scenes in the training scenes' regime.
* ``reference_scene(preset)`` - the reference's own demo scene and parameter
  sets: its defaults (src/fluid.c:88-107), its sizing of box, spacing and
  smoothing radius (src/fluid.c:116-159, src/geometry.c:101-160 on one
  rank) and its lattice (src/geometry.c:29-67), with the presets x / y / a
  / b of src/controls.c:344-401 ("standard fluid", "super goo", "zero g
  high surface tension", "spring gas").
* ``attract_mover_path(tunables, box, frames)`` - the per-frame mover
  centres of the reference's idle "attract mode" (src/renderer.c:493-531),
  which sweeps the sphere mover along a sine through the fluid.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import PARTICLE, make_box, make_tunables, scene_lattice

_CFG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs", "presets.json")


# src/fluid.c:88-107: the compute rank's initial tunables.  The frame time
# step 1/30 is split over steps_per_frame = 4 substeps off the Pi
# (src/fluid.c:102-107).
REFERENCE_DEFAULTS = dict(g=6.0, k=0.2, k_near=6.0, k_spring=10.0, sigma=5.0, beta=0.5,
                          rest_density=30.0, mover_width=2.0, mover_height=2.0, mover_type=0,
                          kill_sim=0, active=1)
REFERENCE_STEPS_PER_FRAME = 4

# src/controls.c:344-401: the presets bound to the controller's x / y / a / b
# buttons (the fields each one sets; the rest keep their values).
REFERENCE_PRESETS = {
    "x": dict(g=6.0, k=0.2, k_near=6.0, k_spring=10.0, sigma=5.0, beta=0.5, rest_density=30.0),
    "y": dict(g=6.0, k=0.1, k_near=3.0, k_spring=-30.0, sigma=100.0, beta=10.0,
              rest_density=30.0),
    "a": dict(g=0.0, k=0.2, k_near=6.0, k_spring=10.0, sigma=20.0, beta=2.0, rest_density=55.0),
    "b": dict(g=6.0, k=0.0, k_near=0.0, k_spring=115.0, sigma=20.0, beta=2.0, rest_density=0.0),
}


def reference_time_step(steps_per_frame=REFERENCE_STEPS_PER_FRAME):
    """1.0f/30.0f divided by (float)steps_per_frame, in float (fluid.c:91, 106)."""
    return np.float32(np.float32(1.0) / np.float32(30.0)) / np.float32(steps_per_frame)


def reference_scene(preset="x", n_global=1500, pixel_dims=(1920, 1080), max_x=15.0):
    """The reference's demo scene on one compute rank.

    src/fluid.c:116-133: box [0, 15] x [0, 15 / aspect], aspect = the render
    window's width / height; the water fills the whole box.
    src/fluid.c:141-144: spacing = pow(area / N, 1/2) (double, stored as a
    float).  src/geometry.c:101-160 with one rank: floor(width / spacing) + 1
    columns, floor(height / spacing) rows, node_start_x / node_end_x the box
    edges.  src/fluid.c:159: h = 2 spacing.  src/fluid.c:156: room for
    2 N_global particles.  src/geometry.c:29-67: the lattice x = i spacing,
    y = j spacing, row-major, id = index, v = a = 0.

    Returns (tunables, box, particles, info) with info holding spacing,
    capacity (the reference's max_fluid_particles_local) and the lattice
    shape.  The mover centre is not set by the compute rank (the render rank
    sends it every frame); it starts at the first point of the attract-mode
    path (attract_mover_path).
    """
    import math
    f32 = np.float32
    aspect = f32(pixel_dims[0]) / f32(pixel_dims[1])
    mx = f32(max_x)
    my = mx / aspect
    area = (mx - f32(0.0)) * (my - f32(0.0))
    spacing = f32(math.pow(float(area / f32(n_global)), 1.0 / 2.0))
    nx = int(math.floor(float((mx - f32(0.0)) / spacing))) + 1
    ny = int(math.floor(float((my - f32(0.0)) / spacing)))
    h = f32(2.0) * spacing
    kw = dict(REFERENCE_DEFAULTS)
    kw.update(REFERENCE_PRESETS[preset])
    kw.update(smoothing_radius=float(h), time_step=float(reference_time_step()),
              node_start_x=0.0, node_end_x=float(mx))
    box = make_box(mx, my)
    cx, cy = attract_mover_path(None, box, 1)[0]
    kw.update(mover_center_x=float(cx), mover_center_y=float(cy))
    parts = scene_lattice(nx, ny, float(spacing), 0.0, seed=0)
    info = {"spacing": float(spacing), "h": float(h), "nx": nx, "ny": ny, "n": nx * ny,
            "capacity": 2 * n_global, "aspect": float(aspect), "box": (float(mx), float(my))}
    return make_tunables(**kw), box, parts, info


def attract_mover_path(tunables, box, frames, gl_x0=-0.95, direction=1):
    """Mover centres of the reference's idle mode, one per frame
    (src/renderer.c:493-531 with src/renderer.c:397-414): each frame the
    centre, in OpenGL coordinates [-1, 1], moves by 0.01 in x, turns round
    past either edge, and takes y = sin(3.14 * 5 x) / 10 - 0.6; in float.
    gl_x0 is where the sweep starts.  Returns a frames x 2 float32 array of
    simulation coordinates."""
    f32 = np.float32
    hw, hh = f32(box.max_x) * f32(0.5), f32(box.max_y) * f32(0.5)
    gl_x = f32(gl_x0)
    out = np.zeros((frames, 2), np.float32)
    for f in range(frames):
        gl_x = f32(gl_x + f32(0.01) * f32(direction))
        if gl_x > f32(1.0) or gl_x < f32(-1.0):
            direction = -direction
        gl_y = f32(np.sin(f32(3.14) * f32(5.0) * gl_x, dtype=np.float32) / f32(10.0)) - f32(0.6)
        out[f] = (gl_x * hw + hw, gl_y * hh + hh)
        gl_x = f32(out[f, 0] / hw - f32(1.0))
    return out


def reference_frames(base, box, frames, cycle=("x", "y", "a", "b"), frames_per_preset=6):
    """The tunables a render rank sends on frames 0..frames-1 (received by the
    compute rank at src/fluid.c:293-294): the sphere mover follows the
    attract-mode path, and the presets of `cycle` take turns every
    frames_per_preset frames, as a user pressing the controller's buttons
    would (src/controls.c:344-401).  base is the run's initial tunables."""
    path = attract_mover_path(None, box, frames)
    out = []
    for f in range(frames):
        t = type(base).from_buffer_copy(bytes(base))
        for k, v in REFERENCE_PRESETS[cycle[(f // frames_per_preset) % len(cycle)]].items():
            setattr(t, k, float(v))
        t.mover_center_x = float(path[f, 0])
        t.mover_center_y = float(path[f, 1])
        out.append(t)
    return out


def presets():
    with open(_CFG) as f:
        cfg = json.load(f)
    out = []
    for p in cfg["presets"]:
        kw = dict(cfg["base"])
        kw.update(p)
        out.append(kw)
    return out


def preset_tunables(i):
    return make_tunables(**presets()[i])


def dam_break(nx=1024, ny=1024, box=(1200.0, 800.0), seed=12345, spacing=0.5, jitter=0.01):
    """Tunables, box and particles of the plain dam break (preset 0)."""
    return preset_tunables(0), make_box(*box), scene_lattice(nx, ny, spacing, jitter, seed=seed)


def batched_scenes(nscenes=256, seed=2026, nmin=2048, nmax=65536, spacing=0.5, jitter=0.01,
                   only=None):
    """List of (tunables, box, particles, preset) for the batched config
    (with `only`, just those scene indices, drawn from the same sequence)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nscenes):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        aspect = float(rng.uniform(1.0, 2.0))
        frac = float(rng.uniform(0.2, 0.5))
        preset = int(rng.integers(0, 4))
        area = n * spacing * spacing / frac
        w = math.sqrt(area * aspect)
        h = w / aspect
        nx = int(math.ceil(math.sqrt(n * aspect)))
        ny = int(math.ceil(n / nx))
        if only is not None and s not in only:
            continue
        parts = scene_lattice(nx, ny, spacing, jitter, seed=seed * 100003 + s,
                              x0=spacing / 2, y0=spacing / 2)[:n].copy()
        assert parts["id"][-1] == n - 1
        out.append((preset_tunables(preset), make_box(np.float32(w), np.float32(h)), parts, preset))
    return out


def concat(scenes):
    """Concatenated particle array and counts for sph_upload_batch."""
    parts = np.concatenate([p for _, _, p, _ in scenes]) if scenes else np.zeros(0, PARTICLE)
    counts = np.array([len(p) for _, _, p, _ in scenes], np.int32)
    return parts, counts


def sloshing_dam_breaks(nscenes, n=16384, box=(128.0, 96.0), seed=7, spacing=0.5, jitter=0.01,
                        amp=(0.0, 2.0), period=(1.0, 4.0), widen=1):
    """List of (tunables, box, particles, amplitude, period).  widen > 1
    makes scenes of n particles whose fluid blocks are `widen` times wider
    than those of n / widen particles (same depth, same local physics) in a
    box `widen` times wider."""
    rng = np.random.default_rng(seed)
    out = []
    n0 = n // widen
    wmax = box[0] / widen * 0.75 / spacing
    for s in range(nscenes):
        nx_lo = int(math.ceil(math.sqrt(n0) * 0.75))
        nx_hi = int(min(wmax, math.sqrt(n0) * 1.5))
        nx = int(rng.integers(nx_lo, max(nx_lo + 1, nx_hi))) * widen
        ny = int(math.ceil(n / nx))
        assert ny * spacing < box[1] * 0.9, "fluid block taller than the box"
        parts = scene_lattice(nx, ny, spacing, jitter, seed=seed * 1000003 + s,
                              x0=spacing / 2, y0=spacing / 2)[:n].copy()
        a = float(rng.uniform(*amp))
        T = float(rng.uniform(*period))
        out.append((preset_tunables(0), make_box(*box), parts, a, T))
    return out
