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
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import PARTICLE, make_box, make_tunables, scene_lattice

_CFG = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs", "presets.json")


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
