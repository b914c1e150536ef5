"""Dense tiles: the gather kernels' candidate paths on one scene.

A workgroup stages the candidate rows of its 256 sorted particles in LDS
when they fit the library's tile (1,024-2,048 records by build; the density
pass, whose records are positions only, twice as many in the same bytes); a
workgroup whose rows fit only one at a time stages and walks them row by row;
one that spans more than two cell rows, or whose rows do not fit even singly,
walks the same candidates in global memory.  All three visit the same chunks
in the same order.  A dense lattice (spacing 0.15h, about 44 particles per
cell) makes tiles of 1,050-1,220 records, and one workgroup spans several
rows; it is compared bit for bit with the oracle.  The same scene then runs on
the test build with 256-record tiles (nearly every workgroup row by row or on
the global-memory walk, XCD runs of one tile) in a child process, and must
equal the default library bit for bit.  An overfull lattice (spacing 0.06h,
about 280 particles per cell) stages the viscosity and relaxation rows one at
a time (the density pass most tiles whole), takes the reorder's ranking past
its staged ids, and exceeds both of the reference's caps; it too is compared
bit for bit, with the overflow counters.  A lattice at 0.1h (about 100
particles per cell) puts most tiles between the two capacities: density
staged whole, the other gathers row by row, bit for bit.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import sph_amd as sa
from oracle import pyoracle as po
from tests.common import assert_bitexact, dam_break_tunables

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SMALL = os.path.join(sa.PKG_DIR, "variants", "libsph_hip_t256_x1.so")

CHILD = r"""
import sys
import numpy as np
import sph_amd as sa
d = np.load(sys.argv[1])
parts = d["parts"]
s = sa.Solver(sa.Tunables.from_buffer_copy(d["tun"].tobytes()), sa.make_box(60, 40), len(parts))
s.upload(parts)
s.step(int(sys.argv[3]))
np.save(sys.argv[2], s.download())
"""


def dense_case():
    parts = sa.scene_lattice(128, 128, 0.15, 0.01, seed=7)
    return parts, dam_break_tunables(), po.make_box(60, 40)


def test_dense_tiles_bitexact_against_oracle():
    parts, tun, box = dense_case()
    s = sa.Solver(sa.Tunables.from_buffer_copy(bytes(tun)), sa.AABB.from_buffer_copy(bytes(box)),
                  len(parts))
    s.upload(parts)
    g = po.GatherSim(box, tun, len(parts))
    g.set_state(parts.view(po.P32))
    for step in range(1, 4):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"dense step {step}")
    c, oc = s.counters(), g.counters()
    assert c["max_neighbors"] == oc["max_neighbors"]
    assert c["neighbor_overflow"] == oc["neighbor_overflow"]


def test_overfull_cells_bitexact_against_oracle():
    """Spacing 0.06h: about 280 particles per cell (past the reference's
    100-slot buckets and 400-entry neighbour lists, both counted).  Cells this
    large reach past the ids the reorder stages around each workgroup (its
    positions +-64), so the ranking's global-memory path runs, and every tile
    is too large for the viscosity and relaxation passes' LDS tiles as a whole
    (they stage its candidate rows one at a time)."""
    parts = sa.scene_lattice(48, 48, 0.06, 0.01, seed=5)
    tun, box = dam_break_tunables(), po.make_box(60, 40)
    s = sa.Solver(sa.Tunables.from_buffer_copy(bytes(tun)), sa.AABB.from_buffer_copy(bytes(box)),
                  len(parts))
    s.upload(parts)
    g = po.GatherSim(box, tun, len(parts))
    g.set_state(parts.view(po.P32))
    for step in range(1, 4):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"overfull step {step}")
    s.step(2)  # the captured fused step
    g.step(1)
    # the device counts a step's post-relaxation lists in the next step's
    # viscosity pass, which walks exactly that set: the neighbour counts are
    # the oracle's before its last hash; the cell counts include every sort
    for ph in range(po.PH_GRAVITY, po.PH_HASH):
        g.phase(ph)
    before_last_hash = g.counters()
    g.phase(po.PH_HASH)
    oc = g.counters()
    assert_bitexact(s.download(), g.get_state(), what="overfull, graph steps 4-5")
    c = s.counters()
    assert c["max_cell_count"] == oc["max_cell_count"] > 64 + 64
    assert c["bucket_overflow"] == oc["bucket_overflow"] > 0
    for k in ("neighbor_overflow", "max_neighbors"):
        assert c[k] == before_last_hash[k], k
    assert before_last_hash["neighbor_overflow"] > 0


def test_tiles_between_the_density_and_record_capacities():
    """Spacing 0.1h: 16,384 particles, about 100 per cell, tiles of 1,200-1,800
    candidates - staged whole by the density pass, row by row by the
    viscosity and relaxation passes - bit-exact against the oracle."""
    parts = sa.scene_lattice(128, 128, 0.1, 0.01, seed=7)
    tun, box = dam_break_tunables(), po.make_box(60, 40)
    s = sa.Solver(sa.Tunables.from_buffer_copy(bytes(tun)), sa.AABB.from_buffer_copy(bytes(box)),
                  len(parts))
    s.upload(parts)
    g = po.GatherSim(box, tun, len(parts))
    g.set_state(parts.view(po.P32))
    for step in range(1, 4):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"0.1h lattice step {step}")
    s.step(2)
    g.step(2)
    assert_bitexact(s.download(), g.get_state(), what="0.1h lattice, graph steps 4-5")


def test_small_window_build_equals_default(tmp_path):
    assert os.path.exists(SMALL), "build the variants first (python -m sph_amd.build)"
    parts, tun, box = dense_case()
    steps = 6
    s = sa.Solver(sa.Tunables.from_buffer_copy(bytes(tun)), sa.AABB.from_buffer_copy(bytes(box)),
                  len(parts))
    s.upload(parts)
    s.step(steps)
    want = s.download()
    src, out = tmp_path / "in.npz", tmp_path / "out.npy"
    np.savez(src, parts=parts, tun=np.frombuffer(bytes(tun), np.uint8))
    env = dict(os.environ, SPH_AMD_LIB=SMALL)
    print("child: 256-record tiles,", len(parts), "particles,", steps, "steps", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, str(src), str(out), str(steps)], cwd=ROOT,
                       env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    got = np.load(out)
    assert_bitexact(got, want, what="256-record tiles vs default")
