"""Lanes per particle of the fused step's gathers (sph_set_gather_lanes).

A context with at most two 256-slot tiles per CU (up to 131,072 particles
on 256 CUs: the C1 dam break, the dense test lattices, small batches, the
65k dam break) runs four lanes per particle by default: the quad
splits each 32-candidate chunk's tests four ways and takes four pairs per
drain step, and every lane applies the four pair terms in canonical order.
Larger contexts run one lane per particle.  Both shapes must give the
oracle's gather form bit for bit, so the small cases here run with the shape
forced each way: the dam break, a dense lattice staged in LDS with and
without a bucket-cap cut inside a row's second chunk, an overfull lattice
whose rows are staged one at a time or walked in global memory, a batch with
a frozen scene, and a switch of shape between captured steps.
"""
import numpy as np
import pytest

import sph_amd as sa
from sph_amd import scenes as sc
from oracle import pyoracle as po
from tests.common import assert_bitexact, dam_break_tunables, golden, parts_from_snap, tun_from_bytes

pytestmark = pytest.mark.gpu


def make_pair(parts, tun, box, lanes, cap=0):
    s = sa.Solver(sa.Tunables.from_buffer_copy(bytes(tun)), sa.AABB.from_buffer_copy(bytes(box)),
                  len(parts))
    s.upload(parts)
    s.set_gather_lanes(lanes)
    if cap:
        s.set_bucket_cap(cap)
    g = po.GatherSim(box, tun, len(parts))
    g.set_cap(cap)
    g.set_state(parts.view(po.P32) if parts.dtype != po.P32 else parts)
    return s, g


def c1_case():
    d = golden("c1_dambreak.npz")
    return parts_from_snap(d["input"]), tun_from_bytes(d["tunables"]), po.make_box(60, 40)


@pytest.mark.parametrize("lanes", [1, 4])
def test_c1_dam_break(lanes):
    parts, tun, box = c1_case()
    s, g = make_pair(parts, tun, box, lanes)
    for upto in (1, 2, 5, 10, 30, 60):
        n = upto - s.counters()["steps"]
        s.step(n)
        g.step(n)
        assert_bitexact(s.download(), g.get_state(), what=f"lanes {lanes}, step {upto}")
    c, oc = s.counters(), g.counters()
    assert c["max_neighbors"] == oc["max_neighbors"]
    assert c["max_cell_count"] == oc["max_cell_count"]


@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("cap", [0, 40])
def test_dense_lattice(lanes, cap):
    """About 44 particles per cell: rows of four and more 32-candidate chunks
    (past the four masks the relaxation replays), tiles of 1,050-1,220
    records; cap 40 cuts each cell inside its second chunk."""
    parts = sa.scene_lattice(128, 128, 0.15, 0.01, seed=7)
    s, g = make_pair(parts, dam_break_tunables(), po.make_box(60, 40), lanes, cap)
    for step in range(1, 3):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(),
                        what=f"dense, lanes {lanes}, cap {cap}, step {step}")
    s.step(2)
    g.step(2)
    assert_bitexact(s.download(), g.get_state(),
                    what=f"dense, lanes {lanes}, cap {cap}, graph steps 3-4")
    c, oc = s.counters(), g.counters()
    assert c["max_cell_count"] == oc["max_cell_count"]


@pytest.mark.parametrize("lanes", [1, 4])
def test_overfull_lattice(lanes):
    """About 280 particles per cell: rows staged one at a time (viscosity and
    relaxation) and both of the reference's caps exceeded (counted)."""
    parts = sa.scene_lattice(48, 48, 0.06, 0.01, seed=5)
    s, g = make_pair(parts, dam_break_tunables(), po.make_box(60, 40), lanes)
    for step in range(1, 3):
        s.step(1, graph=False)
        g.step(1)
        assert_bitexact(s.download(), g.get_state(), what=f"overfull, lanes {lanes}, step {step}")
    s.step(1)
    g.step(1)
    assert_bitexact(s.download(), g.get_state(), what=f"overfull, lanes {lanes}, graph step 3")
    c = s.counters()
    assert c["bucket_overflow"] > 0 and c["neighbor_overflow"] > 0


@pytest.mark.parametrize("lanes", [1, 4])
def test_batch_with_a_frozen_scene(lanes):
    t = sc.preset_tunables(1)
    box = sa.make_box(24, 20)
    ps = [sa.scene_lattice(24, 16, 0.5, 0.01, seed=30 + i, x0=0.25, y0=0.25) for i in range(3)]
    counts = [len(p) for p in ps]
    s = sa.Solver.batch([(t, box, len(p)) for p in ps])
    s.upload_batch(np.concatenate(ps), counts)
    s.set_gather_lanes(lanes)
    s.step(3)
    s.set_frozen([0, 1, 0])
    s.step(5)
    offs = np.concatenate([[0], np.cumsum(counts)])
    got = s.download()
    for i in (0, 2):
        g = po.GatherSim(box, t, len(ps[i]))
        g.set_state(ps[i])
        g.step(8)
        assert_bitexact(got[offs[i]:offs[i + 1]], g.get_state(), what=f"lanes {lanes}, scene {i}")
    g = po.GatherSim(box, t, len(ps[1]))
    g.set_state(ps[1])
    g.step(3)
    frozen = got[offs[1]:offs[2]]
    for f in ("x", "y", "v_x", "v_y"):
        assert np.array_equal(frozen[f], g.get_state()[f]), f


def test_switching_lanes_between_captured_steps():
    """Each change takes the other kernels (the captured steps are rebuilt);
    0 goes back to the size rule."""
    parts, tun, box = c1_case()
    s, g = make_pair(parts, tun, box, 0)
    for k, lanes in enumerate((0, 1, 4, 1, 0)):
        s.set_gather_lanes(lanes)
        s.step(4)
        g.step(4)
        assert_bitexact(s.download(), g.get_state(), what=f"segment {k}, lanes {lanes}")


def test_invalid_lanes_rejected():
    parts, tun, box = c1_case()
    s, _ = make_pair(parts, tun, box, 0)
    for bad in (2, 3, 8, -1):
        with pytest.raises(sa.SphError):
            s.set_gather_lanes(bad)
