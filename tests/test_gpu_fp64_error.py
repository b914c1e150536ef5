"""Error of the GPU's FP32 step against an FP64 CPU run of the same method.

The GPU is bit-exact against the FP32 oracle (test_gpu_parity.py); this file
reports how far that FP32 result is from the same gather-form step evaluated
in float64 (the oracle compiled with REAL = double), on the 4096-particle dam
break from the same initial state.  After one step the per-particle density
and near-density agree within 1e-5 relative, and positions within 4 ulp of
the largest coordinate (float32 positions near x = 30 carry 2e-6 absolute
rounding, which is 1e-3 of one step's typical displacement: the stored
position, not the step's arithmetic, bounds displacement accuracy).  Later
steps are reported, not bounded: the flow is chaotic and the FP32 and FP64
trajectories separate.
"""
import numpy as np
import pytest

import sph_amd as sa
from oracle import pyoracle as po
from tests.common import dam_break_tunables

pytestmark = pytest.mark.gpu


def rel_errors(got, want, x0, y0):
    out = {}
    for f in ("density", "density_near"):
        w = np.asarray(want[f], np.float64)
        g = np.asarray(got[f], np.float64)
        out[f] = float(np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-3)))
    wx, wy = np.asarray(want["x"], np.float64), np.asarray(want["y"], np.float64)
    dw = np.hypot(wx - x0, wy - y0)
    dg = np.hypot(np.asarray(got["x"], np.float64) - wx, np.asarray(got["y"], np.float64) - wy)
    out["position_abs_h"] = float(np.max(dg))
    out["position_ulps"] = float(np.max(dg) / (max(np.abs(wx).max(), np.abs(wy).max()) * 2.0 ** -23))
    out["displacement_rms_rel"] = float(np.sqrt(np.mean(dg ** 2)) / max(float(np.mean(dw)), 1e-12))
    return out


def test_fp32_step_error_against_fp64():
    tun = dam_break_tunables()
    box = po.make_box(60, 40)
    parts = sa.scene_lattice(64, 64, 0.5, 0.01, seed=12345)
    x0 = parts["x"].astype(np.float64)
    y0 = parts["y"].astype(np.float64)
    s = sa.Solver(sa.Tunables.from_buffer_copy(tun), sa.make_box(60, 40), len(parts))
    s.upload(parts)
    g = po.GatherSim(box, tun, len(parts), real="f64")
    g.set_state(np.ascontiguousarray(parts).view(po.P32))
    done = 0
    report = {}
    for k in (1, 10, 100):
        s.step(k - done, graph=False)
        g.step(k - done)
        done = k
        got = s.download()
        want = g.get_state()
        order = np.argsort(want["id"])
        want = want[order]
        report[k] = rel_errors(got, want, x0, y0)
        print(f"step {k}: " + ", ".join(f"{a} {b:.2e}" for a, b in report[k].items()))
    e1 = report[1]
    assert e1["density"] < 1e-5 and e1["density_near"] < 1e-5, e1
    assert e1["position_ulps"] <= 4.0, e1
