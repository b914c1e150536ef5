"""Shared helpers of the test suite: fixtures, scenes and comparisons."""
import os

import numpy as np

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SNAP_FIELDS = ("x", "y", "v_x", "v_y", "x_prev", "y_prev", "density", "density_near")


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def tun_from_bytes(arr):
    return po.Tunables.from_buffer_copy(arr.tobytes())


def parts_from_snap(snap):
    p = np.zeros(snap.shape[1], po.P32)
    for k, f in enumerate(SNAP_FIELDS):
        p[f] = snap[k]
    p["id"] = np.arange(snap.shape[1])
    return p


def dam_break_tunables(**kw):
    """The benchmark configuration: rho0=10, h=1, g=9.8 (applied as -g,
    src/fluid.c:403), dt=1/60, k=0.004, k_near=0.01, sigma=0, beta=0.3,
    k_spring=10 (the reference default, src/fluid.c:94), sphere mover of
    width 0 (no mover: d2 <= 0 && d2 > 0 never holds, src/fluid.c:674)."""
    base = dict(rest_density=10.0, smoothing_radius=1.0, g=9.8, k=0.004, k_near=0.01,
                k_spring=10.0, sigma=0.0, beta=0.3, time_step=1.0 / 60.0, mover_type=0,
                mover_width=0.0, mover_height=0.0, active=1)
    base.update(kw)
    return po.make_tunables(**base)


def bits(a):
    a = np.ascontiguousarray(a, np.float32)
    return a.view(np.uint32)


def assert_bitexact(got, want, fields=SNAP_FIELDS, what=""):
    """Bit-for-bit equality per field (NaN-safe; +0 == -0 is not required
    to differ, so compare with == as well as bit patterns of non-zeros)."""
    for f in fields:
        g, w = np.asarray(got[f], np.float32), np.asarray(want[f], np.float32)
        same = (bits(g) == bits(w)) | ((g == 0) & (w == 0))
        if not same.all():
            idx = np.nonzero(~same)[0]
            raise AssertionError(
                f"{what} field {f}: {len(idx)} of {len(g)} differ; first ids {idx[:5]} "
                f"got {g[idx[:3]]} want {w[idx[:3]]}")


def max_rel(got, want, floor=1e-6):
    g = np.asarray(got, np.float64)
    w = np.asarray(want, np.float64)
    return float(np.max(np.abs(g - w) / np.maximum(np.abs(w), floor))) if len(g) else 0.0


# Phase indices of the reference loop body (src/fluid.c:273-341).
FIRST_HALF = (po.PH_GRAVITY, po.PH_VISCOSITY, po.PH_PREDICT)
SECOND_HALF = (po.PH_HASH_DENSITY, po.PH_RELAX, po.PH_UPDATE_VELOCITIES, po.PH_HASH)


def run_reference_loop(phase, set_tunables, nsteps, frames, steps_per_frame=4, first_step=0):
    """Drive phase(ph) through nsteps steps of the reference loop
    (src/fluid.c:270-372): on the last substep of every frame the next of
    `frames` (tunables) arrives between predict_positions and hash_fluid, as
    the render rank's MPI_Scatterv delivers it (src/fluid.c:293-294).
    first_step counts the steps run before (it sets the substep and the
    frame).  Returns the frames consumed so far."""
    frame = (first_step + 1) // steps_per_frame if first_step else 0
    for step in range(first_step, first_step + nsteps):
        for ph in FIRST_HALF:
            phase(ph)
        if step % steps_per_frame == steps_per_frame - 1 and frame < len(frames):
            set_tunables(frames[frame])
            frame += 1
        for ph in SECOND_HALF:
            phase(ph)
    return frame
