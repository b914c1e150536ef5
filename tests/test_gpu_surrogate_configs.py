"""The learned macro-step at the configurations' own sizes.

* Training batch (64 sloshing dam breaks of 16,384 particles, 1,048,576 in
  all, taken from plain-solver trajectories as the trainer takes them):
  forward, loss, gradients and one Adam step of the whole batch on the
  device against the FP64 oracle.  The oracle evaluates 8 of the 64 scenes;
  the other 56 get their own predicted displacement as target, so their
  error - and their gradient - is exactly zero on the device, and the
  oracle's sum over the 8 scenes is the whole batch's gradient.  BF16
  tolerances as in test_gpu_surrogate.py.
* Inference workload (64 sloshing dam breaks of 65,536 particles, 4,194,304
  in all, the trained weights): judged macro-steps keep the judge's
  invariants - every scene plain on its first macro-step and then on the
  fixed schedule, flags in {0, 1, 2, 3}, shadow statistics counting every
  shadowed scene - and with tolerance 0 every macro-step falls back, which
  leaves exactly the plain run, at full size.
"""
import numpy as np
import pytest

import sph_amd as sa
from oracle import surrogate_ref as sr
from sph_amd import scenes as sc
from sph_amd import surrogate as su
from tests.common import assert_bitexact
from tests.test_gpu_surrogate import DT, device_weights, slosh_accel

pytestmark = pytest.mark.gpu

TREND_REFRESH = 16  # the library's schedule of plain macro-steps


def test_training_batch_64_scenes_of_16k():
    scenes = sc.sloshing_dam_breaks(64, n=16384, box=(128.0, 96.0), seed=7)
    s = su.make_batch(scenes)
    steps = 40
    s.step(steps)  # a state in mid-trajectory, as the trainer samples them
    n = 16384
    assert s.n == 64 * n
    m = sa.Surrogate(seed=9)
    p = sr.random_params(31).astype(np.float32)
    m.params = p
    d_dev = m.predict(s).astype(np.float64)
    state = s.download()
    check = [0, 9, 18, 27, 36, 45, 54, 63]
    rng = np.random.default_rng(8)
    target = d_dev.astype(np.float32).copy()  # zero error everywhere ...
    pd = device_weights(p)
    g_ref = np.zeros(sr.NPARAMS)
    loss_ref = 0.0
    for k in check:  # ... but on the checked scenes
        t, box, parts, amp, per = scenes[k]
        sl = slice(k * n, (k + 1) * n)
        q = state[sl]
        ax = slosh_accel(amp, per, steps, 0)
        args = (q["x"], q["y"], q["v_x"], q["v_y"], float(ax), -float(t.g),
                float(t.smoothing_radius), float(DT))
        d_ref, _, _ = sr.forward(pd, *args)
        err = np.abs(d_dev[sl] - d_ref) / sr.DS
        scale = np.abs(d_ref).max() / sr.DS
        assert err.max() <= 0.02 * (1 + scale), (k, err.max(), scale)
        assert np.sqrt((err ** 2).mean()) <= 0.005 * (1 + scale), k
        target[sl] = (d_ref + rng.normal(0, 0.02, d_ref.shape)).astype(np.float32)
        l, g = sr.loss_and_grads(pd, *args, target[sl], n_total=64 * n)
        loss_ref += l
        g_ref += g
        print(f"scene {k}: forward checked, {len(q)} particles", flush=True)
    m.zero_grads()
    loss = m.backward(s, target)
    g = m.grads().astype(np.float64)
    assert abs(loss - loss_ref) <= 0.01 * loss_ref, (loss, loss_ref)
    blocks = dict(W1=(sr.oW1, sr.oW2), W2=(sr.oW2, sr.ob2), b2=(sr.ob2, sr.oW3),
                  W3=(sr.oW3, sr.oW4), W4=(sr.oW4, sr.ob4), b4=(sr.ob4, sr.NPARAMS))
    for name, (a, b) in blocks.items():
        ga, gr = g[a:b], g_ref[a:b]
        rel = np.linalg.norm(ga - gr) / np.linalg.norm(gr)
        cos = ga @ gr / (np.linalg.norm(ga) * np.linalg.norm(gr))
        assert rel <= 0.05 and cos >= 0.998, (name, rel, cos)
    # one Adam step with lr 1e-3 (the training configuration)
    m.adam(1e-3)
    pp, _, _ = sr.adam(p.astype(np.float64), g, np.zeros(sr.NPARAMS), np.zeros(sr.NPARAMS), 1,
                       1e-3)
    np.testing.assert_allclose(m.params, pp, rtol=2e-6, atol=1e-7)
    assert s.counters()["nonfinite"] == 0


def c5_batch():
    scenes = sc.sloshing_dam_breaks(64, n=65536, box=(512.0, 96.0), seed=11, widen=4)
    s = su.make_batch(scenes)
    s.step(1)
    return s


def test_inference_workload_judge_invariants():
    s = c5_batch()
    assert s.nscenes == 64 and s.n == 64 * 65536
    m = sa.Surrogate(seed=0)
    su.load_weights(m)
    m.set_shadow(8)
    flags = m.macro_step(s, 0.01)
    assert np.all(flags == 2)  # first macro-step: plain everywhere
    shadowed = 0
    for k in range(1, 9):
        flags = m.macro_step(s, 0.01)
        print(f"macro-step {k}: flags {np.bincount(flags, minlength=4)}", flush=True)
        sched = (k + np.arange(64)) % TREND_REFRESH == 0
        assert np.all(flags[sched] == 2) and not np.any(flags[~sched] == 2), k
        assert set(np.unique(flags)) <= {0, 1, 2, 3}
        shadowed += int(np.sum(((k + np.arange(64)) % 8 == 0) & ~sched))
    st = m.shadow_stats()
    assert st["shadowed"] == shadowed > 0, st
    assert st["accepted"] <= st["shadowed"] and st["accepted_over_tol"] <= st["accepted"]
    assert 0.0 <= st["mean_abs_error"] <= st["max_abs_error"]
    c = s.counters()
    assert c["nonfinite"] == 0
    print("shadow statistics", st, flush=True)


def test_inference_workload_tolerance_zero_is_the_plain_run():
    ref = c5_batch()
    ref.step(12)
    want = ref.download()
    ref.close()
    s = c5_batch()
    m = sa.Surrogate(seed=0)
    su.load_weights(m)
    m.set_shadow(4)  # shadow steps on the way must not change the run either
    for k in range(3):
        flags = m.macro_step(s, 0.0)
        assert np.all(flags != 0), k
    assert_bitexact(s.download(), want, fields=("x", "y", "v_x", "v_y", "x_prev", "y_prev"),
                    what="tol 0, 3 macro-steps")
