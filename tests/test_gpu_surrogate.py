"""The learned macro-step on the GPU against the FP64 oracle.

BF16 weights and activations on MFMA make the device forward and gradients
approximate: the tolerances below are BF16-sized (8-bit mantissa, errors
compounding over three layers) and are stated per check.  The solver-side
pieces the surrogate relies on (snapshots, device state load/gather) are
bit-exact checks.
"""
import numpy as np
import pytest

import sph_amd as sa
from sph_amd import scenes as sc
from oracle import surrogate_ref as sr
from tests.common import assert_bitexact

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def bf16_round(a):
    """float32 values rounded to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(a, np.float32).copy()
    b = a.view(np.uint32)
    b += 0x7FFF + ((b >> 16) & 1)
    b &= 0xFFFF0000
    return b.view(np.float32).astype(np.float64)


def device_weights(p):
    """The weights the device multiplies with: bf16 copies of W1, W2 and W3
    (the fp32 master copy for the rest).  Evaluating the oracle with them
    leaves only the activations' bf16 rounding between device and oracle;
    with the fp32 weights, the rounding of the weights on near-constant
    inputs (the constant 1) shifts every edge of a particle alike and biases
    its sum."""
    q = np.asarray(p, np.float64).copy()
    for lo, hi in ((sr.oW1, sr.oW2), (sr.oW2, sr.ob2), (sr.oW3, sr.oW4)):
        q[lo:hi] = bf16_round(q[lo:hi])
    return q


def slosh_accel(amp, period, clock, t0, dt=DT):
    """The device's sloshing acceleration in float32."""
    if amp == 0:
        return np.float32(0)
    omega = np.float32(2.0 * np.pi / period)
    arg = np.float32(np.float32(omega * np.float32(clock + t0)) * dt)
    return np.float32(-np.float32(amp) * np.sin(arg, dtype=np.float32))


def two_scene_batch(steps=5, spacing=0.5):
    """Two scenes (presets 0 and 1, the second sloshing with a clock
    offset); spacing 0.2h packs ~78 neighbours within h per particle, which
    makes the edge passes flush their LDS edge buffer several times per
    16-particle group."""
    t0, t1 = sc.preset_tunables(0), sc.preset_tunables(1)
    b0, b1 = sa.make_box(20, 16), sa.make_box(24, 18)
    j = 0.1 * spacing
    p0 = sa.scene_lattice(24, 20, spacing, j, seed=1, x0=0.25, y0=0.25)
    p1 = sa.scene_lattice(30, 16, spacing, j, seed=2, x0=0.25, y0=0.25)
    s = sa.Solver.batch([(t0, b0, len(p0)), (t1, b1, len(p1))])
    s.upload_batch(np.concatenate([p0, p1]), [len(p0), len(p1)])
    s.set_scene_forcing(1, 1.5, 2.0)
    s.set_scene_time_offset(1, 7)
    s.step(steps)
    return s, [len(p0), len(p1)], [(t0, 0.0, 1.0, 0), (t1, 1.5, 2.0, 7)], steps


def oracle_batch(params, state, counts, meta, clock, target=None):
    """Per-scene oracle forward (and gradients when target is given)."""
    offs = np.concatenate([[0], np.cumsum(counts)])
    ds, grads, loss = [], np.zeros(sr.NPARAMS), 0.0
    ntot = int(offs[-1])
    for k, (t, amp, per, t0) in enumerate(meta):
        q = state[offs[k]:offs[k + 1]]
        ax = slosh_accel(amp, per, clock, t0)
        args = (q["x"], q["y"], q["v_x"], q["v_y"], float(ax), -float(t.g), float(t.smoothing_radius),
                float(DT))
        d, _, _ = sr.forward(params, *args)
        ds.append(d)
        if target is not None:
            l, g = sr.loss_and_grads(params, *args, target[offs[k]:offs[k + 1]], n_total=ntot)
            loss += l
            grads += g
    return np.concatenate(ds), grads, loss


def test_mfma_lane_layout():
    D = sa.mfma_selftest()
    r = np.arange(16)
    A = ((r[:, None] * 16 + r[None, :]) % 7 - 3).astype(np.float64)
    B = ((r[:, None] * 3 + r[None, :]) % 5 - 2).astype(np.float64)
    np.testing.assert_array_equal(D, A @ B)


@pytest.mark.parametrize("spacing", [0.5, 0.2])
def test_forward_matches_oracle(spacing):
    s, counts, meta, clock = two_scene_batch(spacing=spacing)
    m = sa.Surrogate(seed=1)
    p = sr.random_params(21).astype(np.float32)
    m.params = p
    got = m.predict(s).astype(np.float64)
    state = s.download()
    want, _, _ = oracle_batch(device_weights(p), state, counts, meta, clock)
    # compare the network part o = (d - ballistic) / 0.05
    err = np.abs(got - want) / sr.DS
    scale = np.abs(want).max() / sr.DS
    # BF16 tolerance: 2% of the output scale in the worst particle, 0.5% rms
    assert err.max() <= 0.02 * (1 + scale), (err.max(), scale)
    assert np.sqrt((err ** 2).mean()) <= 0.005 * (1 + scale)


@pytest.mark.parametrize("spacing", [0.5, 0.2])
def test_gradients_match_oracle(spacing):
    s, counts, meta, clock = two_scene_batch(spacing=spacing)
    m = sa.Surrogate(seed=2)
    p = sr.random_params(22).astype(np.float32)
    m.params = p
    state = s.download()
    pd = device_weights(p)
    want_d, _, _ = oracle_batch(pd, state, counts, meta, clock)
    rng = np.random.default_rng(5)
    target = (want_d + rng.normal(0, 0.02, want_d.shape)).astype(np.float32)
    _, g_ref, loss_ref = oracle_batch(pd, state, counts, meta, clock, target)
    m.zero_grads()
    loss = m.backward(s, target)
    g = m.grads().astype(np.float64)
    # loss to 1% (BF16 forward), every parameter block's gradient to 5% in
    # norm and to cosine similarity 0.998
    assert abs(loss - loss_ref) <= 0.01 * loss_ref, (loss, loss_ref)
    blocks = dict(W1=(sr.oW1, sr.oW2), W2=(sr.oW2, sr.ob2), b2=(sr.ob2, sr.oW3), W3=(sr.oW3, sr.oW4),
                  W4=(sr.oW4, sr.ob4), b4=(sr.ob4, sr.NPARAMS))
    for name, (a, b) in blocks.items():
        ga, gr = g[a:b], g_ref[a:b]
        rel = np.linalg.norm(ga - gr) / np.linalg.norm(gr)
        cos = ga @ gr / (np.linalg.norm(ga) * np.linalg.norm(gr))
        assert rel <= 0.05 and cos >= 0.998, (name, rel, cos)
    # two accumulations give twice the gradient
    m.backward(s, target)
    np.testing.assert_allclose(m.grads(), 2 * g, rtol=1e-4, atol=1e-7)


def test_adam_matches_oracle():
    s, counts, meta, clock = two_scene_batch()
    m = sa.Surrogate(seed=3)
    p = sr.random_params(23).astype(np.float32)
    m.params = p
    pm = np.zeros(sr.NPARAMS)
    pv = np.zeros(sr.NPARAMS)
    pp = p.astype(np.float64)
    target = np.zeros((sum(counts), 2), np.float32)
    for t in (1, 2):
        m.zero_grads()
        m.backward(s, target)
        g = m.grads().astype(np.float64)
        m.adam(1e-3)
        pp, pm, pv = sr.adam(pp, g, pm, pv, t, 1e-3)
        np.testing.assert_allclose(m.params, pp, rtol=2e-6, atol=1e-7)
        pp = m.params.astype(np.float64)  # keep the fp32 trajectory


def test_apply_moves_by_prediction():
    s, counts, meta, clock = two_scene_batch()
    m = sa.Surrogate(seed=4)
    d = m.predict(s)
    before = s.download()
    m.apply(s)
    after = s.download()
    offs = np.concatenate([[0], np.cumsum(counts)])
    for k, (t, amp, per, t0) in enumerate(meta):
        b, a, dd = before[offs[k]:offs[k + 1]], after[offs[k]:offs[k + 1]], d[offs[k]:offs[k + 1]]
        box = (20, 16) if k == 0 else (24, 18)
        x = np.float32(b["x"] + dd[:, 0])
        y = np.float32(b["y"] + dd[:, 1])
        mx = np.float32(box[0]) - np.float32(0.001)
        my = np.float32(box[1]) - np.float32(0.001)
        x = np.where(x < 0, np.float32(0), np.where(x > box[0], mx, x))
        y = np.where(y < 0, np.float32(0), np.where(y > box[1], my, y))
        assert np.array_equal(a["x"], x.astype(np.float32)) and np.array_equal(a["y"], y.astype(np.float32))
        ax = slosh_accel(amp, per, clock, t0)
        vx = np.clip(dd[:, 0] / (4 * DT) + np.float32(1.5) * ax * DT, -5, 5)
        vy = np.clip(dd[:, 1] / (4 * DT) + np.float32(1.5) * np.float32(-t.g) * DT, -5, 5)
        np.testing.assert_allclose(a["v_x"], vx, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(a["v_y"], vy, rtol=1e-5, atol=1e-5)


def test_snapshot_restore_per_scene():
    t = sc.preset_tunables(2)
    box = sa.make_box(22, 18)
    ps = [sa.scene_lattice(20, 18, 0.5, 0.02, seed=50 + i, x0=0.25, y0=0.25) for i in range(3)]
    counts = [len(p) for p in ps]

    def fresh():
        s = sa.Solver.batch([(t, box, n) for n in counts])
        s.upload_batch(np.concatenate(ps), counts)
        s.set_scene_forcing(0, 1.0, 1.5)
        return s

    ref = fresh()
    ref.step(7)
    want = ref.download()
    s = fresh()
    s.step(3)
    s.snapshot_save()
    s.step(4)
    at7 = s.download()
    assert_bitexact(at7, want, what="straight run")
    s.snapshot_restore([1, 0, 1])
    s.set_frozen([0, 1, 0])
    s.step(4)
    s.set_frozen([0, 0, 0])
    # a frozen scene keeps x, y, v (x_prev and density are not carried)
    assert_bitexact(s.download(), want, fields=("x", "y", "v_x", "v_y"),
                    what="restored + re-taken scenes")


def test_trajectory_load_continues_bit_exact():
    """A state recorded in the trajectory store and loaded into another
    context (ids renumbered in storage order, sloshing clock offset set)
    continues exactly as the original run; targets are frame differences."""
    t = sc.preset_tunables(1)
    box = sa.make_box(22, 18)
    ps = [sa.scene_lattice(22, 16, 0.5, 0.02, seed=60 + i, x0=0.25, y0=0.25) for i in range(2)]
    n = len(ps[0])
    s = sa.Solver.batch([(t, box, n) for _ in ps])
    s.upload_batch(np.concatenate(ps), [n, n])
    s.set_scene_forcing(1, 2.0, 3.0)
    tr = sa.Trajectory(2, n, 3)
    s.step(6)
    tr.record(s, 0)
    st = s.download()
    f0 = tr.read(0)
    assert np.array_equal(f0[:, :, 0].ravel(), st["x"]) and np.array_equal(f0[:, :, 3].ravel(), st["v_y"])
    s.step(4)
    tr.record(s, 1)
    s.step(4)
    tr.record(s, 2)
    want = s.download()
    # load frame 1 of scene 1 and frame 0 of scene 0 into a fresh context
    s2 = sa.Solver.batch([(t, box, n) for _ in ps])
    s2.upload_batch(np.concatenate(ps), [n, n])
    s2.set_scene_forcing(0, 2.0, 3.0)
    tr.load(s2, [1, 0], [1, 0])
    s2.set_scene_time_offset(0, 10)
    s2.set_scene_time_offset(1, 6)
    s2.phase("hash")
    s2.step(4)
    got = s2.download()
    for f in ("x", "y", "v_x", "v_y"):  # scene 1 at step 14 == original scene 1
        assert np.array_equal(got[f][:n], want[f][n:]), f
    f2 = tr.read(2)
    assert np.array_equal(got["x"][:n], f2[1, :, 0])
    # scene 0 of s2 is the original scene 0 at step 10 (= frame 1)
    assert np.array_equal(got["x"][n:], tr.read(1)[0, :, 0])


def test_macro_step_judge_and_fallback():
    """The first macro-step is plain by construction; tol = 0 then rejects
    every moving scene: each is restored and re-taken by four plain steps,
    so the result equals a plain run bit for bit."""
    t = sc.preset_tunables(0)
    box = sa.make_box(22, 18)
    ps = [sa.scene_lattice(20, 16, 0.5, 0.02, seed=70 + i, x0=0.25, y0=0.25) for i in range(2)]
    counts = [len(p) for p in ps]

    def fresh():
        s = sa.Solver.batch([(t, box, n) for n in counts])
        s.upload_batch(np.concatenate(ps), counts)
        return s

    ref = fresh()
    ref.step(10)
    s = fresh()
    s.step(2)
    m = sa.Surrogate(seed=5)
    flags = m.macro_step(s, tol=0.0)
    assert list(flags) == [2, 2]  # the first macro-step of a scene is plain
    flags = m.macro_step(s, tol=0.0)
    assert list(flags) == [1, 1]
    # the judge's density pass overwrites density; the rest is the plain run's
    assert_bitexact(s.download(), ref.download(), fields=("x", "y", "v_x", "v_y", "x_prev", "y_prev"),
                    what="fallback run")
    # a huge tolerance accepts every scene: the surrogate moved them
    flags = m.macro_step(s, tol=1e9)
    assert list(flags) == [0, 0]


def test_trained_network_beats_method_without_it():
    """Train on plain-solver trajectories of small sloshing dam breaks (the
    training pipeline at reduced size) and compare, on held-out scenes, the
    4-step displacement error of the network with that of the method without
    it (ballistic extrapolation: the network's output layer zeroed)."""
    from sph_amd import surrogate as su
    scn = sc.sloshing_dam_breaks(40, n=4096, box=(64.0, 48.0), seed=3)
    train_s, hold_s = scn[:32], scn[32:]
    tr, c = su.generate(train_s, steps=80)
    assert c["nonfinite"] == 0
    hold, _ = su.generate(hold_s, steps=80)
    m = sa.Surrogate(seed=1)
    hist = su.train(m, tr, train_s, iters=400, batch=8, lr=1e-3, seed=2)
    assert np.mean(hist[-50:]) < 0.5 * np.mean(hist[:10])
    net, bal = su.evaluate(m, hold, hold_s, frames=[0, 5, 10, 15, 19], batch=8)
    print(f"held-out 4-step rms: network {net:.5f} h, without it {bal:.5f} h")
    assert net < 0.6 * bal, (net, bal)
