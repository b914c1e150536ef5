"""Pins of the FP64 network oracle (oracle/surrogate_ref.py).

The network has no counterpart in the reference, so the oracle is pinned by
mathematics: its backward against central finite differences of its
forward, its forward against the textbook per-edge form (the second edge
layer applied on every edge before the sum), against the ballistic closed
form when the output layer is zero, and by translation / relabelling
invariance; Adam's first step against its closed form.
"""
import numpy as np
import pytest

from oracle import surrogate_ref as sr


def small_scene(n_side=7, seed=0, spacing=0.5):
    rng = np.random.default_rng(seed)
    g = np.arange(n_side, dtype=np.float64) * spacing
    X, Y = np.meshgrid(g, g)
    x = (X.ravel() + 2.0 + rng.uniform(-0.1, 0.1, X.size)).astype(np.float32)
    y = (Y.ravel() + 2.0 + rng.uniform(-0.1, 0.1, X.size)).astype(np.float32)
    vx = rng.uniform(-2, 2, X.size).astype(np.float32)
    vy = rng.uniform(-2, 2, X.size).astype(np.float32)
    return x, y, vx, vy


def test_pairs_match_brute_force():
    x, y, _, _ = small_scene(9, 1)
    i, j = sr.neighbour_pairs(x, y, 1.0)
    got = set(zip(i.tolist(), j.tolist()))
    want = set()
    for a in range(len(x)):
        for b in range(len(x)):
            if a != b:
                dx = np.float32(x[b] - x[a])
                dy = np.float32(y[b] - y[a])
                if not (np.float32(dx * dx + dy * dy) > np.float32(1.0)):
                    want.add((a, b))
    assert got == want


def test_forward_equals_per_edge_form():
    x, y, vx, vy = small_scene(6, 2)
    p = sr.random_params(3)
    _, o, _ = sr.forward(p, x, y, vx, vy, 0.7, -9.8, 1.0, 1 / 60)
    o2 = sr.forward_per_edge(p, x, y, vx, vy, 0.7, -9.8, 1.0, 1 / 60)
    np.testing.assert_allclose(o, o2, rtol=1e-12, atol=1e-12)


def test_zero_output_layer_is_ballistic():
    x, y, vx, vy = small_scene(5, 4)
    d0 = sr.unpack(sr.random_params(5))
    d0["W4"][:] = 0
    d0["b4"][:] = 0
    dt = 1 / 60
    d, _, _ = sr.forward(sr.pack(d0), x, y, vx, vy, 1.25, -9.8, 1.0, dt)
    np.testing.assert_allclose(d[:, 0], 4 * dt * vx.astype(np.float64) + 10 * dt * dt * 1.25)
    np.testing.assert_allclose(d[:, 1], 4 * dt * vy.astype(np.float64) - 10 * dt * dt * 9.8)


def test_translation_and_relabelling_invariance():
    rng = np.random.default_rng(7)
    n = 64
    # coordinates on a 1/1024 grid: exact in float32 before and after the shift
    x = (rng.integers(0, 8 * 1024, n) / 1024.0).astype(np.float32)
    y = (rng.integers(0, 8 * 1024, n) / 1024.0).astype(np.float32)
    vx = rng.uniform(-1, 1, n).astype(np.float32)
    vy = rng.uniform(-1, 1, n).astype(np.float32)
    p = sr.random_params(8)
    _, o, _ = sr.forward(p, x, y, vx, vy, 0.0, -9.8, 1.0, 1 / 60)
    _, o_s, _ = sr.forward(p, x + np.float32(64), y + np.float32(32), vx, vy, 0.0, -9.8, 1.0, 1 / 60)
    np.testing.assert_allclose(o, o_s, rtol=1e-12, atol=1e-12)
    perm = rng.permutation(n)
    _, o_p, _ = sr.forward(p, x[perm], y[perm], vx[perm], vy[perm], 0.0, -9.8, 1.0, 1 / 60)
    np.testing.assert_allclose(o_p, o[perm], rtol=1e-12, atol=1e-12)


def test_isolated_particle_sees_no_edges():
    x = np.array([1.0, 5.0], np.float32)
    y = np.array([1.0, 5.0], np.float32)
    v = np.zeros(2, np.float32)
    _, _, c = sr.forward(sr.random_params(1), x, y, v, v, 0.0, -9.8, 1.0, 1 / 60)
    assert np.all(c["S"] == 0) and np.all(c["deg"] == 0)


def test_gradient_matches_finite_differences():
    x, y, vx, vy = small_scene(5, 11)
    p = sr.random_params(12)
    rng = np.random.default_rng(13)
    tgt = rng.normal(0, 0.05, (len(x), 2))
    pairs = sr.neighbour_pairs(x, y, 1.0)
    args = (x, y, vx, vy, 0.4, -9.8, 1.0, 1 / 60, tgt)
    loss, g = sr.loss_and_grads(p, *args, pairs=pairs)
    blocks = [(sr.oW1, sr.oW2, sr.KE, 6), (sr.oW2, sr.ob2, None, None), (sr.ob2, sr.oW3, None, None),
              (sr.oW3, sr.oW4, sr.KN, 69), (sr.oW4, sr.ob4, None, None), (sr.ob4, sr.NPARAMS, None, None)]
    idx = []
    for a, b, width, live in blocks:
        cand = np.arange(a, b)
        if width:
            cand = cand[(cand - a) % width < live]
        idx += list(rng.choice(cand, size=min(12, len(cand)), replace=False))
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        lp, _ = sr.loss_and_grads(pp, *args, pairs=pairs)
        lm, _ = sr.loss_and_grads(pm, *args, pairs=pairs)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    # padding columns carry no gradient
    W1 = g[sr.oW1:sr.oW2].reshape(sr.H, sr.KE)
    W3 = g[sr.oW3:sr.oW4].reshape(sr.H, sr.KN)
    assert np.all(W1[:, 6:] == 0) and np.all(W3[:, 69:] == 0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p = sr.random_params(4)
    g = rng.normal(0, 1, sr.NPARAMS)
    z = np.zeros(sr.NPARAMS)
    p1, m1, v1 = sr.adam(p, g, z, z, 1, 1e-3)
    live = np.ones(sr.NPARAMS, bool)
    live[sr.oW1:sr.oW2] = (np.arange(sr.H * sr.KE) % sr.KE) < 6
    live[sr.oW3:sr.oW4] = (np.arange(sr.H * sr.KN) % sr.KN) < 69
    want = np.where(live, p - 1e-3 * g / (np.abs(g) + 1e-8), p)
    np.testing.assert_allclose(p1, want, rtol=1e-12, atol=1e-15)
