"""Training and inference drivers of the learned macro-step (sph_surrogate_*).

Everything runs on the GPU through libsph_hip.so: the trajectories stay in
device memory in the library's own trajectory store (sph_trajectory_*).

* ``generate``  - run the plain GPU solver over a batch of scenes and record
  the state (x, y, v_x, v_y) every MACRO steps: the training trajectories.
* ``train``     - Adam on batches of (scene, frame) samples drawn from the
  trajectories: load the sampled states into a batch context, hash, one
  forward/backward on the device, one Adam step.
* ``evaluate``  - 4-step displacement error of the network against the
  method without it (ballistic extrapolation: the network's output zeroed).
* ``run_judged`` - judged macro-steps (sph_surrogate_macro_step): every
  scene moves by the network, scenes whose mean density moves by more than
  the tolerance are re-taken by the plain solver.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import Solver, Surrogate, Trajectory
from . import scenes as sc

MACRO = 4
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG_DIR, "configs", "surrogate_params.npy")
WEIGHTS_META = os.path.join(PKG_DIR, "configs", "surrogate_params.json")


def make_batch(scenes, upload=True):
    """Batch context of sloshing dam breaks [(tunables, box, parts, amp, period)]."""
    s = Solver.batch([(t, b, len(p)) for t, b, p, _, _ in scenes])
    if upload:
        parts = np.concatenate([p for _, _, p, _, _ in scenes])
        s.upload_batch(parts, [len(p) for _, _, p, _, _ in scenes])
    for k, (_, _, _, a, T) in enumerate(scenes):
        s.set_scene_forcing(k, a, T)
    return s


def generate(scenes, steps=200, every=MACRO):
    """Plain-solver trajectories of equal-size scenes: a Trajectory with
    steps / every + 1 frames, and the solver's health counters."""
    n = len(scenes[0][2])
    assert all(len(p) == n for _, _, p, _, _ in scenes)
    s = make_batch(scenes)
    T = steps // every
    tr = Trajectory(len(scenes), n, T + 1)
    tr.record(s, 0)
    for t in range(1, T + 1):
        s.step(every)
        tr.record(s, t)
    c = s.counters()
    s.close()
    return tr, c


def _load_samples(ctx, tr, scenes, s_sel, f_sel, every=MACRO):
    """Load frames f_sel of scenes s_sel into the batch context (one sample
    per context scene), set each sample's sloshing and clock, and hash."""
    tr.load(ctx, s_sel, f_sel)
    for k, (s, f) in enumerate(zip(s_sel, f_sel)):
        _, _, _, a, T = scenes[int(s)]
        ctx.set_scene_forcing(k, a, T)
        ctx.set_scene_time_offset(k, int(f) * every)
    ctx.phase("hash")


def _batch_ctx(scenes, batch):
    ctx = Solver.batch([(t, b, len(p)) for t, b, p, _, _ in scenes[:batch]])
    ctx.upload_batch(np.concatenate([p for _, _, p, _, _ in scenes[:batch]]),
                     [len(p) for _, _, p, _, _ in scenes[:batch]])
    return ctx


def train(model, tr, scenes, iters, batch=64, lr=1e-3, seed=0, log=None, log_every=100):
    """Adam on random (scene, frame) samples; returns the loss history."""
    rng = np.random.default_rng(seed)
    ctx = _batch_ctx(scenes, batch)
    hist = []
    t0 = time.time()
    for it in range(iters):
        s_sel = rng.choice(tr.nscenes, size=batch, replace=False)
        f_sel = rng.integers(0, tr.frames - 1, size=batch)
        _load_samples(ctx, tr, scenes, s_sel, f_sel)
        model.zero_grads()
        hist.append(model.backward_trajectory(ctx, tr))
        model.adam(lr)
        if log and (it % log_every == 0 or it == iters - 1):
            log(f"iter {it} loss {np.mean(hist[-log_every:]):.4f} ({time.time() - t0:.1f} s)")
    ctx.close()
    return hist


def ballistic_model(model):
    """The method without the network: the same model with its output layer
    (W4, b4: the last 2 x 64 + 2 parameters) zeroed."""
    z = Surrogate(seed=0)
    p = model.params.copy()
    p[-(2 * 64 + 2):] = 0.0
    z.params = p
    return z


def evaluate(model, tr, scenes, frames, batch=None):
    """RMS 4-step displacement error (units of h) of the network and of the
    ballistic extrapolation over all scenes of the trajectory at the given
    frames."""
    S = tr.nscenes
    batch = min(batch or S, S)
    ctx = _batch_ctx(scenes, batch)
    zero = ballistic_model(model)
    se_net = se_bal = 0.0
    cnt = 0
    for f in frames:
        a, b = tr.read(f), tr.read(f + 1)
        for b0 in range(0, S - batch + 1, batch):
            s_sel = np.arange(b0, b0 + batch)
            _load_samples(ctx, tr, scenes, s_sel, np.full(batch, f))
            y = (b[s_sel, :, :2] - a[s_sel, :, :2]).reshape(-1, 2).astype(np.float64)
            d_net = model.predict(ctx).astype(np.float64)
            d_bal = zero.predict(ctx).astype(np.float64)
            se_net += ((d_net - y) ** 2).sum()
            se_bal += ((d_bal - y) ** 2).sum()
            cnt += len(y)
    ctx.close()
    zero.close()
    return float(np.sqrt(se_net / cnt)), float(np.sqrt(se_bal / cnt))


def run_judged(model, solver, n_macro, tol=0.01):
    """n_macro judged macro-steps; returns (seconds, fallback flags per macro-step)."""
    flags = []
    solver.sync()
    t0 = time.perf_counter()
    for _ in range(n_macro):
        flags.append(model.macro_step(solver, tol))
    solver.sync()
    return time.perf_counter() - t0, np.array(flags)


def save_weights(model, meta, path=WEIGHTS, meta_path=WEIGHTS_META):
    np.save(path, model.params.astype(np.float32))
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def load_weights(model, path=WEIGHTS):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path}: train the network first (python -m sph_amd.surrogate)")
    model.params = np.load(path, allow_pickle=False)


def main(argv=None):
    """Train on the sloshing dam breaks (512 scenes x 16384 particles, 200
    plain steps each, batches of 64 scenes, Adam lr 1e-3) and save the
    weights next to this module."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--holdout", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    t0 = time.time()
    scn = sc.sloshing_dam_breaks(args.scenes + args.holdout, n=args.n, seed=args.seed)
    train_s, hold_s = scn[:args.scenes], scn[args.scenes:]
    tr, c = generate(train_s, steps=args.steps)
    log(f"generated {tr.nscenes} x {tr.n} x {tr.frames} frames in {time.time() - t0:.1f} s; "
        f"counters {c}")
    hold, _ = generate(hold_s, steps=args.steps)
    model = Surrogate(seed=args.seed)
    hist = train(model, tr, train_s, args.iters, batch=args.batch, lr=args.lr, seed=args.seed,
                 log=log)
    frames = list(range(0, hold.frames - 1, 5))
    net, bal = evaluate(model, hold, hold_s, frames, batch=min(args.batch, args.holdout))
    log(f"held-out 4-step displacement rms: network {net:.5f} h, ballistic {bal:.5f} h")
    meta = dict(scenes=args.scenes, n=args.n, steps=args.steps, iters=args.iters, batch=args.batch,
                lr=args.lr, seed=args.seed, final_loss=float(np.mean(hist[-100:])),
                holdout_rms_network=net, holdout_rms_ballistic=bal,
                seconds=time.time() - t0)
    save_weights(model, meta, args.out, os.path.splitext(args.out)[0] + ".json")
    log(json.dumps(meta))


if __name__ == "__main__":
    main()
