"""Training and inference drivers of the learned macro-step (sph_surrogate_*).

Everything here runs on the GPU through libsph_hip.so; torch only holds the
trajectory data in HBM and does index arithmetic on it.

* ``generate``  - run the plain GPU solver over a batch of scenes and keep
  the state (x, y, v_x, v_y) every MACRO steps: the training trajectories.
* ``train``     - Adam on batches of (scene, time) samples drawn from the
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

from . import Solver, Surrogate
from . import scenes as sc

MACRO = 4
PKG_DIR = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG_DIR, "configs", "surrogate_params.npy")
WEIGHTS_META = os.path.join(PKG_DIR, "configs", "surrogate_params.json")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned macro-step needs a GPU")
    return torch


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
    """Plain-solver trajectories: a float32 CUDA tensor [steps/every + 1,
    nscenes, n, 4] of (x, y, v_x, v_y) by particle id (all scenes of equal n)."""
    torch = _torch()
    n = len(scenes[0][2])
    assert all(len(p) == n for _, _, p, _, _ in scenes)
    s = make_batch(scenes)
    T = steps // every
    data = torch.empty((T + 1, len(scenes), n, 4), dtype=torch.float32, device="cuda")
    s.gather_state(data[0].view(-1, 4))
    for t in range(1, T + 1):
        s.step(every)
        s.gather_state(data[t].view(-1, 4))
    c = s.counters()
    s.close()
    return data, c


def _load_samples(ctx, data, scenes, s_sel, t_sel, every=MACRO):
    """Load states data[t, s] of the sampled (scene, time) pairs into the batch
    context (one sample per context scene) and hash them; returns the target
    4-step displacements as a CUDA tensor [B * n, 2]."""
    torch = _torch()
    ts = torch.as_tensor(t_sel, device="cuda")
    ss = torch.as_tensor(s_sel, device="cuda")
    xs = data[ts, ss].reshape(-1, 4).contiguous()
    tgt = (data[ts + 1, ss, :, :2].reshape(-1, 2) - xs[:, :2]).contiguous()
    ctx.load_state(xs)
    for k, (s, t) in enumerate(zip(s_sel, t_sel)):
        _, _, _, a, T = scenes[int(s)]
        ctx.set_scene_forcing(k, a, T)
        ctx.set_scene_time_offset(k, int(t) * every)
    ctx.phase("hash")
    return tgt


def train(model, data, scenes, iters, batch=64, lr=1e-3, seed=0, log=None, log_every=100):
    """Adam on random (scene, time) samples; returns the loss history."""
    rng = np.random.default_rng(seed)
    T = data.shape[0] - 1
    S = data.shape[1]
    ctx = make_batch(scenes[:batch], upload=False)
    # one upload to give every scene its particle count
    ctx.upload_batch(np.concatenate([p for _, _, p, _, _ in scenes[:batch]]),
                     [len(p) for _, _, p, _, _ in scenes[:batch]])
    hist = []
    t0 = time.time()
    for it in range(iters):
        s_sel = rng.choice(S, size=batch, replace=False)
        t_sel = rng.integers(0, T, size=batch)
        tgt = _load_samples(ctx, data, scenes, s_sel, t_sel)
        model.zero_grads()
        loss = model.backward(ctx, tgt)
        model.adam(lr)
        hist.append(loss)
        if log and (it % log_every == 0 or it == iters - 1):
            log(f"iter {it} loss {np.mean(hist[-log_every:]):.4f} ({time.time() - t0:.1f} s)")
    ctx.close()
    return hist


def evaluate(model, data, scenes, times, batch=None):
    """RMS 4-step displacement error (units of h) of the network and of the
    ballistic extrapolation (4 dt v + 10 dt^2 a: the network's output
    zeroed) over the given scenes and times."""
    torch = _torch()
    S = data.shape[1]
    batch = batch or S
    ctx = make_batch(scenes[:batch], upload=False)
    ctx.upload_batch(np.concatenate([p for _, _, p, _, _ in scenes[:batch]]),
                     [len(p) for _, _, p, _, _ in scenes[:batch]])
    zero = Surrogate(seed=0)
    p0 = model.params
    pz = p0.copy()
    n_out = 2 * 64 + 2  # W4 and b4 are the last 130 parameters
    pz[-n_out:] = 0.0
    zero.params = pz
    se_net = se_bal = 0.0
    cnt = 0
    for t in times:
        for b0 in range(0, S, batch):
            s_sel = np.arange(b0, min(S, b0 + batch))
            if len(s_sel) < batch:
                break
            tgt = _load_samples(ctx, data, scenes, s_sel, np.full(len(s_sel), t))
            y = tgt.cpu().numpy().astype(np.float64)
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
    torch = _torch()
    flags = []
    solver.sync()
    torch.cuda.synchronize()
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
    data, c = generate(train_s, steps=args.steps)
    log(f"generated {tuple(data.shape)} in {time.time() - t0:.1f} s; counters {c}")
    hold, _ = generate(hold_s, steps=args.steps)
    model = Surrogate(seed=args.seed)
    hist = train(model, data, train_s, args.iters, batch=args.batch, lr=args.lr, seed=args.seed,
                 log=log)
    times = list(range(0, hold.shape[0] - 1, 5))
    net, bal = evaluate(model, hold, hold_s, times, batch=min(args.batch, args.holdout))
    log(f"held-out 4-step displacement rms: network {net:.5f} h, ballistic {bal:.5f} h")
    meta = dict(scenes=args.scenes, n=args.n, steps=args.steps, iters=args.iters, batch=args.batch,
                lr=args.lr, seed=args.seed, final_loss=float(np.mean(hist[-100:])),
                holdout_rms_network=net, holdout_rms_ballistic=bal,
                seconds=time.time() - t0, dropped_edges=model.dropped_edges())
    save_weights(model, meta, args.out, os.path.splitext(args.out)[0] + ".json")
    log(json.dumps(meta))


if __name__ == "__main__":
    main()
