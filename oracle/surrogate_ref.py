"""FP64 CPU restatement of the learned macro-step network - test infrastructure.

Only tests/, the smoke check and the bench's CPU leg use this module; the
product path (libsph_hip.so) never does.  It has no counterpart in the
reference (AdamSimpson/SPH has no learned component): it restates the network
as libsph_hip.so's sph_mlp.hip defines it, in float64 and without the BF16
rounding, so the device forward, loss, gradients and Adam step can be checked
against it to BF16 tolerances.

Network (per particle i, neighbours j with |x_j - x_i| <= h, j != i):
    f_ij = (dx/h, dy/h, 0.2 dvx, 0.2 dvy, r/h, 1)
    S_i  = sum_j SiLU(W1 f_ij)                     (64)
    m_i  = W2 S_i + deg_i b2                       (= sum_j (W2 SiLU(W1 f_ij) + b2))
    u_i  = [m_i, 0.2 v_i, 0.1 a_i, 1]              (69 of 80 columns)
    o_i  = W4 SiLU(W3 u_i) + b4                    (2)
    d_i  = 4 dt v_i + 10 dt^2 a_i + 0.05 o_i
    loss = mean_i |d_i - y_i|^2 / 0.05^2
a_i = (sloshing acceleration, -g) of the particle's scene.

Pins (tests/test_surrogate_oracle.py): the backward against central finite
differences of this forward; the forward against the per-edge textbook form
(layer 2 applied on every edge, then summed), the ballistic closed form with
a zero output layer, and translation / relabelling invariance.
"""
from __future__ import annotations

import numpy as np

H = 64
KE = 16
KN = 80
VS, AS, DS = 0.2, 0.1, 0.05

oW1 = 0
oW2 = oW1 + H * KE
ob2 = oW2 + H * H
oW3 = ob2 + H
oW4 = oW3 + H * KN
ob4 = oW4 + 2 * H
NPARAMS = ob4 + 2


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[oW1:oW2].reshape(H, KE), W2=p[oW2:ob2].reshape(H, H), b2=p[ob2:oW3],
                W3=p[oW3:oW4].reshape(H, KN), W4=p[oW4:ob4].reshape(2, H), b4=p[ob4:ob4 + 2])


def pack(d):
    p = np.zeros(NPARAMS)
    p[oW1:oW2] = d["W1"].ravel()
    p[oW2:ob2] = d["W2"].ravel()
    p[ob2:oW3] = d["b2"]
    p[oW3:oW4] = d["W3"].ravel()
    p[oW4:ob4] = d["W4"].ravel()
    p[ob4:ob4 + 2] = d["b4"]
    return p


def random_params(seed=0, out_scale=1.0):
    """Random parameters with the library's zero padding columns."""
    rng = np.random.default_rng(seed)
    d = dict(W1=np.zeros((H, KE)), W2=rng.uniform(-1, 1, (H, H)) / 8.0, b2=rng.uniform(-0.2, 0.2, H),
             W3=np.zeros((H, KN)), W4=rng.uniform(-1, 1, (2, H)) / 8.0 * out_scale,
             b4=rng.uniform(-0.1, 0.1, 2))
    d["W1"][:, :6] = rng.uniform(-1, 1, (H, 6)) / np.sqrt(6.0)
    d["W3"][:, :69] = rng.uniform(-1, 1, (H, 69)) / np.sqrt(69.0)
    return pack(d)


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def neighbour_pairs(x, y, h):
    """All ordered pairs (i, j), i != j, with r^2 <= h^2 evaluated in float32
    as the device evaluates it (dx = x_j - x_i, r2 = dx*dx + dy*dy)."""
    from scipy.spatial import cKDTree
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    pts = np.stack([x, y], 1).astype(np.float64)
    if len(pts) < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    pr = cKDTree(pts).query_pairs(float(h) * (1 + 1e-4), output_type="ndarray")
    i = np.concatenate([pr[:, 0], pr[:, 1]]).astype(np.int64)
    j = np.concatenate([pr[:, 1], pr[:, 0]]).astype(np.int64)
    dx = (x[j] - x[i]).astype(np.float32)
    dy = (y[j] - y[i]).astype(np.float32)
    r2 = (dx * dx + dy * dy).astype(np.float32)
    keep = ~(r2 > np.float32(h) * np.float32(h))
    return i[keep], j[keep]


def edge_features(x, y, vx, vy, i, j, h):
    x, y, vx, vy = (np.asarray(a, np.float32) for a in (x, y, vx, vy))
    dx = (x[j] - x[i]).astype(np.float32)
    dy = (y[j] - y[i]).astype(np.float32)
    r2 = (dx * dx + dy * dy).astype(np.float32)
    f = np.zeros((len(i), KE))
    hr = 1.0 / float(h)
    f[:, 0] = dx * hr
    f[:, 1] = dy * hr
    f[:, 2] = (vx[j] - vx[i]).astype(np.float64) * VS
    f[:, 3] = (vy[j] - vy[i]).astype(np.float64) * VS
    f[:, 4] = np.sqrt(r2.astype(np.float64)) * hr
    f[:, 5] = 1.0
    return f


def forward(p, x, y, vx, vy, ax, ay, h, dt, pairs=None):
    """Returns (d, o, cache).  ax/ay: per-particle acceleration (scalar or array)."""
    P = unpack(p)
    n = len(x)
    i, j = neighbour_pairs(x, y, h) if pairs is None else pairs
    f = edge_features(x, y, vx, vy, i, j, h)
    pre1 = f @ P["W1"].T
    S = np.zeros((n, H))
    np.add.at(S, i, silu(pre1))
    deg = np.bincount(i, minlength=n).astype(np.float64)
    m = S @ P["W2"].T + deg[:, None] * P["b2"][None, :]
    ax = np.broadcast_to(np.asarray(ax, np.float64), (n,))
    ay = np.broadcast_to(np.asarray(ay, np.float64), (n,))
    vx64 = np.asarray(vx, np.float64)
    vy64 = np.asarray(vy, np.float64)
    u = np.zeros((n, KN))
    u[:, :H] = m
    u[:, 64] = vx64 * VS
    u[:, 65] = vy64 * VS
    u[:, 66] = ax * AS
    u[:, 67] = ay * AS
    u[:, 68] = 1.0
    pre3 = u @ P["W3"].T
    z = silu(pre3)
    o = z @ P["W4"].T + P["b4"][None, :]
    ball = np.stack([4 * dt * vx64 + 10 * dt * dt * ax, 4 * dt * vy64 + 10 * dt * dt * ay], 1)
    d = ball + DS * o
    cache = dict(P=P, i=i, j=j, f=f, pre1=pre1, S=S, deg=deg, u=u, pre3=pre3, z=z, n=n)
    return d, o, cache


def forward_per_edge(p, x, y, vx, vy, ax, ay, h, dt):
    """The textbook form: the whole edge MLP on every edge, then the sum."""
    P = unpack(p)
    n = len(x)
    i, j = neighbour_pairs(x, y, h)
    f = edge_features(x, y, vx, vy, i, j, h)
    e = silu(f @ P["W1"].T) @ P["W2"].T + P["b2"][None, :]
    m = np.zeros((n, H))
    for k in range(len(i)):  # plain loop: no vectorised scatter
        m[i[k]] += e[k]
    out = np.zeros((n, 2))
    for q in range(n):
        u = np.concatenate([m[q], [float(vx[q]) * VS, float(vy[q]) * VS,
                                   float(np.broadcast_to(ax, (n,))[q]) * AS,
                                   float(np.broadcast_to(ay, (n,))[q]) * AS, 1.0],
                            np.zeros(KN - 69)])
        out[q] = P["W4"] @ silu(P["W3"] @ u) + P["b4"]
    return out


def loss_and_grads(p, x, y, vx, vy, ax, ay, h, dt, target, n_total=None, pairs=None):
    """Loss (mean over n_total particles) and its gradient w.r.t. p (NPARAMS)."""
    d, o, c = forward(p, x, y, vx, vy, ax, ay, h, dt, pairs)
    P = c["P"]
    n = c["n"]
    nt = n if n_total is None else n_total
    e = (d - np.asarray(target, np.float64)) / DS
    loss = float((e * e).sum() / nt)
    dout = 2.0 * e / nt                       # d loss / d o   (d = ball + DS o)
    g = {}
    g["W4"] = dout.T @ c["z"]
    g["b4"] = dout.sum(0)
    dz = dout @ P["W4"]
    dpre3 = dz * silu_grad(c["pre3"])
    g["W3"] = dpre3.T @ c["u"]
    g["W3"][:, 69:] = 0.0
    du = dpre3 @ P["W3"]
    dm = du[:, :H]
    g["W2"] = dm.T @ c["S"]
    g["b2"] = (c["deg"][:, None] * dm).sum(0)
    dS = dm @ P["W2"]
    dpre1 = dS[c["i"]] * silu_grad(c["pre1"])
    g["W1"] = dpre1.T @ c["f"]
    g["W1"][:, 6:] = 0.0
    return loss, pack(g)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t = 1-based step count) on the live parameters; the
    padding columns of W1 and W3 are never updated."""
    live = np.ones(NPARAMS, bool)
    live[oW1:oW2] = (np.arange(H * KE) % KE) < 6
    live[oW3:oW4] = (np.arange(H * KN) % KN) < 69
    m = np.where(live, b1 * m + (1 - b1) * g, m)
    v = np.where(live, b2 * v + (1 - b2) * g * g, v)
    c1 = 1 - b1 ** t
    c2 = 1 - b2 ** t
    p = np.where(live, p - lr * (m / c1) / (np.sqrt(v / c2) + eps), p)
    return p, m, v
