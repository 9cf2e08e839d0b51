"""CPU statement of the message-passing warm-start network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
forward and backward kernels.  The network has no counterpart in the
reference; this file fixes its definition:

Per sample (one p' system A x = b on a mesh of N cells, ELL width W):
  s_b = sqrt(mean(b^2)) + 1e-300, s_a = mean(ac), v = mean(vol)
  features  X0[i] = [b_i/s_b, ac_i/s_a, nnb_i/W, vol_i/v, xp_i s_a/s_b, 0, 0, 0]
            (xp = the previous p' solution, mean-free, or 0: the optional
            fifth input of the network built with use_prev)
  edges     e[j,i] = -anb[j,i]/ac_i on slots with a neighbour, else 0
  target    t = (x - mean(x)) * s_a/s_b
Network (hidden H = 64, L layers, every weight used in BF16, sums in FP32):
  H0 = bf16(relu(X0 We + be))
  layer l: M = bf16(sum_j e[j,i] H[nbr_j(i)]);  z = [H, M] W_l + b_l
           H = bf16(H + relu(z))
  y = H wd + bd;  yc = y - mean(y)
  loss = sum over samples and cells of (yc - t)^2 / (total cells)
  prediction x_hat = yc * s_b/s_a
Backward: straight-through BF16 (gradients of the master FP32 weights), with
dz rounded to BF16 where the device feeds it to the matrix cores.

``bf16=False`` runs the same definition without any rounding (for the
finite-difference checks that pin the backward formulas).
"""
from __future__ import annotations

import numpy as np

HIDDEN = 64
FEATURES = 8


def bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).reshape(a.shape)


def param_layout(layers=3, hidden=HIDDEN):
    """(name, shape) of the flat parameter vector, in device order."""
    lay = [("We", (FEATURES, hidden)), ("be", (hidden,))]
    for l in range(layers):
        lay += [(f"W{l}", (2 * hidden, hidden)), (f"b{l}", (hidden,))]
    lay += [("wd", (hidden,)), ("bd", (1,))]
    return lay


def unflatten(theta, layers=3):
    out, o = {}, 0
    for name, shape in param_layout(layers):
        k = int(np.prod(shape))
        out[name] = theta[o:o + k].reshape(shape)
        o += k
    return out


def flatten(d, layers=3):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in param_layout(layers)])


def n_params(layers=3):
    return sum(int(np.prod(s)) for _, s in param_layout(layers))


def features(nbr0, b, ac, anb, vol, x=None, xprev=None):
    """Per-sample features, edge weights and target (float64).
    nbr0: (W, N) 0-based neighbours, -1 = none."""
    W, N = nbr0.shape
    sb = np.sqrt(np.mean(b * b)) + 1e-300
    sa = np.mean(ac)
    X0 = np.zeros((N, FEATURES))
    X0[:, 0] = b / sb
    X0[:, 1] = ac / sa
    X0[:, 2] = (nbr0 >= 0).sum(0) / W
    X0[:, 3] = vol / np.mean(vol)
    if xprev is not None:
        X0[:, 4] = (xprev - np.mean(xprev)) * sa / sb
    e = np.where(nbr0 >= 0, -anb / ac[None, :], 0.0)
    t = None if x is None else (x - np.mean(x)) * sa / sb
    return X0, e, t, sb, sa


def _agg(nbr0, e, H):
    W, N = nbr0.shape
    M = np.zeros_like(H)
    for j in range(W):
        ok = nbr0[j] >= 0
        M[ok] += e[j, ok, None] * H[nbr0[j, ok]]
    return M


def _agg_t(nbr0, e, D):
    """Transpose of _agg: out[k] = sum over (j, i) with nbr0[j, i] = k of e[j, i] D[i]."""
    W, N = nbr0.shape
    out = np.zeros_like(D)
    for j in range(W):
        ok = np.nonzero(nbr0[j] >= 0)[0]
        np.add.at(out, nbr0[j, ok], e[j, ok, None] * D[ok])
    return out


def forward(theta, samples, layers=3, bf16_round=True):
    """samples: list of (nbr0, X0, e, t).  Returns (loss, caches, ys)."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    P = unflatten(np.asarray(theta, np.float64), layers)
    Wq = {k: (bf16(v) if bf16_round else v) for k, v in P.items() if k.startswith(("W", "wd"))}
    total = sum(s[1].shape[0] for s in samples)
    loss = 0.0
    caches, ys = [], []
    dt = np.float32 if bf16_round else np.float64
    for nbr0, X0, e, t in samples:
        X0 = X0.astype(dt)
        e = e.astype(dt)
        z0 = X0 @ Wq["We"].astype(dt) + P["be"].astype(dt)
        H = rnd(np.maximum(z0, 0))
        c = {"X0": X0, "H0": H, "layers": []}
        for l in range(layers):
            M = rnd(_agg(nbr0, e, H))
            z = np.concatenate([H, M], 1) @ Wq[f"W{l}"].astype(dt) + P[f"b{l}"].astype(dt)
            Hn = rnd(H + np.maximum(z, 0))
            c["layers"].append({"H": H, "M": M, "mask": z > 0})
            H = Hn
        y = H @ Wq["wd"].astype(dt) + dt(P["bd"][0])
        yc = y - y.mean()
        c["H3"] = H
        c["yc"] = yc
        ys.append(yc)
        if t is not None:
            loss += float(np.sum((yc.astype(np.float64) - t) ** 2))
        caches.append(c)
    return loss / total, caches, ys


def backward(theta, samples, caches, layers=3, bf16_round=True):
    """Gradient of the loss of forward() w.r.t. the flat master parameters."""
    rnd = bf16 if bf16_round else (lambda a: a)
    P = unflatten(np.asarray(theta, np.float64), layers)
    Wq = {k: (bf16(v) if bf16_round else v) for k, v in P.items() if k.startswith(("W", "wd"))}
    dt = np.float32 if bf16_round else np.float64
    total = sum(s[1].shape[0] for s in samples)
    g = {name: np.zeros(shape) for name, shape in param_layout(layers)}
    for (nbr0, X0, e, t), c in zip(samples, caches):
        e = e.astype(dt)
        dy = (2.0 * (c["yc"].astype(np.float64) - t) / total).astype(dt)
        g["wd"] += c["H3"].T.astype(np.float64) @ dy
        g["bd"][0] += float(np.sum(dy))
        dH = np.outer(dy, Wq["wd"]).astype(dt)
        for l in reversed(range(layers)):
            L = c["layers"][l]
            dz = rnd(np.where(L["mask"], dH, 0).astype(dt))
            HM = np.concatenate([L["H"], L["M"]], 1)
            g[f"W{l}"] += HM.T.astype(np.float64) @ dz.astype(np.float64)
            g[f"b{l}"] += dz.astype(np.float64).sum(0)
            D = dz @ Wq[f"W{l}"].astype(dt).T
            dH = dH + D[:, :HIDDEN] + _agg_t(nbr0, e, D[:, HIDDEN:])
        dz0 = np.where(c["H0"] > 0, dH, 0)
        g["We"] += c["X0"].T.astype(np.float64) @ dz0.astype(np.float64)
        g["be"] += dz0.astype(np.float64).sum(0)
    return flatten(g, layers)


def init_params(seed, layers=3):
    """He-style initialisation (the device draws its own; tests pass parameters explicitly)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in param_layout(layers):
        if name.startswith("W"):
            d[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])
        elif name == "wd":
            d[name] = rng.standard_normal(shape) * np.sqrt(1.0 / shape[0])
        else:
            d[name] = np.zeros(shape)
    return flatten(d, layers)
