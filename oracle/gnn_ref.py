"""ORACLE -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the learned pressure-guess network of
include/cfd_gnn.h (no counterpart in the reference, whose pressure solve
starts from x0 = 0, src/navier_pressure.cpp:21): forward pass, loss and the
gradient of the loss with respect to every parameter.

bf16=True rounds to bfloat16 (nearest even) exactly where the device does:
the padded input features, the MLP weights, every hidden activation after its
ReLU, every message, and the masked gradients dZ that enter the MFMA
products; accumulation is FP32. bf16=False computes in float64 with no
rounding: a smooth (piecewise linear) function, which the tests check against
central finite differences. Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

IN, H, L = 8, 64, 3


def layout():
    """Offsets in the packed parameter vector:
    W0[8][64] c0[64] (W_l[128][64] c_l[64]) x3 w_out[64] c_out[1]."""
    off, o = {}, 0
    off["W0"] = o; o += IN * H
    off["c0"] = o; o += H
    for l in range(L):
        off[f"W{l + 1}"] = o; o += 2 * H * H
        off[f"c{l + 1}"] = o; o += H
    off["wout"] = o; o += H
    off["cout"] = o; o += 1
    off["N"] = o
    return off


OFF = layout()
NPARAM = OFF["N"]


def unpack(p):
    p = np.asarray(p)
    W0 = p[OFF["W0"]:OFF["W0"] + IN * H].reshape(IN, H)
    c0 = p[OFF["c0"]:OFF["c0"] + H]
    Ws = [p[OFF[f"W{l + 1}"]:OFF[f"W{l + 1}"] + 2 * H * H].reshape(2 * H, H) for l in range(L)]
    cs = [p[OFF[f"c{l + 1}"]:OFF[f"c{l + 1}"] + H] for l in range(L)]
    return W0, c0, Ws, cs, p[OFF["wout"]:OFF["wout"] + H], p[OFF["cout"]]


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[OFF["W0"]:OFF["W0"] + IN * H] = rng.normal(0, np.sqrt(2 / IN), IN * H)
    for l in range(L):
        o = OFF[f"W{l + 1}"]
        p[o:o + 2 * H * H] = rng.normal(0, np.sqrt(2 / (2 * H)), 2 * H * H)
    p[OFF["wout"]:OFF["wout"] + H] = rng.normal(0, np.sqrt(2 / H), H)
    return p


def message(col, w, h):
    """m_i = sum_k w[k][i] h[col[k][i]], skipping padding (col = self or w = 0)."""
    K, n = col.shape
    m = np.zeros_like(h)
    for k in range(K):
        c = col[k]
        wk = np.where((c == np.arange(n)) | (w[k] == 0), 0.0, w[k]).astype(h.dtype)
        m += wk[:, None] * h[c]
    return m


def message_T(col, w, g):
    """Transpose of message(): out_j = sum over (k, i) with col[k][i] = j of
    w[k][i] g_i (a scatter, independent of any symmetry of the graph)."""
    K, n = col.shape
    out = np.zeros_like(g)
    for k in range(K):
        c = col[k]
        wk = np.where((c == np.arange(n)) | (w[k] == 0), 0.0, w[k]).astype(g.dtype)
        np.add.at(out, c, wk[:, None] * g)
    return out


def forward(p, col, w, feat, bf16_mode=True):
    """Returns y and the cache of activations."""
    dt = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda x: x)
    W0, c0, Ws, cs, wout, cout = unpack(np.asarray(p, dtype=dt))
    col = np.asarray(col)
    w = np.asarray(w, dtype=dt)
    x0 = rnd(np.asarray(feat, dtype=dt))
    h = [rnd(np.maximum(x0 @ rnd(W0) + c0, 0))]
    m = []
    for l in range(L):
        ml = rnd(message(col, w, h[l]))
        m.append(ml)
        z = np.concatenate([h[l], ml], axis=1) @ rnd(Ws[l]) + cs[l]
        h.append(rnd(np.maximum(z, 0)))
    y = h[L] @ wout + cout
    return y, dict(x0=x0, h=h, m=m)


def loss_grad(p, col, w, feat, t, bf16_mode=True):
    """loss = mean(e^2), e = (y - mean y) - (t - mean t); returns (loss, grad)."""
    dt = np.float32 if bf16_mode else np.float64
    rnd = bf16 if bf16_mode else (lambda x: x)
    y, cache = forward(p, col, w, feat, bf16_mode)
    W0, c0, Ws, cs, wout, cout = unpack(np.asarray(p, dtype=dt))
    col = np.asarray(col)
    w = np.asarray(w, dtype=dt)
    t = np.asarray(t, dtype=dt)
    n = y.shape[0]
    e = (y - y.mean()) - (t - t.mean())
    loss = float(np.mean(e.astype(np.float64) ** 2))
    dy = (2.0 * e / n).astype(dt)
    g = np.zeros(NPARAM, dtype=np.float64)
    hL = cache["h"][L]
    g[OFF["wout"]:OFF["wout"] + H] = hL.T @ dy
    g[OFF["cout"]] = dy.sum()
    dh = np.outer(dy, wout).astype(dt)
    for l in reversed(range(L)):
        hl, ml, hn = cache["h"][l], cache["m"][l], cache["h"][l + 1]
        dz = rnd(np.where(hn > 0, dh, 0))
        X = np.concatenate([hl, ml], axis=1)
        o = OFF[f"W{l + 1}"]
        g[o:o + 2 * H * H] = (X.T @ dz).reshape(-1)
        o = OFF[f"c{l + 1}"]
        g[o:o + H] = dz.sum(axis=0)
        dX = dz @ rnd(Ws[l]).T
        dh = dX[:, :H] + message_T(col, w, dX[:, H:])
    dz = rnd(np.where(cache["h"][0] > 0, dh, 0))
    g[OFF["W0"]:OFF["W0"] + IN * H] = (cache["x0"].T @ dz).reshape(-1)
    g[OFF["c0"]:OFF["c0"] + H] = dz.sum(axis=0)
    return loss, g


def features(diag, val, b, pprev, vol, cx):
    """Cell features and message weights of one pressure system (float64):
    [vol/mean vol, centroid scaled to the bounding box of the centroids,
    b/rms b, p_prev/rms p_prev, a_ii/rms diag, 1]; w_ik = a_ik / |a_ii|."""
    n = diag.shape[0]
    f = np.zeros((n, IN))
    f[:, 0] = vol / vol.mean()
    lo, hi = cx.min(axis=0), cx.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    f[:, 1:4] = (cx - lo) / span
    rb = np.sqrt(np.mean(b * b))
    f[:, 4] = b / rb if rb > 0 else 0
    if pprev is not None:
        rp = np.sqrt(np.mean(pprev * pprev))
        f[:, 5] = pprev / rp if rp > 0 else 0
    rd = np.sqrt(np.mean(diag * diag))
    f[:, 6] = diag / rd if rd > 0 else 0
    f[:, 7] = 1
    ad = np.abs(diag)
    w = np.where(ad > 0, val / np.where(ad > 0, ad, 1), 0)
    return f, w


def galerkin(A, b, y, p):
    """A-norm optimal x0 in span{y, p} (either may be None): solves the 2x2
    (or 1x1) projected system V^T A V c = V^T b."""
    V = [v for v in (y, p) if v is not None]
    if not V:
        return np.zeros_like(b)
    V = np.stack(V, axis=1)
    c = np.linalg.lstsq(V.T @ (A @ V), V.T @ b, rcond=None)[0]
    return V @ c
