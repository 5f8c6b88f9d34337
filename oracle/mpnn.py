"""CPU reference of the learned-bound network -- TEST INFRASTRUCTURE ONLY.

Same network as gpu4gst_code_amd/csrc/gst_mpnn.hip (no counterpart in the reference):

    x_v  = [log1p(deg)/8, log1p(min incident weight)/8, in group 0..7, 0 x 6]
    h1_v = relu(W1 [x_v ; mean_{u in N(v)} x_u] + b1)        W1: 64 x 32
    h2_v = relu(W2 [h1_v ; mean_{u in N(v)} h1_u] + b2)      W2: 64 x 128
    s_v  = w3 . h2_v + b3
    L    = mean_v wt_v (softplus(s_v) - y_v s_v),  wt_v = pos_weight if y_v else 1

`grads(..., bf16=True)` rounds to bfloat16 exactly where the GPU stores bfloat16 (inputs,
aggregates, activations, weights fed to the matrix cores, pre-activation gradients), so the GPU
gradient can be compared tightly; `bf16=False` is the plain float math, pinned against finite
differences of `loss` in float64 (tests/test_mpnn_cpu.py).  The analytical backward pass written
here is the thing under test there, not a restatement of the GPU code.
"""
from __future__ import annotations

import numpy as np

FIN, HID = 16, 64
K1, K2 = 2 * FIN, 2 * HID
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * K1
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * K2
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAM = OFF_B3 + 1


def bf16(x, enabled=True):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if not enabled:
        return x
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, K1), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, K2),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def features(row_ptr, w, gm, rnd=True):
    deg = np.diff(row_ptr).astype(np.float32)
    n = deg.shape[0]
    minw = np.zeros(n, np.float32)
    for v in range(n):
        if deg[v] > 0:
            minw[v] = np.min(w[row_ptr[v]:row_ptr[v + 1]].astype(np.float32))
    x = np.zeros((n, FIN), np.float32)
    x[:, 0] = np.log1p(deg) * np.float32(0.125)
    x[:, 1] = np.log1p(minw) * np.float32(0.125)
    for i in range(8):
        x[:, 2 + i] = ((gm >> i) & 1).astype(np.float32)
    return bf16(x, rnd)


def _mean_agg(row_ptr, col, h, rnd):
    n = row_ptr.shape[0] - 1
    deg = np.diff(row_ptr)
    src = np.repeat(np.arange(n), deg)
    s = np.zeros_like(h, dtype=np.float32)
    np.add.at(s, src, h[col].astype(np.float32))
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0).astype(np.float32)
    return bf16(s * inv[:, None], rnd)


def _mean_agg_t(row_ptr, col, d):
    """transpose of the mean aggregation on a symmetric graph: out[u] = sum_{v in N(u)} d[v]/deg(v)"""
    n = row_ptr.shape[0] - 1
    deg = np.diff(row_ptr)
    src = np.repeat(np.arange(n), deg)
    scaled = d / np.maximum(deg, 1)[:, None]
    out = np.zeros_like(d)
    np.add.at(out, src, scaled[col])
    return out


def forward(p, row_ptr, col, w, gm, rnd=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    f = np.float32
    x = features(row_ptr, w, gm, rnd)
    ax = _mean_agg(row_ptr, col, x, rnd)
    X1 = np.concatenate([x, ax], axis=1)
    h1 = bf16(np.maximum(X1 @ bf16(W1.astype(f), rnd).T + b1.astype(f), 0), rnd)
    ah1 = _mean_agg(row_ptr, col, h1, rnd)
    X2 = np.concatenate([h1, ah1], axis=1)
    h2 = bf16(np.maximum(X2 @ bf16(W2.astype(f), rnd).T + b2.astype(f), 0), rnd)
    s = h2 @ w3.astype(f) + f(b3)
    return dict(x=x, ax=ax, X1=X1, h1=h1, ah1=ah1, X2=X2, h2=h2, s=s.astype(np.float32))


def loss(p, row_ptr, col, w, gm, y, pos_weight):
    """Loss in float64 without any rounding (for finite differences)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = features(row_ptr, w, gm, False).astype(np.float64)
    n = x.shape[0]
    ax = _mean_agg64(row_ptr, col, x)
    h1 = np.maximum(np.concatenate([x, ax], 1) @ W1.T + b1, 0)
    ah1 = _mean_agg64(row_ptr, col, h1)
    h2 = np.maximum(np.concatenate([h1, ah1], 1) @ W2.T + b2, 0)
    s = h2 @ w3 + b3
    wt = np.where(y > 0, pos_weight, 1.0) / n
    sp = np.where(s > 0, s + np.log1p(np.exp(-np.abs(s))), np.log1p(np.exp(s)))
    return float(np.sum(wt * (sp - y * s)))


def _mean_agg64(row_ptr, col, h):
    n = row_ptr.shape[0] - 1
    deg = np.diff(row_ptr)
    src = np.repeat(np.arange(n), deg)
    s = np.zeros_like(h, dtype=np.float64)
    np.add.at(s, src, h[col])
    return s / np.maximum(deg, 1)[:, None]


def grads(p, row_ptr, col, w, gm, y, pos_weight, rnd=True):
    """Analytical gradient of the loss (float32 math; bf16 rounding at the GPU's points if rnd)."""
    f = np.float32
    W1, b1, W2, b2, w3, b3 = unpack(p)
    A = forward(p, row_ptr, col, w, gm, rnd)
    n = A["s"].shape[0]
    y = y.astype(f)
    wt = (np.where(y > 0, f(pos_weight), f(1.0)) / f(n)).astype(f)
    sig = (1 / (1 + np.exp(-A["s"].astype(np.float64)))).astype(f)
    dS = wt * (sig - y)
    h2 = A["h2"]
    g_w3 = (dS[:, None] * h2).sum(0)
    g_b3 = dS.sum()
    dZ2 = bf16(np.where(h2 > 0, dS[:, None] * w3.astype(f)[None, :], 0), rnd)
    g_b2 = dZ2.sum(0)
    g_W2 = dZ2.T @ A["X2"]
    dX2 = dZ2 @ bf16(W2.astype(f), rnd)
    dH1 = dX2[:, :HID] + _mean_agg_t(row_ptr, col, dX2[:, HID:])
    dZ1 = bf16(np.where(A["h1"] > 0, dH1, 0), rnd)
    g_b1 = dZ1.sum(0)
    g_W1 = dZ1.T @ A["X1"]
    g = np.zeros(NPARAM, np.float64)
    g[OFF_W1:OFF_B1] = g_W1.ravel()
    g[OFF_B1:OFF_W2] = g_b1
    g[OFF_W2:OFF_B2] = g_W2.ravel()
    g[OFF_B2:OFF_W3] = g_b2
    g[OFF_W3:OFF_B3] = g_w3
    g[OFF_B3] = g_b3
    return g, A
