"""CPU: pin the network reference (oracle/mpnn.py): its analytical backward pass against finite
differences of the float64 loss, and its bf16 rounding helper against exact values."""
import numpy as np
import pytest

import gpu4gst_code_amd as gst
from oracle import mpnn


def _case(seed, n=60, e=180):
    G = gst.gen_random_connected(n, e, 1, 30, seed=seed)
    groups = gst.gen_groups(n, 5, 2, 6, seed=seed + 1)
    gm = groups.masks(n)
    rng = np.random.default_rng(seed)
    y = (rng.random(n) < 0.2).astype(np.uint8)
    p = rng.normal(0, 0.3, mpnn.NPARAM)
    p[mpnn.OFF_B1:mpnn.OFF_W2] += 0.2   # keep most ReLUs active
    p[mpnn.OFF_B2:mpnn.OFF_W3] += 0.2
    return G, gm, y, p


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-5, 65504.0], np.float32)
    r = mpnn.bf16(x)
    # exactly representable values stay; 1.005859375 = 1 + 3/512 rounds to 1 + 1/128 (nearest even)
    assert r[0] == 1.0 and r[1] == np.float32(1.0078125) or r[1] == np.float32(1.0)
    assert r[3] == -2.5
    u = r.view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_backward_matches_finite_differences(seed):
    G, gm, y, p = _case(seed)
    g, _ = mpnn.grads(p, G.row_ptr, G.col, G.w, gm, y, 4.0, rnd=False)
    rng = np.random.default_rng(seed + 100)
    idx = np.concatenate([rng.choice(np.arange(mpnn.OFF_W1, mpnn.OFF_B1), 12, replace=False),
                          rng.choice(np.arange(mpnn.OFF_B1, mpnn.OFF_W2), 4, replace=False),
                          rng.choice(np.arange(mpnn.OFF_W2, mpnn.OFF_B2), 12, replace=False),
                          rng.choice(np.arange(mpnn.OFF_B2, mpnn.OFF_W3), 4, replace=False),
                          rng.choice(np.arange(mpnn.OFF_W3, mpnn.OFF_B3), 4, replace=False), [mpnn.OFF_B3]])
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mpnn.loss(pp, G.row_ptr, G.col, G.w, gm, y, 4.0) - mpnn.loss(pm, G.row_ptr, G.col, G.w, gm, y, 4.0)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * max(1e-3, abs(fd)) + 1e-7, (i, fd, g[i])


def test_planted_backward_mistake_is_caught():
    """Dropping the transposed aggregation term (a plausible backward bug) fails the FD check."""
    G, gm, y, p = _case(4)
    g, A = mpnn.grads(p, G.row_ptr, G.col, G.w, gm, y, 4.0, rnd=False)
    # recompute the W1 gradient without the neighbour path through layer 2
    W1, b1, W2, b2, w3, b3 = mpnn.unpack(p)
    f = np.float32
    n = A["s"].shape[0]
    wt = (np.where(y > 0, 4.0, 1.0) / n).astype(f)
    dS = wt * ((1 / (1 + np.exp(-A["s"]))) - y)
    dZ2 = np.where(A["h2"] > 0, dS[:, None] * w3[None, :], 0)
    dX2 = dZ2 @ W2
    dZ1 = np.where(A["h1"] > 0, dX2[:, :64], 0)        # missing the aggregation transpose
    bad = (dZ1.T @ A["X1"]).ravel()
    i = int(np.argmax(np.abs(g[mpnn.OFF_W1:mpnn.OFF_B1] - bad)))
    h = 1e-5
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (mpnn.loss(pp, G.row_ptr, G.col, G.w, gm, y, 4.0) - mpnn.loss(pm, G.row_ptr, G.col, G.w, gm, y, 4.0)) / (2 * h)
    assert abs(fd - g[i]) < 1e-4 * max(1e-3, abs(fd)) + 1e-7
    assert abs(fd - bad[i]) > 1e-3 * max(1e-3, abs(fd))
