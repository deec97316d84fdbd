"""CPU oracle of the learned schedule -- TEST INFRASTRUCTURE ONLY.

The reference (escalab/SIMD2) has no learned schedule, so nothing here
restates a reference file; this is the specification the GPU kernels of
simd2_amd/csrc/learned.hip are checked against, in float64 numpy:

  * graph_features: the 24 features, from their definitions;
  * levels / order: breadth-first levels from a pseudo-peripheral vertex;
  * mlp_forward / mlp_loss_grad: the 24 -> 64 -> 64 -> 2 ReLU network and
    the gradient of L = (1/B) sum_b sum_o w_o (y_bo - t_bo)^2;
  * adam_step: Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).

Pinning (tests/test_learned_oracle.py): the analytic gradient against
central finite differences; BFS depths and reach against
scipy.sparse.csgraph unweighted shortest paths; degree / weight / locality
statistics against direct numpy expressions on hand-built graphs (a path, a
star, a complete graph) whose values are known in closed form.

Only tests/ import this module.
"""
from __future__ import annotations

import math

import numpy as np

NF, NH, NO = 24, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NO * NH
NPARAM = OFF_B3 + NO
FLT_MAX = float(np.finfo(np.float32).max)
NSRC = 8


def edge_mask(X):
    """Edges: off-diagonal entries that are finite weights (not an identity
    +-inf, not the FLT_MAX 'no edge' fill)."""
    X = np.asarray(X, np.float32)
    m = np.abs(X.astype(np.float64)) < FLT_MAX
    np.fill_diagonal(m, False)
    return m


def bfs(adj, src):
    """Levels of a breadth-first search over a boolean adjacency (row i =
    out-edges of i); -1 = unreached.  Returns (levels, depth, reached)."""
    n = adj.shape[0]
    level = np.full(n, -1, np.int64)
    level[src] = 0
    frontier = np.zeros(n, bool)
    frontier[src] = True
    visited = frontier.copy()
    depth = 0
    while True:
        nxt = adj[frontier].any(axis=0) & ~visited
        if not nxt.any():
            break
        depth += 1
        level[nxt] = depth
        visited |= nxt
        frontier = nxt
    return level, depth, int(visited.sum())


def graph_features(op, X, epi):
    X = np.asarray(X, np.float32)
    n = X.shape[0]
    E = edge_mask(X)
    W = np.where(E, np.abs(X.astype(np.float64)), 0.0)
    W2 = np.where(E, X.astype(np.float64) ** 2, 0.0)
    deg = E.sum(1).astype(np.float64)
    indeg = E.sum(0).astype(np.float64)
    m = deg.sum()
    mean = m / n
    sd_out = math.sqrt(max((deg ** 2).sum() / n - mean * mean, 0.0))
    sd_in = math.sqrt(max((indeg ** 2).sum() / n - mean * mean, 0.0))
    density = m / (n * (n - 1)) if n > 1 else 0.0
    wmax = W.max() if m > 0 else 0.0
    wmean = W.sum() / m if m > 0 else 0.0
    wsd = math.sqrt(max(W2.sum() / m - wmean * wmean, 0.0)) if m > 0 else 0.0
    wmin = np.abs(X.astype(np.float64))[E].min() if m > 0 else 0.0
    ii, jj = np.nonzero(E)
    d = np.abs(ii - jj).astype(np.float64)
    step = max(1, n // 256)
    rows = np.arange(0, n, step)
    se = E[rows].sum()
    ss = (E[rows] & E.T[rows]).sum()
    depths, reach, width = [], 0.0, 0.0
    for k in range(NSRC):
        src = (k * n) // NSRC
        _, dep, r = bfs(E, src)
        depths.append(dep)
        reach += r / n / NSRC
        width += r / (dep + 1) / n / NSRC
    dmean = sum(depths) / NSRC
    dmx = max(depths)
    f = np.zeros(NF)
    f[0] = math.log2(n) / 16
    f[1] = density
    f[2] = (math.log10(density + 1e-9) + 9) / 9
    f[3] = math.log2(1 + mean) / 16
    f[4] = sd_out / (1 + mean)
    f[5] = math.log2(1 + deg.max()) / 16
    f[6] = (deg == 0).sum() / n
    f[7] = (indeg == 0).sum() / n
    f[8] = sd_in / (1 + mean)
    f[9] = wmean / wmax if wmax > 0 else 0.0
    f[10] = min(wsd / (wmean + 1e-12), 10.0) if m > 0 else 0.0
    f[11] = wmin / wmax if (m > 0 and wmax > 0) else 0.0
    f[12] = math.log10(1 + wmax) / 10
    f[13] = ss / se if se > 0 else 0.0
    f[14] = d.sum() / m / n if m > 0 else 0.0
    f[15] = d.max() / n if m > 0 else 0.0
    f[16] = (d <= 16).sum() / m if m > 0 else 0.0
    f[17] = math.log2(1 + dmean) / 16
    f[18] = math.log2(1 + dmx) / 16
    f[19] = reach
    f[20] = width
    f[21] = 1.0 if op in ("max_plus", "max_min") else 0.0
    f[22] = 1.0 if epi else 0.0
    f[23] = (math.ceil(math.log2(max(dmx, 1))) + 1) / 16
    return f


def levels(X):
    """Levels of the search from the lowest-numbered vertex of the last level
    of a search from vertex 0."""
    E = edge_mask(X)
    lv0, dep, _ = bfs(E, 0)
    far = int(np.nonzero(lv0 == dep)[0].min())
    lv, _, _ = bfs(E, far)
    return lv


def order(X):
    lv = levels(X)
    n = lv.shape[0]
    key = np.where(lv < 0, 32768 + 1, lv)
    return np.argsort(key, kind="stable").astype(np.int32)


def unpack(P):
    P = np.asarray(P, np.float64)
    W1 = P[OFF_W1:OFF_B1].reshape(NH, NF)
    b1 = P[OFF_B1:OFF_W2]
    W2 = P[OFF_W2:OFF_B2].reshape(NH, NH)
    b2 = P[OFF_B2:OFF_W3]
    W3 = P[OFF_W3:OFF_B3].reshape(NO, NH)
    b3 = P[OFF_B3:NPARAM]
    return W1, b1, W2, b2, W3, b3


def mlp_forward(P, x):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3


def mlp_loss(P, x, t, w):
    y = mlp_forward(P, x)
    B = x.shape[0]
    return float((np.asarray(w, np.float64) * (y - t) ** 2).sum() / B)


def mlp_loss_grad(P, x, t, w):
    """(loss, gradient) by backpropagation; ReLU derivative 1 where the
    activation is positive, 0 elsewhere."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    w = np.asarray(w, np.float64)
    B = x.shape[0]
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ W3.T + b3
    loss = float((w * (y - t) ** 2).sum() / B)
    dy = 2.0 * w * (y - t) / B
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ W3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def finite_difference_grad(P, x, t, w, idx, eps=1e-6):
    """Central differences of mlp_loss at the parameter indices idx."""
    P = np.asarray(P, np.float64).copy()
    out = np.zeros(len(idx))
    for k, i in enumerate(idx):
        old = P[i]
        P[i] = old + eps
        lp = mlp_loss(P, x, t, w)
        P[i] = old - eps
        lm = mlp_loss(P, x, t, w)
        P[i] = old
        out[k] = (lp - lm) / (2 * eps)
    return out


def adam_step(P, m, v, g, step, lr):
    m = 0.9 * m + 0.1 * g
    v = 0.999 * v + 0.001 * g * g
    mh = m / (1 - 0.9 ** step)
    vh = v / (1 - 0.999 ** step)
    return P - lr * mh / (np.sqrt(vh) + 1e-8), m, v


def init_params(seed):
    """He-uniform initial weights, zero biases."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM)
    for off, fan_in, rows in ((OFF_W1, NF, NH), (OFF_W2, NH, NH), (OFF_W3, NH, NO)):
        lim = math.sqrt(6.0 / fan_in)
        P[off:off + rows * fan_in] = rng.uniform(-lim, lim, rows * fan_in)
    return P
