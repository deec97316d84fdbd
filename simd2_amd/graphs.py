"""Seeded synthetic graph generators on the GPU (torch), in the adjacency
conventions of the reference apps: weights as FP32, a fill value where there
is no edge, a fixed diagonal (apps/apsp/cuASR/apsp_cuASR.cu:83-91)."""
from __future__ import annotations


def er_graph(torch, n, p, seed, wlo=1, whi=100, fill=float("inf"), diag=0.0, device="cuda"):
    """Directed Erdos-Renyi graph: each edge present with probability p,
    integer weight uniform in [wlo, whi]."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    mask = torch.rand((n, n), generator=g, device=device) < p
    w = torch.randint(wlo, whi + 1, (n, n), generator=g, device=device).float()
    X = torch.where(mask, w, torch.full_like(w, fill))
    X.fill_diagonal_(diag)
    return X


FLT_MAX = 3.4028234663852886e38

# Problem mix of the batched workload: (op, do_epilogue, check, share).
BATCH_MIX = [("min_plus", True, "threshold", 0.4),   # APSP
             ("max_min", True, "threshold", 0.2),    # MCP
             ("min_max", False, "threshold", 0.2),   # MST minimax
             ("or_and", True, "exact", 0.2)]         # transitive closure


def app_graph(torch, op, n, p, seed, wlo=1, whi=100, device="cuda"):
    """One Erdos-Renyi graph in the fill convention of the reference app that
    uses op: APSP +inf / diagonal 0, MCP -FLT_MAX / diagonal -FLT_MAX, MST
    symmetric FLT_MAX / diagonal 0, GTC bit-packed reachability with the
    diagonal set (int32 words, ceil(n/32) per row)."""
    from . import pack_bits
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    mask = torch.rand((n, n), generator=g, device=device) < p
    w = torch.randint(wlo, whi + 1, (n, n), generator=g, device=device).float()
    if op == "min_max":
        mask = torch.triu(mask, 1)
        mask = mask | mask.T
        w = torch.triu(w, 1)
        w = w + w.T
    if op == "or_and":
        m = mask.clone()
        m.fill_diagonal_(True)
        return pack_bits(m.to(torch.int32))
    if op == "max_min":
        X = torch.where(mask, w, torch.full_like(w, -FLT_MAX))
        X.fill_diagonal_(-FLT_MAX)
        return X
    fill = float("inf") if op == "min_plus" else FLT_MAX
    X = torch.where(mask, w, torch.full_like(w, fill))
    X.fill_diagonal_(0.0)
    return X


def batch_specs(count, seed, nmin=32, nmax=512, pmin=0.01, pmax=0.2, mix=None):
    """Seeded (n, op, do_epilogue, check, p, graph_seed) per graph: n uniform
    in [nmin, nmax], p uniform in [pmin, pmax], ops drawn by the mix shares."""
    import numpy as np
    mix = BATCH_MIX if mix is None else mix
    rng = np.random.default_rng(seed)
    shares = np.array([m[3] for m in mix], np.float64)
    kinds = rng.choice(len(mix), size=count, p=shares / shares.sum())
    ns = rng.integers(nmin, nmax + 1, size=count)
    ps = rng.uniform(pmin, pmax, size=count)
    out = []
    for k in range(count):
        op, epi, check, _ = mix[kinds[k]]
        out.append((int(ns[k]), op, epi, check, float(ps[k]), int(seed * 100003 + k)))
    return out


def _fill_value(op):
    return {"min_plus": float("inf"), "max_min": float("-inf"),
            "min_max": float("inf"), "max_plus": float("-inf")}[op]


def _finish(torch, op, mask, w, diag=True):
    """Adjacency of `op` from an edge mask and weights: no edge = the
    (+)-identity; diagonal 0 for the min-based ops (APSP, MST), left as no
    edge for max_min (MCP) and max_plus (APLP)."""
    X = torch.where(mask, w, torch.full_like(w, _fill_value(op)))
    if diag and op in ("min_plus", "min_max"):
        X.fill_diagonal_(0.0)
    return X


def er_family(torch, op, n, p, seed, wlo=1, whi=100, device="cuda"):
    """Directed Erdos-Renyi (symmetric for min_max), integer weights."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    mask = torch.rand((n, n), generator=g, device=device) < p
    w = torch.randint(wlo, whi + 1, (n, n), generator=g, device=device).float()
    if op == "min_max":
        mask = torch.triu(mask, 1)
        mask = mask | mask.T
        w = torch.triu(w, 1)
        w = w + w.T
    mask.fill_diagonal_(False)
    return _finish(torch, op, mask, w)


def grid_family(torch, op, n, seed, diag_p=0.3, wlo=1, whi=1000, device="cuda"):
    """2D grid of n vertices, floor(sqrt(n)) per row (the last row partial),
    numbered row by row; each vertex is joined to its right and lower
    neighbours and, with probability diag_p, to one of its two lower diagonal
    neighbours; undirected, integer weights in [wlo, whi]."""
    import math
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    cols = max(1, int(math.isqrt(n)))
    idx = torch.arange(n, device=device)
    r, c = idx // cols, idx % cols
    src, dst = [], []
    right = (c + 1 < cols) & (idx + 1 < n)
    src.append(idx[right]); dst.append(idx[right] + 1)
    down = idx + cols < n
    src.append(idx[down]); dst.append(idx[down] + cols)
    pick = torch.rand(n, generator=g, device=device) < diag_p
    left = torch.rand(n, generator=g, device=device) < 0.5
    dr = pick & ~left & (c + 1 < cols) & (idx + cols + 1 < n)
    dl = pick & left & (c > 0) & (idx + cols - 1 < n)
    src.append(idx[dr]); dst.append(idx[dr] + cols + 1)
    src.append(idx[dl]); dst.append(idx[dl] + cols - 1)
    s = torch.cat(src); d = torch.cat(dst)
    wts = torch.randint(wlo, whi + 1, (s.numel(),), generator=g, device=device).float()
    mask = torch.zeros((n, n), dtype=torch.bool, device=device)
    w = torch.zeros((n, n), device=device)
    mask[s, d] = True; mask[d, s] = True
    w[s, d] = wts; w[d, s] = wts
    return _finish(torch, op, mask, w)


def rgg_family(torch, op, n, degree, seed, wlo=1, whi=100, device="cuda"):
    """Random geometric graph: n points uniform in the unit square (vertex ids
    in the order drawn), joined when closer than the radius that gives the
    requested mean degree (pi r^2 n = degree); undirected, integer weights."""
    import math
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    pts = torch.rand((n, 2), generator=g, device=device)
    rad = math.sqrt(degree / (math.pi * n))
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    mask = d2 < rad * rad
    mask.fill_diagonal_(False)
    w = torch.randint(wlo, whi + 1, (n, n), generator=g, device=device).float()
    w = torch.triu(w, 1)
    w = w + w.T
    return _finish(torch, op, mask, w)
