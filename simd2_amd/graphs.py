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
