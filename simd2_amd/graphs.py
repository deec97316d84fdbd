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
