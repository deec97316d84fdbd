"""Seeded synthetic point sets on the GPU (torch) for the kNN workloads."""
from __future__ import annotations


def gaussian_clusters(torch, n, dim, clusters=64, seed=0, center_std=10.0, sigma=1.0,
                      centers=None, device="cuda"):
    """n points in `clusters` Gaussian clusters: centres ~ N(0, center_std^2)
    per coordinate, points ~ N(centre, sigma^2).  Returns (points, centres)."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    if centers is None:
        centers = torch.randn((clusters, dim), generator=g, device=device) * center_std
    lab = torch.randint(0, centers.shape[0], (n,), generator=g, device=device)
    pts = centers[lab] + torch.randn((n, dim), generator=g, device=device) * sigma
    return pts.contiguous(), centers
