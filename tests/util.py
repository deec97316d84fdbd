"""Helpers shared by the tests: golden-case iteration, graph readers, inputs."""
import os

import numpy as np

from conftest import GOLDEN

OPS_F32 = ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul",
           "min_max", "minus_square"]
EXACT_OPS = ["min_plus", "max_plus", "max_min", "max_mul", "min_mul", "min_max"]
FLT_MAX = np.float32(np.finfo(np.float32).max)


def golden_cases(g):
    tags = sorted({k.split("/")[0] for k in g})
    for t in tags:
        M, N, K, epi = (int(x) for x in g[f"{t}/shape"])
        yield t, M, N, K, bool(epi)


def read_edges(name):
    path = os.path.join(GOLDEN, "graphs", name)
    with open(path) as f:
        toks = f.read().split()
    v, e = int(toks[0]), int(toks[1])
    rest = toks[2:]
    width = 3 if len(rest) >= 3 * e else 2
    edges = []
    for i in range(e):
        u, w_ = int(rest[width * i]), int(rest[width * i + 1])
        w = float(rest[width * i + 2]) if width == 3 else 1.0 + ((u * 131 + w_ * 71) % 100)
        edges.append((u, w_, w))
    return v, edges


def app_matrix(v, edges, app):
    """Adjacency matrix with the fill convention of each reference app."""
    if app == "apsp":
        X = np.full((v, v), FLT_MAX, np.float32)
        for u, w_, w in edges:
            X[u, w_] = w
        np.fill_diagonal(X, 0)
    elif app == "mcp":
        X = np.full((v, v), -FLT_MAX, np.float32)
        for u, w_, w in edges:
            X[u, w_] = w
        np.fill_diagonal(X, -FLT_MAX)
    elif app == "mst":
        X = np.full((v, v), FLT_MAX, np.float32)
        for u, w_, w in edges:
            X[u, w_] = w
            X[w_, u] = w
        np.fill_diagonal(X, 0)
    elif app == "gtc":
        X = np.zeros((v, v), np.int32)
        for u, w_, w in edges:
            X[u, w_] = 1
        np.fill_diagonal(X, 1)
    else:
        raise ValueError(app)
    return X


APP_OP = {"apsp": ("min_plus", True), "mcp": ("max_min", True),
          "mst": ("min_max", False), "gtc": ("or_and", True)}


def er_graph(n, p, seed, wlo=1, whi=100, fill=np.inf, diag=0.0):
    """Directed Erdos-Renyi adjacency: integer weights in [wlo, whi] as FP32,
    `fill` where there is no edge, `diag` on the diagonal."""
    rng = np.random.default_rng(seed)
    mask = rng.random((n, n)) < p
    w = rng.integers(wlo, whi + 1, size=(n, n)).astype(np.float32)
    X = np.where(mask, w, np.float32(fill)).astype(np.float32)
    np.fill_diagonal(X, diag)
    return X


def semiring_inputs(op, M, N, K, seed, holes=0.1):
    """Inputs of the single-srgemm workload: integer-valued U{1..100} with
    +inf (min-based) or -inf (max-based) holes; (0,1] for max_mul/min_mul;
    no holes for plus_mul / minus_square."""
    rng = np.random.default_rng(seed)
    if op in ("max_mul", "min_mul"):
        A = (1.0 - rng.random((M, K))).astype(np.float32)
        B = (1.0 - rng.random((K, N))).astype(np.float32)
        C = (1.0 - rng.random((M, N))).astype(np.float32)
        return A, B, C
    A = rng.integers(1, 101, (M, K)).astype(np.float32)
    B = rng.integers(1, 101, (K, N)).astype(np.float32)
    C = rng.integers(1, 101, (M, N)).astype(np.float32)
    if op in ("plus_mul", "minus_square"):
        return A, B, C
    hole = np.float32(np.inf if op.startswith("min") else -np.inf)
    A[rng.random((M, K)) < holes] = hole
    B[rng.random((K, N)) < holes] = hole
    return A, B, C
