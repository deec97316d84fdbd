"""Write the golden fixtures under tests/golden/ from the reference's own CPU oracle.

Run in a container that holds the reference tree (after ``make -C oracle``):

    python tests/golden/make_golden.py

What it produces (all data, no code):

* ``host_srgemm_cases.npz`` -- inputs and outputs of the reference's
  ``mmo::host_srgemm`` (kernel/host_srgemm.h:16-49, compiled into
  oracle/_ref/) for all nine operators, with and without the epilogue, on the
  inputs tests/cuAsr_kernel_test.cu builds: shapes and A seeds drawn from its
  unseeded glibc ``rand()`` sequence (:416-428), B seeded 3080 and C seeded 3090
  through the reference's ``rng_init_matrix`` (utils/init_mat.h:3-11), or_and
  on ``(int)x % 2`` (:297-314), D pre-filled as that test pre-fills it.
  A few edge shapes (K=0, 1x1x1, empty) use the true additive identity.
* ``converge_cases.npz`` -- the repeated-squaring driver
  (apps/*/cuASR/*_cuASR.cu, kernel/converge.cuh) run with the reference's
  host_srgemm as the squaring, on the reference's small graph fixtures copied
  into graphs/, with each app's fill convention.  For n=256 only the iteration
  count and a SHA-256 of the result bytes are stored.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

FLT_MAX = np.float32(np.finfo(np.float32).max)
FLT_MIN = np.float32(np.finfo(np.float32).tiny)  # numeric_limits<float>::min()

# The reference test's D pre-fill per op (tests/cuAsr_kernel_test.cu:67-69,
# 125-127, ...): FLT_MAX for min-based, numeric_limits<float>::min() for
# max-based, 0 for plus / or.
REF_TEST_PREFILL = {
    "min_plus": FLT_MAX, "plus_mul": np.float32(0), "max_plus": FLT_MIN,
    "max_min": FLT_MIN, "max_mul": FLT_MIN, "min_mul": FLT_MAX,
    "min_max": FLT_MAX, "minus_square": np.float32(0),
}

# Cases taken from the reference test's 25-shape sequence (indices into it),
# chosen small enough to commit.
SEQ_CASES = [6, 8, 12, 16]
EDGE_SHAPES = [(1, 1, 1), (7, 9, 0), (0, 5, 3), (3, 0, 4), (17, 33, 65)]


def ref_case(M, N, K, a_seed, epi, out, tag, prefill):
    A = oracle.ref_rng_init_matrix(M * K, a_seed).reshape(M, K)
    B = oracle.ref_rng_init_matrix(K * N, 3080).reshape(K, N)
    C = oracle.ref_rng_init_matrix(M * N, 3090).reshape(M, N)
    out[f"{tag}/A"] = A
    out[f"{tag}/B"] = B
    out[f"{tag}/C"] = C
    out[f"{tag}/shape"] = np.array([M, N, K, int(epi)], np.int32)
    for op in oracle.OPS:
        if op == "or_and":
            Ai = (A.astype(np.int32) % 2).astype(np.int32)
            Bi = (B.astype(np.int32) % 2).astype(np.int32)
            Ci = (C.astype(np.int32) % 2).astype(np.int32)
            D0 = np.zeros((M, N), np.int32)
            out[f"{tag}/or_and"] = oracle.ref_srgemm_i32(Ai, Bi, Ci, D0, epi)
            continue
        fill = prefill(op)
        D0 = np.full((M, N), fill, np.float32)
        D = oracle.ref_srgemm_f32(op, A, B, C, D0, epi)
        if K > 0 and prefill is ref_prefill:
            # The true identity gives the same result on these inputs.
            Dt = oracle.ref_srgemm_f32(op, A, B, C,
                                       np.full((M, N), oracle.add_identity(op), np.float32), epi)
            assert np.array_equal(D.view(np.uint32), Dt.view(np.uint32)), (tag, op)
        out[f"{tag}/{op}"] = D


def ref_prefill(op):
    return REF_TEST_PREFILL[op]


def true_prefill(op):
    return oracle.add_identity(op)


def read_edges(path):
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
    """Adjacency with each app's fill convention (§ apps/<app>/cuASR mains)."""
    if app == "apsp":  # apps/apsp/cuASR/apsp_cuASR.cu:83-91
        X = np.full((v, v), FLT_MAX, np.float32)
        for u, w_, w in edges:
            X[u, w_] = w
        np.fill_diagonal(X, 0)
    elif app == "mcp":  # apps/mcp/cuASR/mcp_cuASR.cu:87-96
        X = np.full((v, v), -FLT_MAX, np.float32)
        for u, w_, w in edges:
            X[u, w_] = w
        np.fill_diagonal(X, -FLT_MAX)
    elif app == "mst":  # apps/mst/cuASR/mst_cuASR.cu (FLT_MAX fill, diagonal 0)
        X = np.full((v, v), FLT_MAX, np.float32)
        for u, w_, w in edges:
            X[u, w_] = w
            X[w_, u] = w
        np.fill_diagonal(X, 0)
    elif app == "gtc":  # apps/gtc/cuASR/gtc_cuASR.cu:96-111
        X = np.zeros((v, v), np.int32)
        for u, w_, w in edges:
            X[u, w_] = 1
        np.fill_diagonal(X, 1)
    else:
        raise ValueError(app)
    return X


APP_OP = {"apsp": ("min_plus", True), "mcp": ("max_min", True),
          "mst": ("min_max", False), "gtc": ("or_and", True)}


def ref_converge(X, app):
    """The driver loop of apps/<app>/cuASR with the reference host_srgemm as
    the squaring and comp_update (kernel/converge.cuh) as the check."""
    op, epi = APP_OP[app]
    X = X.copy()
    itr, run = 0, True
    while run and itr < 500:
        itr += 1
        if op == "or_and":
            Xd = oracle.ref_srgemm_i32(X, X, X, np.zeros_like(X), epi)
            run = bool(np.any(X != Xd))
        else:
            Xd = oracle.ref_srgemm_f32(op, X, X, X,
                                       np.full(X.shape, oracle.add_identity(op), np.float32), epi)
            with np.errstate(invalid="ignore", over="ignore"):
                run = bool(np.any(np.abs(X - Xd) > 0.0001))
        X = Xd
    return X, itr


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref not built: run make -C oracle with the reference tree present")
    out = {}
    seq = oracle.ref_test_sequence(25)
    for idx in SEQ_CASES:
        M, N, K, s_noepi, s_epi = (int(x) for x in seq[idx])
        ref_case(M, N, K, s_noepi, False, out, f"seq{idx}_noepi", ref_prefill)
        ref_case(M, N, K, s_epi, True, out, f"seq{idx}_epi", ref_prefill)
    for j, (M, N, K) in enumerate(EDGE_SHAPES):
        for epi in (False, True):
            ref_case(M, N, K, 1000 + j, epi, out, f"edge{j}_{'epi' if epi else 'noepi'}",
                     true_prefill)
    np.savez_compressed(os.path.join(HERE, "host_srgemm_cases.npz"), **out)

    conv = {}
    for gname in ("16_LD", "rmat_16", "rmat_256"):
        v, edges = read_edges(os.path.join(HERE, "graphs", f"{gname}.txt"))
        for app in APP_OP:
            X0 = app_matrix(v, edges, app)
            Xf, itr = ref_converge(X0, app)
            key = f"{gname}/{app}"
            conv[f"{key}/itr"] = np.array([itr], np.int32)
            conv[f"{key}/sha256"] = np.frombuffer(
                hashlib.sha256(np.ascontiguousarray(Xf).tobytes()).digest(), np.uint8)
            if v <= 16:
                conv[f"{key}/out"] = Xf
    np.savez_compressed(os.path.join(HERE, "converge_cases.npz"), **conv)
    for k in sorted(conv):
        if k.endswith("/itr"):
            print(k, int(conv[k][0]))


if __name__ == "__main__":
    main()
