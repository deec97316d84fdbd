"""CPU tests of the oracle (the C restatement of kernel/host_srgemm.h and
kernel/converge.cuh): pinned against the reference's own host_srgemm (golden
fixtures, and the live build in oracle/_ref where the reference tree exists),
closed forms, cross-algorithm checks and FP64 tolerance."""
import numpy as np
import pytest

from util import (APP_OP, EXACT_OPS, FLT_MAX, OPS_F32, app_matrix, er_graph, golden_cases,
                  read_edges)


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32 if x.dtype == np.float32 else np.int32)


# ---- against the reference's own host_srgemm ---------------------------------

def test_restatement_matches_golden(oracle_mod, golden_srgemm):
    g = golden_srgemm
    n = 0
    for tag, M, N, K, epi in golden_cases(g):
        A, B, C = g[f"{tag}/A"], g[f"{tag}/B"], g[f"{tag}/C"]
        for op in OPS_F32:
            D = oracle_mod.srgemm_f32(op, A, B, C, epi)
            assert np.array_equal(bits(D), bits(g[f"{tag}/{op}"])), (tag, op)
            n += 1
        Ai, Bi, Ci = ((x.astype(np.int32) % 2).astype(np.int32) for x in (A, B, C))
        Di = oracle_mod.srgemm_i32(Ai, Bi, Ci, epi)
        assert np.array_equal(Di, g[f"{tag}/or_and"]), (tag, "or_and")
    assert n > 50


@pytest.mark.skipif("not __import__('oracle').ref_available()",
                    reason="reference tree not present (oracle/_ref not built)")
def test_restatement_matches_reference_all_test_shapes(oracle_mod):
    """All 25 shapes of tests/cuAsr_kernel_test.cu:416-428 with its inputs,
    every op, both epilogue settings, against the live reference build."""
    seq = oracle_mod.ref_test_sequence(25)
    for M, N, K, s0, s1 in seq:
        M, N, K = int(M), int(N), int(K)
        for epi, seed in ((False, int(s0)), (True, int(s1))):
            A = oracle_mod.ref_rng_init_matrix(M * K, seed).reshape(M, K)
            B = oracle_mod.ref_rng_init_matrix(K * N, 3080).reshape(K, N)
            C = oracle_mod.ref_rng_init_matrix(M * N, 3090).reshape(M, N)
            for op in OPS_F32:
                D0 = np.full((M, N), oracle_mod.add_identity(op), np.float32)
                want = oracle_mod.ref_srgemm_f32(op, A, B, C, D0, epi)
                got = oracle_mod.srgemm_f32(op, A, B, C, epi, threads=4)
                assert np.array_equal(bits(want), bits(got)), (M, N, K, epi, op)
            Ai, Bi, Ci = ((x.astype(np.int32) % 2).astype(np.int32) for x in (A, B, C))
            want = oracle_mod.ref_srgemm_i32(Ai, Bi, Ci, np.zeros((M, N), np.int32), epi)
            assert np.array_equal(want, oracle_mod.srgemm_i32(Ai, Bi, Ci, epi))


def test_openmp_rows_bit_identical(oracle_mod):
    rng = np.random.default_rng(3)
    A = rng.random((37, 53), dtype=np.float32)
    B = rng.random((53, 29), dtype=np.float32)
    C = rng.random((37, 29), dtype=np.float32)
    for op in OPS_F32:
        a = oracle_mod.srgemm_f32(op, A, B, C, True, threads=1)
        b = oracle_mod.srgemm_f32(op, A, B, C, True, threads=5)
        assert np.array_equal(bits(a), bits(b)), op


# ---- closed forms -------------------------------------------------------------

IDENTITY = {  # (diagonal, off-diagonal) of the semiring identity matrix
    "min_plus": (0.0, np.inf), "plus_mul": (1.0, 0.0), "max_plus": (0.0, -np.inf),
    "max_min": (np.inf, -np.inf), "max_mul": (1.0, -np.inf), "min_mul": (1.0, np.inf),
    "min_max": (-np.inf, np.inf),
}


@pytest.mark.parametrize("op", list(IDENTITY))
def test_identity_matrix(oracle_mod, op):
    rng = np.random.default_rng(1)
    A = rng.integers(1, 100, (23, 17)).astype(np.float32)
    d, o = IDENTITY[op]
    I = np.full((17, 17), o, np.float32)
    np.fill_diagonal(I, d)
    assert np.array_equal(oracle_mod.srgemm_f32(op, A, I), A)
    I2 = np.full((23, 23), o, np.float32)
    np.fill_diagonal(I2, d)
    assert np.array_equal(oracle_mod.srgemm_f32(op, I2, A), A)


@pytest.mark.parametrize("op", EXACT_OPS)
def test_annihilator(oracle_mod, op):
    rng = np.random.default_rng(2)
    A = rng.integers(1, 100, (9, 11)).astype(np.float32)
    C = rng.integers(1, 100, (9, 7)).astype(np.float32)
    ident = oracle_mod.add_identity(op)
    if op in ("max_mul", "min_mul"):
        pytest.skip("inf * finite is not absorbing for the multiplicative ops")
    B = np.full((11, 7), ident, np.float32)
    assert np.array_equal(oracle_mod.srgemm_f32(op, A, B, C, True), C)
    assert np.all(oracle_mod.srgemm_f32(op, A, B, C, False) == ident)


def test_plus_mul_vs_fp64(oracle_mod):
    rng = np.random.default_rng(4)
    A = rng.integers(1, 101, (64, 300)).astype(np.float32)
    B = rng.integers(1, 101, (300, 48)).astype(np.float32)
    D = oracle_mod.srgemm_f32("plus_mul", A, B)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.max(np.abs(D - ref) / ref) <= 1e-5
    # minus_square: sum_k (a_ik - b_kj)^2
    Dm = oracle_mod.srgemm_f32("minus_square", A, B)
    refm = ((A.astype(np.float64)[:, :, None] - B.astype(np.float64)[None]) ** 2).sum(1)
    assert np.max(np.abs(Dm - refm) / refm) <= 1e-5


def test_or_and_is_boolean_product(oracle_mod):
    rng = np.random.default_rng(5)
    A = (rng.random((19, 40)) < 0.2).astype(np.int32)
    B = (rng.random((40, 23)) < 0.2).astype(np.int32)
    got = oracle_mod.srgemm_i32(A, B)
    assert np.array_equal(got, ((A.astype(np.int64) @ B) > 0).astype(np.int32))


def test_add_identity_values(oracle_mod):
    assert oracle_mod.add_identity("min_plus") == np.inf
    assert oracle_mod.add_identity("max_min") == -np.inf
    assert oracle_mod.add_identity("plus_mul") == 0


# ---- convergence --------------------------------------------------------------

def test_comp_update_semantics(oracle_mod):
    inf = np.float32(np.inf)
    prev = np.array([inf, inf, 1.0, 2.0], np.float32)
    cur = np.array([inf, 5.0, 1.00005, 2.0], np.float32)
    ch, p, c = oracle_mod.comp_update(prev, cur)            # 1e-4 threshold
    assert ch and np.array_equal(p, cur)
    ch, _, _ = oracle_mod.comp_update(np.array([inf, 1.0], np.float32),
                                      np.array([inf, 1.00005], np.float32))
    assert not ch                                            # inf-inf is NaN: unchanged
    ch, _, _ = oracle_mod.comp_update(np.array([1.0], np.float32),
                                      np.array([1.00005], np.float32), "exact")
    assert ch
    ch, p, c = oracle_mod.comp_update(np.array([-FLT_MAX, 3.0], np.float32),
                                      np.array([0.5, 3.0], np.float32), "aplp")
    assert not ch and c[0] == -FLT_MAX and p[0] == -FLT_MAX
    ch, _, _ = oracle_mod.comp_update(np.array([1, 2], np.int32), np.array([1, 3], np.int32))
    assert ch


def test_converge_golden(oracle_mod, golden_converge):
    import hashlib
    g = golden_converge
    for gname, fname in (("16_LD", "16_LD.txt"), ("rmat_16", "rmat_16.txt"),
                         ("rmat_256", "rmat_256.txt")):
        v, edges = read_edges(fname)
        for app, (op, epi) in APP_OP.items():
            X0 = app_matrix(v, edges, app)
            if op == "or_and":
                X, itr = oracle_mod.converge_i32(X0)
            else:
                X, itr = oracle_mod.converge_f32(op, X0, epilogue=epi, check="threshold")
            key = f"{gname}/{app}"
            assert itr == int(g[f"{key}/itr"][0]), key
            assert hashlib.sha256(X.tobytes()).digest() == g[f"{key}/sha256"].tobytes(), key


def test_path_graph_closed_form(oracle_mod):
    """apps/data/Dia_input/16_LD.txt is the chain i -> i+1: distances are
    prefix-sum differences and the loop takes ceil(log2(n-1)) + 1 squarings."""
    v, edges = read_edges("16_LD.txt")
    X0 = app_matrix(v, edges, "apsp")
    X, itr = oracle_mod.converge_f32("min_plus", X0, check="threshold")
    w = np.zeros(v)
    for u, t, wt in edges:
        assert t == u + 1
        w[t] = wt
    pref = np.cumsum(w)
    for i in range(v):
        for j in range(v):
            want = 0.0 if i == j else (pref[j] - pref[i] if j > i else FLT_MAX)
            assert X[i, j] == np.float32(want)
    assert itr == int(np.ceil(np.log2(v - 1))) + 1


def test_apsp_matches_dijkstra(oracle_mod):
    from scipy.sparse.csgraph import dijkstra
    X0 = er_graph(96, 0.06, seed=11)
    X, itr = oracle_mod.converge_f32("min_plus", X0, check="exact", threads=4)
    W = np.where(np.isinf(X0), 0, X0).astype(np.float64)
    D = dijkstra(W, directed=True)
    assert np.array_equal(X.astype(np.float64), D)
    assert itr >= 2


def test_mcp_matches_widest_path(oracle_mod):
    """max_min squaring = widest (bottleneck) path; brute-force check by a
    Floyd-Warshall-style max-min closure on a small graph."""
    n = 24
    X0 = er_graph(n, 0.15, seed=12, fill=-np.inf, diag=-np.inf)
    X, _ = oracle_mod.converge_f32("max_min", X0, check="exact")
    W = X0.astype(np.float64).copy()
    for k in range(n):
        W = np.maximum(W, np.minimum(W[:, k:k + 1], W[k:k + 1, :]))
    assert np.array_equal(X.astype(np.float64), W)


def test_converge_cap(oracle_mod):
    """max_plus on a positive cycle never converges: the loop stops at the cap."""
    X0 = np.array([[0, 1], [1, 0]], np.float32)
    _, itr = oracle_mod.converge_f32("max_plus", X0, check="exact", max_itr=7)
    assert itr == 7


# ---- FP16 emulation restatement (kernel/tensor_srgemm.cuh) ---------------------

def test_fp16_emulation_exact_inputs_equal_fp32_oracle(oracle_mod):
    """Small integers make every half operation exact, so the emulation must
    equal host_srgemm (max-based results are >= 0 here, where the 0
    initial value is harmless)."""
    from oracle import fp16_emulation as emu
    rng = np.random.default_rng(8)
    A = rng.integers(0, 16, (32, 48)).astype(np.float32)
    B = rng.integers(0, 16, (48, 16)).astype(np.float32)
    C = rng.integers(0, 16, (32, 16)).astype(np.float32)
    for op in ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul", "min_max",
               "minus_square"]:
        for epi in (False, True):
            want = oracle_mod.srgemm_f32(op, A, B, C, epi)
            got = emu.tensor_srgemm(A.astype(np.float16), B.astype(np.float16), C, op, epi)
            assert np.array_equal(got, want), (op, epi)


def test_fp16_emulation_quirks():
    from oracle import fp16_emulation as emu
    # max-based accumulators start at __float2half(FLT_MIN) == 0: all-negative
    # products are clamped to 0.
    A = np.full((16, 16), -2.0, np.float16)
    B = np.full((16, 16), -3.0, np.float16)
    C = np.zeros((16, 16), np.float32)
    assert np.all(emu.tensor_srgemm(A, B, C, "max_plus", False) == 0.0)
    # 2048 + 1 is not a half: the sum rounds (to even) before the float min.
    A = np.full((16, 16), 2048.0, np.float16)
    B = np.full((16, 16), 1.0, np.float16)
    assert np.all(emu.tensor_srgemm(A, B, C, "min_plus", False) == 2048.0)
    assert np.all(emu.tensor_srgemm(A, B, C, "max_plus", False) == 2048.0)
    # or_and treats > 0 as true
    A = np.eye(16, dtype=np.float16) * 0.5
    got = emu.tensor_srgemm(A, A, C, "or_and", False)
    assert np.array_equal(got, np.eye(16, dtype=np.float32))


def test_knn_insertion_sort_is_stable_topk(oracle_mod):
    """The restated modified_insertion_sort_linear equals a stable argsort:
    k smallest per column, ties in index order."""
    rng = np.random.default_rng(9)
    D = rng.integers(0, 20, (300, 70)).astype(np.float32)  # many ties
    for k in (1, 5, 16, 32):
        dist, idx = oracle_mod.knn_insertion_sort(D, k, do_sqrt=False)
        order = np.argsort(D, axis=0, kind="stable")[:k]
        assert np.array_equal(idx, order.astype(np.int32)), k
        assert np.array_equal(dist, np.take_along_axis(D, order, axis=0)), k
    dist, idx = oracle_mod.knn_insertion_sort(D, 4, do_sqrt=True)
    assert np.allclose(dist, np.sqrt(np.sort(D, axis=0)[:4]))
