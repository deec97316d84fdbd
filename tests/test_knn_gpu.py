"""GPU kNN: the reference's PLD post-processing (bit-exact vs the oracle) and
the BF16-MFMA precision form (top-k agreement and distance error vs FP64)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fp64_topk(Qp, Rp, k, extra=8, chunk=512):
    """Ground truth: FP64 norm expansion to shortlist, then the direct FP64
    sum of squares, ordered by (distance, index)."""
    Q64, R64 = Qp.astype(np.float64), Rp.astype(np.float64)
    qn, rn = (Q64 ** 2).sum(1), (R64 ** 2).sum(1)
    out_d = np.zeros((len(Q64), k))
    out_i = np.zeros((len(Q64), k), np.int64)
    for s in range(0, len(Q64), chunk):
        D = qn[s:s + chunk, None] + rn[None, :] - 2.0 * (Q64[s:s + chunk] @ R64.T)
        cand = np.argpartition(D, min(k + extra, D.shape[1] - 1), axis=1)[:, :k + extra]
        for row, c in enumerate(cand):
            dd = ((R64[c] - Q64[s + row]) ** 2).sum(1)
            order = np.lexsort((c, dd))[:k]
            out_i[s + row] = c[order]
            out_d[s + row] = dd[order]
    return out_d, out_i


def clustered(torch, Q, R, d, seed):
    from simd2_amd.datasets import gaussian_clusters
    ref, centers = gaussian_clusters(torch, R, d, 64, seed=seed)
    qry, _ = gaussian_clusters(torch, Q, d, 64, seed=seed + 1, centers=centers)
    return qry, ref


@pytest.mark.parametrize("R,Q,k", [(300, 70, 16), (1000, 130, 32), (64, 64, 1), (257, 5, 7),
                                   (2000, 100, 64), (2000, 67, 100), (40, 3, 40), (33, 65, 2),
                                   (3000, 70, 16), (5000, 130, 1), (2049, 65, 8),
                                   (8192, 33, 64)])
def test_knn_select_vs_oracle(gpu, oracle_mod, R, Q, k):
    """Integer distances (ties on purpose; the earlier index wins).  Rows of
    at least 4k groups of 32 take the one-pass collect (groups holding two
    or more candidates read again; 2049 ends in a group of one reference),
    shorter rows the second streaming pass."""
    import simd2_amd.knn as K
    rng = np.random.default_rng(R + Q + k)
    D = rng.integers(0, 50, (R, Q)).astype(np.float32)
    say(f"select R={R} Q={Q} k={k}")
    want_d, want_i = oracle_mod.knn_insertion_sort(D, k, do_sqrt=True)
    gd, gi = K.knn_select(gpu.from_numpy(D).cuda(), k, sqrt=True)
    gpu.cuda.synchronize()
    assert np.array_equal(gi.cpu().numpy(), want_i)
    assert np.array_equal(gd.cpu().numpy(), want_d)


@pytest.mark.parametrize("R,Q,k,kind", [(3000, 70, 700, "uniform"),     # LDS sort of 4096
                                        (10000, 3, 9000, "uniform"),    # global sort, 16384
                                        (5000, 130, 300, "decreasing"),
                                        (4096, 64, 16, "constant"),
                                        (700, 90, 50, "mixed_inf"),
                                        (4100, 40, 32, "mixed_inf"),
                                        (5000, 50, 16, "decreasing")])
def test_knn_select_large_k_and_adversarial_rows(gpu, oracle_mod, R, Q, k, kind):
    """Large k (the candidate sort in LDS and in global memory), rows whose
    distances fall with the index (every element enters the reference's
    list), all-equal rows, and rows with +inf entries."""
    import simd2_amd.knn as K
    rng = np.random.default_rng(R * 3 + k)
    if kind == "uniform":
        D = rng.random((R, Q)).astype(np.float32)
    elif kind == "decreasing":
        D = np.repeat(np.arange(R, 0, -1, dtype=np.float32)[:, None], Q, axis=1)
        D += rng.integers(0, 3, (R, Q)).astype(np.float32)
    elif kind == "constant":
        D = np.full((R, Q), 7.0, np.float32)
    else:
        D = rng.integers(0, 20, (R, Q)).astype(np.float32)
        D[rng.random((R, Q)) < 0.5] = np.inf
    say(f"select {kind} R={R} Q={Q} k={k}")
    want_d, want_i = oracle_mod.knn_insertion_sort(D, k, do_sqrt=False)
    gd, gi = K.knn_select(gpu.from_numpy(D).cuda(), k, sqrt=False)
    gpu.cuda.synchronize()
    assert np.array_equal(gi.cpu().numpy(), want_i)
    assert np.array_equal(gd.cpu().numpy().view(np.uint32), want_d.view(np.uint32))


def test_knn_select_batches_of_queries(gpu, oracle_mod):
    """k = 1000 over R = 16384: the workspace holds about 1984 queries, so
    Q = 4100 runs in three batches; sampled queries of each against the
    oracle."""
    import simd2_amd.knn as K
    torch = gpu
    R, Q, k = 16384, 4100, 1000
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    D = torch.randint(0, 100000, (R, Q), generator=g, device="cuda").float()
    say(f"select batched R={R} Q={Q} k={k}")
    gd, gi = K.knn_select(D, k, sqrt=False)
    torch.cuda.synchronize()
    cols = np.array([0, 1, 1983, 1984, 1985, 2500, 3967, 3968, 4099])
    sub = D[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    want_d, want_i = oracle_mod.knn_insertion_sort(sub, k, do_sqrt=False)
    assert np.array_equal(gi.cpu().numpy()[:, cols], want_i)
    assert np.array_equal(gd.cpu().numpy()[:, cols], want_d)


def test_knn_select_benchmark_size(gpu, oracle_mod):
    """The PLD workload's distance matrix (16384 x 16384 from d = 128
    clustered points, minus_square drop-in), k = 16: sampled queries against
    the oracle's insertion sort of the same columns."""
    import simd2_amd
    import simd2_amd.knn as K
    torch = gpu
    qry, ref = clustered(torch, 16384, 16384, 128, seed=100)
    D = simd2_amd.srgemm("minus_square", ref, qry.t().contiguous())   # R x Q
    say("select 16384 x 16384 k=16")
    gd, gi = K.knn_select(D, 16, sqrt=True)
    torch.cuda.synchronize()
    cols = np.linspace(0, 16383, 24).astype(np.int64)
    sub = D[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    want_d, want_i = oracle_mod.knn_insertion_sort(sub, 16, do_sqrt=True)
    assert np.array_equal(gi.cpu().numpy()[:, cols], want_i)
    assert np.array_equal(gd.cpu().numpy()[:, cols], want_d)


def test_knn_mxu_dropin_vs_oracle(oracle_mod):
    """knn_mxu with the reference's host layouts; integer coordinates make
    every squared distance exact, so indices and distances are bit-exact."""
    import simd2_amd.knn as K
    rng = np.random.default_rng(2)
    R, Q, dim = 500, 300, 40
    ref = rng.integers(0, 10, (R, dim)).astype(np.float32)        # ref_nb x dim
    query = rng.integers(0, 10, (dim, Q)).astype(np.float32)      # dim x query_nb
    D = oracle_mod.srgemm_f32("minus_square", ref, query)          # R x Q
    for k in (16, 64, 100, 500):
        say(f"knn_mxu k={k}")
        gd, gi, secs = K.knn_mxu(ref, query, k)
        want_d, want_i = oracle_mod.knn_insertion_sort(D, k, do_sqrt=True)
        assert np.array_equal(gi, want_i) and np.array_equal(gd, want_d)
        assert secs > 0


def fp64_topk_full(Qp, Rp, k):
    """Ground truth for small problems: every FP64 distance, ordered by
    (distance, index) -- exact on ties."""
    Q64, R64 = Qp.astype(np.float64), Rp.astype(np.float64)
    out_d = np.zeros((len(Q64), k))
    out_i = np.zeros((len(Q64), k), np.int64)
    for q in range(len(Q64)):
        dd = ((R64 - Q64[q]) ** 2).sum(1)
        order = np.lexsort((np.arange(len(R64)), dd))[:k]
        out_i[q], out_d[q] = order, dd[order]
    return out_d, out_i


@pytest.mark.parametrize("Q,R,d,k,data", [(2048, 4096, 128, 16, "clusters"),
                                          (300, 1000, 50, 5, "gauss"),
                                          (130, 70, 96, 16, "gauss"),
                                          (200, 2100, 64, 16, "clusters"),
                                          (100, 40000, 32, 8, "clusters"),
                                          (256, 10000, 128, 4, "clusters"),
                                          (200, 12000, 200, 8, "gauss"),
                                          (1000, 3000, 128, 31, "offset"),
                                          (600, 5000, 128, 64, "clusters"),
                                          (300, 3000, 64, 100, "gauss"),
                                          (200, 2000, 96, 256, "clusters"),
                                          (500, 2500, 300, 31, "clusters"),
                                          (512, 3000, 4096, 10, "uniform"),
                                          (256, 1000, 1000, 20, "offset")])
def test_knn_bf16_vs_fp64(gpu, Q, R, d, k, data):
    """Top-k agreement with FP64 and the distance error, over the register
    (d <= 160) and the streamed (d > 160) passes, each with 16-reference tile
    minima (R < 256 (k + 16)) and 64-reference ones; "offset" data sit 1000 away
    from the origin (the centring keeps the expansion's error at the data's
    own scale); "uniform" is the reference PLD test's data (coordinates
    uniform in [0, 10), apps/pld/cuASR/test.cpp) at d = 4096."""
    import simd2_amd.knn as K
    torch = gpu
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    if data in ("clusters", "offset"):
        qry, ref = clustered(torch, Q, R, d, seed=5)
        if data == "offset":
            qry, ref = qry + 1000.0, ref + 1000.0
    elif data == "uniform":
        qry = 10.0 * torch.rand((Q, d), generator=g, device="cuda")
        ref = 10.0 * torch.rand((R, d), generator=g, device="cuda")
    else:
        qry = torch.randn((Q, d), generator=g, device="cuda")
        ref = torch.randn((R, d), generator=g, device="cuda")
    say(f"knn_bf16 {data} Q={Q} R={R} d={d} k={k}")
    dist, idx, stats = K.knn_bf16(qry, ref, k)
    torch.cuda.synchronize()
    want_d, want_i = fp64_topk(qry.cpu().numpy(), ref.cpu().numpy(), k)
    got_i = idx.cpu().numpy()
    agree = np.mean(np.all(got_i == want_i, axis=1))
    got_d = dist.cpu().numpy().astype(np.float64)
    rel = np.max(np.abs(got_d - want_d) / np.maximum(want_d, 1e-30))
    say(f"  exact top-{k} agreement {agree:.5f}, max rel err {rel:.3e}, stats {stats.tolist()}")
    assert agree == 1.0
    assert rel <= 2e-6


def test_knn_bf16_duplicates_take_the_exact_scan(gpu):
    """3000 copies of one point among 1000 others: queries next to it have
    thousands of equidistant candidates, more than a candidate list holds;
    they are answered by the exact scan, ties in index order."""
    import simd2_amd.knn as K
    torch = gpu
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    d, k = 64, 12
    ref = torch.randn((4000, d), generator=g, device="cuda")
    ref[500:3500] = ref[500]
    qry = torch.cat([ref[500:501] + 0.01 * torch.randn((40, d), generator=g, device="cuda"),
                     torch.randn((60, d), generator=g, device="cuda")])
    say(f"knn_bf16 duplicates Q={qry.shape[0]} R={ref.shape[0]} k={k}")
    dist, idx, stats = K.knn_bf16(qry, ref, k)
    torch.cuda.synchronize()
    want_d, want_i = fp64_topk_full(qry.cpu().numpy(), ref.cpu().numpy(), k)
    assert np.array_equal(idx.cpu().numpy(), want_i)
    rel = np.max(np.abs(dist.cpu().numpy() - want_d) / np.maximum(want_d, 1e-30))
    assert rel <= 2e-6
    assert stats.tolist()[2] >= 40, stats.tolist()   # every query next to the copies


def test_knn_bf16_near_ties_use_fp64(gpu):
    """References at distances that differ in the 8th digit from the query:
    FP32 cannot order them, FP64 does; integer-spaced decoys elsewhere."""
    import simd2_amd.knn as K
    torch = gpu
    rng = np.random.default_rng(3)
    d, k = 40, 8
    q = rng.standard_normal(d)
    dirs = rng.standard_normal((20, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    radii = 3.0 + 1e-7 * rng.permutation(20)
    ref = np.concatenate([q + dirs * radii[:, None], q + 10 + rng.standard_normal((500, d))])
    qry = torch.from_numpy(q[None, :].astype(np.float32)).cuda()
    reft = torch.from_numpy(ref.astype(np.float32)).cuda()
    dist, idx, stats = K.knn_bf16(qry, reft, k)
    torch.cuda.synchronize()
    want_d, want_i = fp64_topk_full(qry.cpu().numpy(), reft.cpu().numpy(), k)
    assert np.array_equal(idx.cpu().numpy(), want_i)
    assert stats.tolist()[0] == 1 and stats.tolist()[1] >= k


def test_knn_direct_fp32_path(gpu):
    """The plain path: minus_square srgemm (FP32, VALU) into the Q x R matrix,
    then the reference top-k selection."""
    import simd2_amd
    import simd2_amd.knn as K
    torch = gpu
    Q, R, d, k = 1024, 2048, 128, 16
    qry, ref = clustered(torch, Q, R, d, seed=11)
    # rows of D = references: D (R x Q) = ref (R x d) (-)^2 query^T (d x Q)
    D = simd2_amd.srgemm("minus_square", ref, qry.t().contiguous())
    dist, idx = K.knn_select(D, k, sqrt=False)
    torch.cuda.synchronize()
    want_d, want_i = fp64_topk(qry.cpu().numpy(), ref.cpu().numpy(), k)
    got_i = idx.cpu().numpy().T
    agree = np.mean(np.all(got_i == want_i, axis=1))
    rel = np.max(np.abs(dist.cpu().numpy().T - want_d) / want_d)
    say(f"  direct FP32: agreement {agree:.5f}, max rel err {rel:.3e}")
    assert agree >= 0.99 and rel <= 1e-5


def test_tensor_srgemm_vs_restatement(gpu):
    import simd2_amd
    from oracle import fp16_emulation as emu
    rng = np.random.default_rng(4)
    for (m, n, k) in [(16, 16, 16), (48, 32, 64), (128, 96, 80)]:
        A = (0.1 + 2.4 * rng.random((m, k))).astype(np.float16)
        B = (0.1 + 2.4 * rng.random((k, n))).astype(np.float16)
        C = (0.1 + 2.4 * rng.random((m, n))).astype(np.float32)
        tA = gpu.from_numpy(A).cuda()
        tB = gpu.from_numpy(B).cuda()
        tC = gpu.from_numpy(C).cuda()
        for op in emu.OPS:
            for epi in (False, True):
                Ai, Bi = (A, B) if op != "or_and" else ((A > 1.3).astype(np.float16),
                                                        (B > 1.3).astype(np.float16))
                say(f"tensor_srgemm {op} {m}x{n}x{k} epi={epi}")
                D = simd2_amd.tensor_srgemm(op, gpu.from_numpy(Ai).cuda(),
                                            gpu.from_numpy(Bi).cuda(), tC, epi)
                gpu.cuda.synchronize()
                want = emu.tensor_srgemm(Ai, Bi, C, op, epi)
                assert np.array_equal(D.cpu().numpy().view(np.uint32),
                                      want.view(np.uint32)), (op, epi)
        del tA, tB
    # dims not multiple of 16 are refused as the reference refuses them
    lib = simd2_amd.lib()
    assert lib.simd2_tensor_srgemm(None, None, None, None, 20, 16, 16, 0, 0, None) == -1


def test_f2h_h2f_roundtrip(gpu):
    import simd2_amd
    x = gpu.randn(10007, device="cuda") * 1000
    h = simd2_amd.f2h(x)
    assert gpu.equal(h, x.half())
    assert gpu.equal(simd2_amd.h2f(h), x.half().float())
