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


def test_knn_select_vs_oracle(gpu, oracle_mod):
    import simd2_amd.knn as K
    rng = np.random.default_rng(1)
    for (R, Q, k) in [(300, 70, 16), (1000, 130, 32), (64, 64, 1), (257, 5, 7)]:
        D = rng.integers(0, 50, (R, Q)).astype(np.float32)  # ties on purpose
        say(f"select R={R} Q={Q} k={k}")
        want_d, want_i = oracle_mod.knn_insertion_sort(D, k, do_sqrt=True)
        gd, gi = K.knn_select(gpu.from_numpy(D).cuda(), k, sqrt=True)
        gpu.cuda.synchronize()
        assert np.array_equal(gi.cpu().numpy(), want_i)
        assert np.array_equal(gd.cpu().numpy(), want_d)


def test_knn_mxu_dropin_vs_oracle(oracle_mod):
    """knn_mxu with the reference's host layouts; integer coordinates make
    every squared distance exact, so indices and distances are bit-exact."""
    import simd2_amd.knn as K
    rng = np.random.default_rng(2)
    R, Q, dim, k = 500, 300, 40, 16
    ref = rng.integers(0, 10, (R, dim)).astype(np.float32)        # ref_nb x dim
    query = rng.integers(0, 10, (dim, Q)).astype(np.float32)      # dim x query_nb
    say("knn_mxu")
    gd, gi, secs = K.knn_mxu(ref, query, k)
    D = oracle_mod.srgemm_f32("minus_square", ref, query)          # R x Q
    want_d, want_i = oracle_mod.knn_insertion_sort(D, k, do_sqrt=True)
    assert np.array_equal(gi, want_i) and np.array_equal(gd, want_d)
    assert secs > 0


@pytest.mark.parametrize("Q,R,d,k,clusters", [(2048, 4096, 128, 16, True),
                                              (300, 1000, 50, 5, False),
                                              (130, 70, 96, 16, False),
                                              (200, 2100, 64, 16, True),
                                              (100, 40000, 32, 8, True)])
def test_knn_bf16_vs_fp64(gpu, Q, R, d, k, clusters):
    import simd2_amd.knn as K
    torch = gpu
    if clusters:
        qry, ref = clustered(torch, Q, R, d, seed=5)
    else:
        g = torch.Generator(device="cuda")
        g.manual_seed(7)
        qry = torch.randn((Q, d), generator=g, device="cuda")
        ref = torch.randn((R, d), generator=g, device="cuda")
    say(f"knn_bf16 Q={Q} R={R} d={d} k={k}")
    dist, idx, stats = K.knn_bf16(qry, ref, k)
    torch.cuda.synchronize()
    want_d, want_i = fp64_topk(qry.cpu().numpy(), ref.cpu().numpy(), k)
    got_i = idx.cpu().numpy()
    agree = np.mean(np.all(got_i == want_i, axis=1))
    got_d = dist.cpu().numpy().astype(np.float64)
    rel = np.max(np.abs(got_d - want_d) / np.maximum(want_d, 1e-30))
    say(f"  exact top-{k} agreement {agree:.5f}, max rel err {rel:.3e}, fp64 {stats.tolist()}")
    assert agree == 1.0
    assert rel <= 2e-6


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
