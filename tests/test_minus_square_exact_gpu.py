"""simd2_minus_square_exact (minus_square on the int8 matrix cores) against
the oracle, bit for bit.

Tiles whose operands are integers in [-128, 127] and whose sums stay at or
below 2^24 run on v_mfma_i32_16x16x64_i8 as |a|^2 + |b|^2 - 2 a.b; there the
oracle's k-by-k FP32 sum is the same integer.  Every other tile runs the
direct FP32 kernel (the one cuasr_plusmiussquare_srsgemm runs).  The tests
check the result against the oracle (or the direct kernel where the oracle
would be slow) and, through simd2_minus_square_tiles, that each tile took the
path its data calls for: fractional or out-of-range entries, infinities, and
sums past 2^24 (where the oracle's additions round, so a tile written from
the exact integer would differ); and the caller-owned workspace's checks."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def tiles(M, N):
    return ((M + 63) // 64) * ((N + 63) // 64)


def run(torch, A, B, C, epi, direct=False, stream=None, workspace=None):
    """D and the (int8, direct) tile counts of the exact entry point, or D of
    the drop-in (the direct kernel) when `direct`."""
    import simd2_amd
    tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(x)).to("cuda") for x in (A, B, C))
    out = torch.empty((A.shape[0], B.shape[1]), device="cuda")
    if stream is not None:
        stream.wait_stream(torch.cuda.current_stream())
    counts = None
    if direct:
        D = simd2_amd.srgemm("minus_square", tA, tB, tC, epi, out=out, stream=stream)
    else:
        D, ws = simd2_amd.minus_square_exact(tA, tB, tC, epi, out=out, workspace=workspace,
                                             stream=stream)
        M, K = A.shape
        counts = simd2_amd.minus_square_tiles(ws, M, B.shape[1], K, stream)
    torch.cuda.synchronize()
    return D.cpu().numpy(), counts


def ints(rng, shape, lo, hi):
    return rng.integers(lo, hi + 1, shape).astype(np.float32)


@pytest.mark.parametrize("M,N,K", [(128, 128, 64), (256, 384, 512), (1000, 777, 300),
                                   (64, 64, 1), (1, 200, 70), (129, 65, 63), (640, 512, 4096)])
def test_integer_data_runs_on_matrix_cores(gpu, oracle_mod, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A, B, C = ints(rng, (M, K), 1, 100), ints(rng, (K, N), 1, 100), ints(rng, (M, N), 1, 100)
    for epi in (False, True):
        say(f"int8 path U{{1..100}} {M}x{N}x{K} epi={epi}")
        D, (mf, va) = run(gpu, A, B, C, epi)
        want = oracle_mod.srgemm_f32("minus_square", A, B, C, epi, threads=8)
        assert np.array_equal(bits(D), bits(want)), (M, N, K, epi)
        assert (mf, va) == (tiles(M, N), 0), (mf, va)


def test_full_int8_range_and_signed_zero(gpu, oracle_mod):
    """Entries over all of [-128, 127] (differences up to 255), -0.0 among
    them, K small enough that every sum stays below 2^24."""
    rng = np.random.default_rng(5)
    M, N, K = 300, 260, 200
    A, B = ints(rng, (M, K), -128, 127), ints(rng, (K, N), -128, 127)
    A[rng.random((M, K)) < 0.05] = -0.0
    A[0, :] = -128.0
    B[:, 0] = 127.0
    C = rng.standard_normal((M, N)).astype(np.float32)
    for epi in (False, True):
        say(f"int8 path full range epi={epi}")
        D, (mf, va) = run(gpu, A, B, C, epi)
        want = oracle_mod.srgemm_f32("minus_square", A, B, C, epi, threads=8)
        assert np.array_equal(bits(D), bits(want))
        assert (mf, va) == (tiles(M, N), 0)


def test_rows_and_columns_that_fail_the_check(gpu, oracle_mod):
    """A fractional entry in one row of A, an infinity in one column of B, a
    128 (just out of range) in another row: exactly the tiles touching those
    rows and columns go to the direct kernel."""
    rng = np.random.default_rng(11)
    M, N, K = 512, 448, 256
    A, B, C = ints(rng, (M, K), 1, 100), ints(rng, (K, N), 1, 100), ints(rng, (M, N), 1, 100)
    A[70, 33] = 2.5          # tile row 1
    A[400, 255] = 128.0      # tile row 6
    B[17, 300] = np.inf      # tile column 4
    for epi in (False, True):
        say(f"int8 path with failing rows/columns epi={epi}")
        D, (mf, va) = run(gpu, A, B, C, epi)
        want = oracle_mod.srgemm_f32("minus_square", A, B, C, epi, threads=8)
        assert np.array_equal(bits(D), bits(want))
        tr, tc = M // 64, N // 64
        bad = 2 * tc + tr - 2  # two tile rows, one tile column, counted once
        assert (mf, va) == (tr * tc - bad, bad), (mf, va)


def test_sums_past_two_to_the_24(gpu, oracle_mod):
    """U{1..100} with K = 16384: sums near 2.7e7, where the oracle's FP32
    additions round; every tile must take the direct kernel.  Then a matrix
    where only some columns are large: only their tiles do."""
    rng = np.random.default_rng(13)
    M, N, K = 64, 96, 16384
    A, B, C = ints(rng, (M, K), 1, 100), ints(rng, (K, N), 1, 100), ints(rng, (M, N), 1, 100)
    say("int8 path past 2^24, all tiles")
    D, (mf, va) = run(gpu, A, B, C, True)
    want = oracle_mod.srgemm_f32("minus_square", A, B, C, True, threads=8)
    assert float(want.max()) > 2.0 ** 24
    assert np.array_equal(bits(D), bits(want))
    assert (mf, va) == (0, tiles(M, N))
    M, N, K = 128, 256, 2048
    A = ints(rng, (M, K), -10, 10)
    B = ints(rng, (K, N), -10, 10)
    B[:, 192:] = ints(rng, (K, 64), 100, 127) * np.where(rng.random((K, 64)) < 0.5, 1, -1)
    C = ints(rng, (M, N), 1, 100)
    say("int8 path past 2^24 in the last tile column")
    D, (mf, va) = run(gpu, A, B, C, True)
    want = oracle_mod.srgemm_f32("minus_square", A, B, C, True, threads=8)
    assert float(want[:, 192:].min()) > 2.0 ** 24 and float(want[:, :192].max()) < 2.0 ** 24
    assert np.array_equal(bits(D), bits(want))
    assert (mf, va) == (6, 2), (mf, va)


def test_fractional_data_takes_the_direct_kernel(gpu):
    """Gaussian data (the kNN workload's kind): every tile on the direct
    kernel, identical to the explicit direct configuration."""
    torch = gpu
    rng = np.random.default_rng(3)
    A = rng.standard_normal((384, 128)).astype(np.float32)
    B = rng.standard_normal((128, 320)).astype(np.float32)
    C = rng.standard_normal((384, 320)).astype(np.float32)
    D, (mf, va) = run(torch, A, B, C, True)
    E, _ = run(torch, A, B, C, True, direct=True)
    assert np.array_equal(bits(D), bits(E))
    assert (mf, va) == (0, tiles(384, 320))


def test_large_square_matches_direct_kernel(gpu):
    """2048^3 U{1..100} (the benchmark's data at half the size): the int8
    path against the drop-in's direct FP32 kernel, bit for bit, C aliased to
    D."""
    torch = gpu
    import simd2_amd
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    n = 2048
    A = torch.randint(1, 101, (n, n), device="cuda", generator=g).float()
    B = torch.randint(1, 101, (n, n), device="cuda", generator=g).float()
    C = torch.randint(1, 101, (n, n), device="cuda", generator=g).float()
    say("int8 path 2048^3 vs direct kernel")
    D1 = C.clone()
    _, ws = simd2_amd.minus_square_exact(A, B, D1, True, out=D1)
    mf, va = simd2_amd.minus_square_tiles(ws, n, n, n)
    D2 = simd2_amd.srgemm("minus_square", A, B, C, True)
    torch.cuda.synchronize()
    assert torch.equal(D1.view(torch.int32), D2.view(torch.int32))
    assert (mf, va) == (tiles(n, n), 0)


def test_k_past_the_int8_limit_takes_the_direct_kernel(gpu, oracle_mod):
    """K = 70000 (past the 65536 the int8 path accepts): no workspace is
    needed, the result is the oracle's and every tile is reported direct."""
    import simd2_amd
    rng = np.random.default_rng(23)
    M, N, K = 3, 70, 70000
    assert simd2_amd.minus_square_workspace(M, N, K) == 0
    A, B, C = ints(rng, (M, K), 0, 1), ints(rng, (K, N), 0, 1), ints(rng, (M, N), 0, 1)
    say("K = 70000, direct kernel")
    D, (mf, va) = run(gpu, A, B, C, True)
    want = oracle_mod.srgemm_f32("minus_square", A, B, C, True, threads=8)
    assert np.array_equal(bits(D), bits(want))
    assert (mf, va) == (0, tiles(M, N)), (mf, va)


def test_workspace_checks_and_reuse(gpu, oracle_mod):
    """A missing or short workspace is refused with SIMD2_ERROR_WORKSPACE and
    nothing is written; one workspace sized for the largest problem serves
    smaller ones, on the default stream and on a second stream in turn."""
    torch = gpu
    import simd2_amd
    rng = np.random.default_rng(17)
    M, N, K = 300, 200, 500
    need = simd2_amd.minus_square_workspace(M, N, K)
    assert need >= (M + N) * K
    A, B, C = ints(rng, (M, K), 1, 100), ints(rng, (K, N), 1, 100), ints(rng, (M, N), 1, 100)
    tA, tB, tC = (torch.from_numpy(x).cuda() for x in (A, B, C))
    out = torch.full((M, N), 7.0, device="cuda")
    short = torch.empty(need - 1, dtype=torch.uint8, device="cuda")
    for ws in (short,):
        with pytest.raises(simd2_amd.Simd2Error) as err:
            simd2_amd.minus_square_exact(tA, tB, tC, True, out=out, workspace=ws)
        assert err.value.status == simd2_amd.ERROR_WORKSPACE
    st = simd2_amd.lib().simd2_minus_square_exact(N, M, K, tB.data_ptr(), N, tA.data_ptr(), K,
                                                  tC.data_ptr(), N, out.data_ptr(), N, 1, None,
                                                  need, torch.cuda.current_stream().cuda_stream)
    assert st == simd2_amd.ERROR_WORKSPACE
    # a workspace inside the output buffer
    big_out = torch.full((M + need // (4 * N) + 2, N), 7.0, device="cuda")
    st = simd2_amd.lib().simd2_minus_square_exact(N, M, K, tB.data_ptr(), N, tA.data_ptr(), K,
                                                  tC.data_ptr(), N, big_out.data_ptr(), N, 1,
                                                  big_out.data_ptr() + 4 * N * 2, need,
                                                  torch.cuda.current_stream().cuda_stream)
    assert st == simd2_amd.ERROR_ALIASING
    # a workspace not 256-byte aligned
    spare = torch.empty(need + 256, dtype=torch.uint8, device="cuda")
    st = simd2_amd.lib().simd2_minus_square_exact(N, M, K, tB.data_ptr(), N, tA.data_ptr(), K,
                                                  tC.data_ptr(), N, out.data_ptr(), N, 1,
                                                  spare.data_ptr() + 16, need,
                                                  torch.cuda.current_stream().cuda_stream)
    assert st == simd2_amd.ERROR_MISALIGNED_OPERAND
    torch.cuda.synchronize()
    assert bool((out == 7.0).all()) and bool((big_out == 7.0).all())
    big = torch.empty(need, dtype=torch.uint8, device="cuda")
    s2 = torch.cuda.Stream()
    for stream in (None, s2):
        for m, n, k in ((M, N, K), (100, 90, 80), (64, 64, 64)):
            A, B, C = ints(rng, (m, k), 1, 100), ints(rng, (k, n), 1, 100), ints(rng, (m, n), 1, 100)
            say(f"workspace reuse {m}x{n}x{k} stream={'s2' if stream else 'default'}")
            D, (mf, va) = run(torch, A, B, C, True, stream=stream, workspace=big)
            want = oracle_mod.srgemm_f32("minus_square", A, B, C, True, threads=8)
            assert np.array_equal(bits(D), bits(want))
            assert (mf, va) == (tiles(m, n), 0)


def test_captured_into_a_graph(gpu):
    """The entry point allocates nothing, so it can be captured: three replays
    of a captured call, with eager work between them, each equal to the eager
    result (inputs changed in place between replays)."""
    torch = gpu
    import simd2_amd
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    M, N, K = 384, 320, 700
    A = torch.randint(1, 101, (M, K), device="cuda", generator=g).float()
    B = torch.randint(1, 101, (K, N), device="cuda", generator=g).float()
    C = torch.randint(1, 101, (M, N), device="cuda", generator=g).float()
    D = torch.empty((M, N), device="cuda")
    ws = torch.empty(simd2_amd.minus_square_workspace(M, N, K), dtype=torch.uint8, device="cuda")
    simd2_amd.minus_square_exact(A, B, C, True, out=D, workspace=ws)  # warm-up
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        simd2_amd.minus_square_exact(A, B, C, True, out=D, workspace=ws)
    for rep in range(3):
        A.copy_(torch.randint(1, 101, (M, K), device="cuda", generator=g).float())
        C.mul_(1.0).add_(rep)
        torch.cuda.synchronize()
        say(f"graph replay {rep}")
        graph.replay()
        torch.cuda.synchronize()
        got = D.clone()
        want = simd2_amd.srgemm("minus_square", A, B, C, True)
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32)), rep
        assert simd2_amd.minus_square_tiles(ws, M, N, K) == (tiles(M, N), 0)
