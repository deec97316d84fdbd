"""GPU parity of the semiring GEMM entry points against the oracle.

Bit-exact for the six min/max semirings and or_and (and for minus_square on
integer-valued data); plus_mul (on the matrix cores for every shape) and
the fused minus_square on fractional data within 1e-5 relative of FP64
(tolerance stated per test)."""
import sys

import numpy as np
import pytest

from util import EXACT_OPS, OPS_F32, golden_cases, semiring_inputs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def run(torch, op, A, B, C, epi):
    import simd2_amd
    dt = torch.int32 if op == "or_and" else torch.float32
    tA = torch.from_numpy(np.ascontiguousarray(A)).to("cuda", dt)
    tB = torch.from_numpy(np.ascontiguousarray(B)).to("cuda", dt)
    tC = torch.from_numpy(np.ascontiguousarray(C)).to("cuda", dt)
    D = simd2_amd.srgemm(op, tA, tB, tC, epi)
    torch.cuda.synchronize()
    return D.cpu().numpy()


def fp64_ref(op, A, B, C, epi):
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    if op == "plus_mul":
        R = A64 @ B64
    else:
        R = (A64 ** 2).sum(1)[:, None] + (B64 ** 2).sum(0)[None, :] - 2 * A64 @ B64
    return R + C.astype(np.float64) if epi else R


def dot_err(D, A, B, C, epi):
    """max |D - FP64 product| / (|A| |B| [+ |C|]): the error of a float sum
    relative to the sum of its terms' magnitudes (order-independent scale)."""
    if D.size == 0:
        return 0.0
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = A64 @ B64
    scale = np.abs(A64) @ np.abs(B64)
    if epi:
        ref = ref + C.astype(np.float64)
        scale = scale + np.abs(C.astype(np.float64))
    return float(np.max(np.abs(D.astype(np.float64) - ref) / np.maximum(scale, 1e-30)))


def test_golden_cases_all_ops(gpu, golden_srgemm):
    """The reference test's inputs and the reference host_srgemm's outputs."""
    g = golden_srgemm
    for tag, M, N, K, epi in golden_cases(g):
        A, B, C = g[f"{tag}/A"], g[f"{tag}/B"], g[f"{tag}/C"]
        for op in OPS_F32:
            say(f"golden {tag} {op} M={M} N={N} K={K} epi={epi}")
            D = run(gpu, op, A, B, C, epi)
            want = g[f"{tag}/{op}"]
            if op == "minus_square":
                # fused (a-b)^2 + acc: one rounding fewer than the oracle
                np.testing.assert_allclose(D, want, rtol=1e-5, atol=1e-6)
            elif op == "plus_mul":
                # MFMA accumulation order: within 1e-5 of the sum of |terms|
                assert dot_err(D, A, B, C, epi) <= 1e-5, (tag, op)
            else:
                assert np.array_equal(bits(D), bits(want)), (tag, op)
        Ai, Bi, Ci = ((x.astype(np.int32) % 2).astype(np.int32) for x in (A, B, C))
        say(f"golden {tag} or_and")
        assert np.array_equal(run(gpu, "or_and", Ai, Bi, Ci, epi), g[f"{tag}/or_and"])


def test_reference_test_shapes(gpu, oracle_mod):
    """The 25 random shapes of tests/cuAsr_kernel_test.cu (glibc rand, dims
    < 256, some on and off the tile grid) with holes and integer data."""
    shapes = [(103, 198, 105), (205, 186, 171), (84, 248, 27), (99, 51, 159), (49, 88, 163),
              (233, 94, 212), (84, 17, 14), (135, 112, 233), (62, 1, 126), (56, 92, 42),
              (60, 84, 236), (67, 251, 250), (5, 60, 124), (249, 92, 187), (241, 179, 5),
              (229, 202, 11), (31, 35, 30), (115, 90, 197), (100, 36, 17), (242, 27, 16),
              (72, 71, 21), (155, 125, 245), (248, 41, 248), (232, 152, 50), (188, 95, 78)]
    for t, (M, N, K) in enumerate(shapes):
        for op in EXACT_OPS:
            epi = bool(t % 2)
            A, B, C = semiring_inputs(op, M, N, K, seed=100 + t)
            say(f"shape {t} {op} {M}x{N}x{K} epi={epi}")
            want = oracle_mod.srgemm_f32(op, A, B, C, epi, threads=8)
            assert np.array_equal(bits(run(gpu, op, A, B, C, epi)), bits(want)), (t, op)


@pytest.mark.parametrize("M,N,K", [(128, 128, 16), (256, 384, 512), (512, 512, 512),
                                   (1024, 256, 64)])
def test_tile_grid_exact_ops(gpu, oracle_mod, M, N, K):
    """Shapes on the 128x128x16 grid (fast path, both tile configurations)."""
    for op in EXACT_OPS:
        for epi in (False, True):
            A, B, C = semiring_inputs(op, M, N, K, seed=M + N + K)
            say(f"grid {op} {M}x{N}x{K} epi={epi}")
            want = oracle_mod.srgemm_f32(op, A, B, C, epi, threads=8)
            assert np.array_equal(bits(run(gpu, op, A, B, C, epi)), bits(want)), op


@pytest.mark.parametrize("M,N,K", [(256, 256, 256), (512, 384, 1024), (200, 300, 100)])
def test_plus_mul_and_minus_square_tolerance(gpu, M, N, K):
    """plus_mul (MFMA) and minus_square (fused FMA)
    within 1e-5 relative of FP64; inputs U{1..100} without holes."""
    for op in ("plus_mul", "minus_square"):
        for epi in (False, True):
            A, B, C = semiring_inputs(op, M, N, K, seed=7)
            say(f"tol {op} {M}x{N}x{K} epi={epi}")
            D = run(gpu, op, A, B, C, epi).astype(np.float64)
            ref = fp64_ref(op, A, B, C, epi)
            rel = np.max(np.abs(D - ref) / np.abs(ref))
            assert rel <= 1e-5, (op, rel)


def test_minus_square_bit_exact_on_integer_data(gpu, oracle_mod):
    """On integer-valued inputs (the 4096^3 benchmark's U{1..100}) every
    (a - b)^2 is an exact integer, so the fused acc + d*d rounds exactly as
    the oracle's separate product and sum; with the same k order the result
    is the oracle's bit for bit.  K = 16384 takes the sums past 2^24, where
    every addition rounds."""
    rng = np.random.default_rng(13)
    M, N, K = 64, 96, 16384
    A = rng.integers(1, 101, (M, K)).astype(np.float32)
    B = rng.integers(1, 101, (K, N)).astype(np.float32)
    C = rng.integers(1, 101, (M, N)).astype(np.float32)
    for epi in (False, True):
        say(f"minus_square integer {M}x{N}x{K} epi={epi}")
        D = run(gpu, "minus_square", A, B, C, epi)
        want = oracle_mod.srgemm_f32("minus_square", A, B, C, epi, threads=8)
        assert float(want.max()) > 2.0 ** 24
        assert np.array_equal(bits(D), bits(want))


def test_plus_mul_mfma_fractional(gpu):
    """MFMA layout check with non-integer data (an index or transpose slip in
    the operand layout cannot pass)."""
    rng = np.random.default_rng(9)
    A = rng.standard_normal((256, 128)).astype(np.float32)
    B = rng.standard_normal((128, 384)).astype(np.float32)
    C = rng.standard_normal((256, 384)).astype(np.float32)
    D = run(gpu, "plus_mul", A, B, C, True).astype(np.float64)
    ref = A.astype(np.float64) @ B.astype(np.float64) + C
    assert np.max(np.abs(D - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.parametrize("M,N,K,pad", [(1000, 777, 300, 0), (130, 70, 65, 3), (1, 5, 3, 1),
                                       (257, 129, 1, 0), (64, 300, 17, 2), (128, 128, 16, 1)])
def test_plus_mul_mfma_every_shape(gpu, M, N, K, pad):
    """plus_mul runs on the matrix cores for every shape: off the 128x128x16
    grid, K not a multiple of 4, and leading dimensions padded by `pad`
    floats (unaligned rows take the element-wise loads)."""
    import simd2_amd
    torch = gpu
    rng = np.random.default_rng(M + N + K + pad)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = rng.standard_normal((M, N)).astype(np.float32)
    Ap = np.zeros((M, K + pad), np.float32)
    Bp = np.zeros((K, N + pad), np.float32)
    Cp = np.zeros((M, N + pad), np.float32)
    Ap[:, :K], Bp[:, :N], Cp[:, :N] = A, B, C
    tA, tB, tC = (torch.from_numpy(x).cuda() for x in (Ap, Bp, Cp))
    tD = torch.full((M, N + pad), 7.0, device="cuda")
    for epi in (False, True):
        say(f"plus_mul mfma {M}x{N}x{K} pad={pad} epi={epi}")
        # row-major A @ B as the column-major call (N, M, K, B, ldb, A, lda, C, ldc, D)
        st = simd2_amd.cuasr_srsgemm("plus_mul", N, M, K, tB, N + pad, tA, K + pad, tC, N + pad,
                                     tD, epi)
        assert st == 0
        torch.cuda.synchronize()
        D = tD.cpu().numpy()
        assert dot_err(D[:, :N], A, B, C, epi) <= 1e-5
        assert np.all(D[:, N:] == 7.0)  # padding untouched


def test_plus_mul_tile_sizes_identical(gpu):
    """The MFMA kernel's 64x64 tiles (problems with fewer 128x128 tiles than
    CUs) and 128x128 tiles sum every element over the same k order: rows of a
    2048 x 2048 product (128x128 tiles) equal the same rows computed as a
    256-row problem (64x64 tiles), bit for bit."""
    import simd2_amd
    torch = gpu
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    A = torch.randn((2048, 1000), generator=g, device="cuda")
    B = torch.randn((1000, 2048), generator=g, device="cuda")
    C = torch.randn((2048, 2048), generator=g, device="cuda")
    say("plus_mul 2048^2 x 1000 in both tile sizes")
    full = simd2_amd.srgemm("plus_mul", A, B, C, True)
    part = simd2_amd.srgemm("plus_mul", A[512:768].contiguous(), B, C[512:768].contiguous(), True)
    torch.cuda.synchronize()
    assert torch.equal(full[512:768].view(torch.int32), part.view(torch.int32))


def test_large_rows_sampled(gpu, oracle_mod):
    """min_plus and max_min at 2048^3: rows sampled against the oracle."""
    import torch
    import simd2_amd
    M = N = K = 2048
    for op in ("min_plus", "max_min"):
        A, B, C = semiring_inputs(op, M, N, K, seed=21)
        say(f"large {op} {M}^3")
        D = run(gpu, op, A, B, C, True)
        rows = [0, 1, 127, 128, 1000, 2047]
        for r in rows:
            want = oracle_mod.srgemm_f32(op, A[r:r + 1], B, C[r:r + 1], True)
            assert np.array_equal(bits(D[r:r + 1]), bits(want)), (op, r)
    del torch, simd2_amd


def test_dropin_column_major_strided(gpu, oracle_mod):
    """Column-major call with leading dimensions larger than the matrices,
    exactly as kernel/srgemm.cuh takes them."""
    import torch
    import simd2_amd
    M, N, K = 70, 50, 90
    lda, ldb, ldc = 80, 100, 72
    rng = np.random.default_rng(5)
    Acm = rng.integers(1, 100, (K, lda)).astype(np.float32)   # column k at row k
    Bcm = rng.integers(1, 100, (N, ldb)).astype(np.float32)
    Ccm = rng.integers(1, 100, (N, ldc)).astype(np.float32)
    tA, tB, tC = (torch.from_numpy(x).cuda() for x in (Acm, Bcm, Ccm))
    tD = torch.full((N, ldc), -7.0, device="cuda")
    st = simd2_amd.cuasr_minplus_srsgemm(M, N, K, tA, lda, tB, ldb, tC, ldc, tD, True)
    assert st == 0
    torch.cuda.synchronize()
    A = Acm[:, :M].T  # A(i,k)
    B = Bcm[:, :K].T  # B(k,j)
    C = Ccm[:, :M].T
    want = oracle_mod.srgemm_f32("min_plus", A, B, C, True)
    got = tD.cpu().numpy()
    assert np.array_equal(got[:, :M].T, want)
    assert np.all(got[:, M:] == -7.0)  # padding between columns untouched


def test_c_equals_d(gpu, oracle_mod):
    import torch
    import simd2_amd
    A, B, C = semiring_inputs("min_plus", 128, 128, 64, seed=3)
    tA, tB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    tC = torch.from_numpy(C).cuda()
    simd2_amd.srgemm("min_plus", tA, tB, tC, True, out=tC)
    torch.cuda.synchronize()
    assert np.array_equal(tC.cpu().numpy(), oracle_mod.srgemm_f32("min_plus", A, B, C, True))


def test_or_and_int_and_packed(gpu, oracle_mod):
    import torch
    import simd2_amd
    for (M, N, K, p) in [(300, 200, 500, 0.02), (64, 2048, 64, 0.1), (1, 1, 1, 1.0),
                         (257, 95, 33, 0.3)]:
        rng = np.random.default_rng(M + N)
        A = (rng.random((M, K)) < p).astype(np.int32)
        B = (rng.random((K, N)) < p).astype(np.int32)
        C = (rng.random((M, N)) < p).astype(np.int32)
        for epi in (False, True):
            say(f"or_and {M}x{N}x{K} p={p} epi={epi}")
            want = oracle_mod.srgemm_i32(A, B, C, epi)
            assert np.array_equal(run(gpu, "or_and", A, B, C, epi), want)
            pa, pb, pc = (simd2_amd.pack_bits(torch.from_numpy(x).cuda()) for x in (A, B, C))
            D = simd2_amd.orand_packed(pa, pb, N, pc, epi)
            torch.cuda.synchronize()
            assert np.array_equal(simd2_amd.unpack_bits(D, N).cpu().numpy(), want)


def test_k_zero_and_empty(gpu):
    import torch
    import simd2_amd
    A = torch.empty((5, 0), device="cuda")
    B = torch.empty((0, 6), device="cuda")
    C = torch.arange(30, dtype=torch.float32, device="cuda").view(5, 6)
    D = simd2_amd.srgemm("min_plus", A, B, C, True)
    assert torch.equal(D, C)
    D = simd2_amd.srgemm("max_plus", A, B, C, False)
    assert torch.all(D == float("-inf"))
    D = simd2_amd.srgemm("plus_mul", A, B, C, False)
    assert torch.all(D == 0)


@pytest.mark.parametrize("M,N,K", [(256, 384, 128), (200, 130, 70)])
def test_every_tile_config_identical(gpu, oracle_mod, M, N, K):
    """Each VALU tile configuration (128x128, 64x64, 128x64 tiles) returns the
    oracle's bits, on and off the tile grid."""
    import simd2_amd
    n_cfg = simd2_amd.lib().simd2_tile_config_count()
    assert n_cfg == 3
    for op in ("min_plus", "max_min", "min_mul"):
        A, B, C = semiring_inputs(op, M, N, K, seed=M * 7 + K)
        want = oracle_mod.srgemm_f32(op, A, B, C, True, threads=8)
        tA, tB, tC = (gpu.from_numpy(x).cuda() for x in (A, B, C))
        for cfg in range(n_cfg):
            say(f"cfg {cfg} {op} {M}x{N}x{K}")
            D = simd2_amd.srgemm(op, tA, tB, tC, True, tile_config=cfg)
            gpu.cuda.synchronize()
            assert np.array_equal(bits(D.cpu().numpy()), bits(want)), (op, cfg)


@pytest.mark.parametrize("M,N,K", [(200, 132, 68), (129, 68, 20), (64, 64, 20), (300, 64, 4),
                                   (1000, 772, 300)])
def test_split_form_off_grid(gpu, oracle_mod, M, N, K):
    """Off-grid shapes with 16-byte-aligned rows (K and N multiples of 4): the
    tiles inside the matrix take the 128-bit path with a checked last K slab,
    the edge tiles the checked path.  Every op, every tile configuration,
    with and without C, and C and D wider than N (leading dimension N + 4,
    D's padding untouched)."""
    import simd2_amd
    for op in EXACT_OPS:
        A, B, C = semiring_inputs(op, M, N, K, seed=M + 3 * N + 7 * K)
        tA, tB, tC = (gpu.from_numpy(x).cuda() for x in (A, B, C))
        for epi in (False, True):
            want = bits(oracle_mod.srgemm_f32(op, A, B, C, epi, threads=8))
            for cfg in range(simd2_amd.lib().simd2_tile_config_count()):
                say(f"split {op} {M}x{N}x{K} epi={epi} cfg={cfg}")
                D = simd2_amd.srgemm(op, tA, tB, tC, epi, tile_config=cfg)
                gpu.cuda.synchronize()
                assert np.array_equal(bits(D.cpu().numpy()), want), (op, epi, cfg)
        # the drop-in's C and D share one leading dimension (kernel/srgemm.cuh)
        Cw = gpu.full((M, N + 4), -3.0, device="cuda")
        Cw[:, :N] = tC
        Dw = gpu.full((M, N + 4), -5.0, device="cuda")
        assert simd2_amd.cuasr_srsgemm(op, N, M, K, tB, N, tA, K, Cw, N + 4, Dw, True) == 0
        gpu.cuda.synchronize()
        got = Dw.cpu().numpy()
        assert np.array_equal(bits(got[:, :N]), bits(oracle_mod.srgemm_f32(op, A, B, C, True,
                                                                          threads=8))), op
        assert np.all(got[:, N:] == -5.0), op


def test_autotune_records_winner(gpu, oracle_mod):
    import simd2_amd
    M = N = K = 512
    A, B, C = semiring_inputs("min_plus", M, N, K, seed=17)
    tA, tB, tC = (gpu.from_numpy(x).cuda() for x in (A, B, C))
    D, best, ms = simd2_amd.autotune("min_plus", tA, tB, tC, True)
    gpu.cuda.synchronize()
    assert best in (0, 1, 2) and all(t > 0 for t in ms)
    assert ms[best] == min(ms)
    want = oracle_mod.srgemm_f32("min_plus", A, B, C, True, threads=8)
    assert np.array_equal(bits(D.cpu().numpy()), bits(want))
    assert simd2_amd.lib().simd2_pick_tile_config(0, N, M, K) == best


@pytest.mark.parametrize("M,N,K", [(1024, 1024, 1024), (512, 768, 64), (256, 256, 4096),
                                   (1536, 1536, 256), (2048, 2048, 64), (1000, 777, 300),
                                   (1000, 776, 300), (130, 70, 65), (33, 50, 17)])
def test_small_problems_split_k(gpu, oracle_mod, M, N, K):
    """Problems with at most four 64x64 tiles per CU take the split-K form
    (four groups over contiguous slab ranges up to two tiles per CU, two
    groups up to four: 1536^2 and 2048^2; partials combined in order) for
    the six min/max semirings -- the checked form when rows are not 16-byte
    aligned (N = 777, 70, 50) or the shape is off the tile grid (1000 x 776):
    edge tiles, K past the last whole slab, group ranges past K --: the oracle's bits with and without C, and the same bits
    as the unsplit 64x64 kernel (tile_config 1) on data with signed zeros,
    whose ties only an order-independent (+) leaves alone."""
    import simd2_amd
    for op in EXACT_OPS:
        for epi in (False, True):
            A, B, C = semiring_inputs(op, M, N, K, seed=M + 3 * N + K)
            say(f"split-K {op} {M}x{N}x{K} epi={epi}")
            want = oracle_mod.srgemm_f32(op, A, B, C, epi, threads=8)
            assert np.array_equal(bits(run(gpu, op, A, B, C, epi)), bits(want)), (op, epi)
        rng = np.random.default_rng(M + K)
        A, B, C = semiring_inputs(op, M, N, K, seed=5)
        for X in (A, B, C):
            z = rng.random(X.shape) < 0.3
            X[z] = np.where(rng.random(int(z.sum())) < 0.5, np.float32(0.0), np.float32(-0.0))
        tA, tB, tC = (gpu.from_numpy(x).cuda() for x in (A, B, C))
        say(f"split-K signed zeros {op} {M}x{N}x{K}")
        Dsplit = simd2_amd.srgemm(op, tA, tB, tC, True)
        Dseq = simd2_amd.srgemm(op, tA, tB, tC, True, tile_config=1)
        gpu.cuda.synchronize()
        assert np.array_equal(bits(Dsplit.cpu().numpy()), bits(Dseq.cpu().numpy())), op
