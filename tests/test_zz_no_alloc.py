"""simd2_minus_square_exact_noalloc -- minus_square on the int8 matrix cores
with the library's own scratch instead of a workspace argument -- against
simd2_minus_square_exact (the caller-workspace entry point), bit for bit,
through the C ABI.

Two calls in a row on one device share the scratch; the second is smaller,
so a stale byte of the first call's quantized operands, norms or tile gates
that the second call read would show as a difference.  Each call mixes
tiles the int8 path writes with tiles it leaves to the direct kernel
(fractional entries in chosen rows), and the second call also runs with the
epilogue.  A problem whose workspace exceeds the scratch is refused."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WORKSPACE_ERROR = 15
MAX_BYTES = 40 << 20


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bind():
    import simd2_amd
    L = simd2_amd.lib()
    i, p = ctypes.c_int, ctypes.c_void_p
    L.simd2_minus_square_exact_noalloc.argtypes = [i, i, i, p, i, p, i, p, i, p, i, i, p]
    L.simd2_minus_square_exact_noalloc.restype = i
    L.simd2_minus_square_exact.argtypes = [i, i, i, p, i, p, i, p, i, p, i, i, p,
                                           ctypes.c_size_t, p]
    L.simd2_minus_square_exact.restype = i
    L.simd2_minus_square_workspace.argtypes = [i, i, i]
    L.simd2_minus_square_workspace.restype = ctypes.c_size_t
    return L


def case(rng, M, N, K):
    """Row-major A (M x K), B (K x N), C (M x N): small integers, with
    fractions in a few rows of A so some tiles take the direct kernel."""
    A = rng.integers(-20, 21, (M, K)).astype(np.float32)
    B = rng.integers(-20, 21, (K, N)).astype(np.float32)
    A[rng.choice(M, 3, replace=False), :] += 0.25
    C = rng.integers(0, 50, (M, N)).astype(np.float32)
    return A, B, C


def call(torch, L, A, B, C, epi, noalloc):
    """Row-major D = [C +] A (-)^2 B as the column-major call (N, M, K, B, A)."""
    M, K = A.shape
    N = B.shape[1]
    tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (A, B, C))
    D = torch.full((M, N), float("nan"), device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        st = L.simd2_minus_square_exact_noalloc(N, M, K, tB.data_ptr(), N, tA.data_ptr(), K,
                                                tC.data_ptr(), N, D.data_ptr(), N, int(epi), s)
    else:
        need = L.simd2_minus_square_workspace(N, M, K)
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
        st = L.simd2_minus_square_exact(N, M, K, tB.data_ptr(), N, tA.data_ptr(), K,
                                        tC.data_ptr(), N, D.data_ptr(), N, int(epi),
                                        ws.data_ptr(), need, s)
    assert st == 0, st
    torch.cuda.synchronize()
    return D.cpu().numpy().view(np.uint32)


def test_noalloc_matches_workspace_entry_point(gpu):
    L = bind()
    rng = np.random.default_rng(11)
    # First call: 1000 x 900 x 700; second, smaller, with the epilogue.
    for (M, N, K), epi in (((1000, 900, 700), False), ((300, 200, 130), True)):
        say(f"noalloc vs workspace minus_square {M}x{N}x{K} epi={epi}")
        A, B, C = case(rng, M, N, K)
        want = call(gpu, L, A, B, C, epi, noalloc=False)
        got = call(gpu, L, A, B, C, epi, noalloc=True)
        assert np.array_equal(got, want), (M, N, K, epi)


def test_noalloc_repeats_on_the_same_scratch(gpu):
    """The same small problem twice after a large one: equal results."""
    L = bind()
    rng = np.random.default_rng(12)
    A, B, C = case(rng, 1200, 1100, 900)
    say("noalloc large call 1200x1100x900")
    call(gpu, L, A, B, C, False, noalloc=True)
    A, B, C = case(rng, 130, 70, 65)
    want = call(gpu, L, A, B, C, True, noalloc=False)
    for r in range(2):
        say(f"noalloc small call 130x70x65 repeat {r}")
        assert np.array_equal(call(gpu, L, A, B, C, True, noalloc=True), want)


def test_noalloc_refuses_problems_past_its_scratch(gpu):
    L = bind()
    M = N = 8192
    K = 4096
    assert L.simd2_minus_square_workspace(N, M, K) > MAX_BYTES
    torch = gpu
    t = torch.zeros(16, device="cuda")
    p = t.data_ptr()
    # The size check comes before any device work; the pointers are never read.
    st = L.simd2_minus_square_exact_noalloc(N, M, K, p, N, p, K, p, N, p + 64, N, 0,
                                            torch.cuda.current_stream().cuda_stream)
    assert st == WORKSPACE_ERROR, st
