"""CPU tests of the C ABI: libsimd2.so loads, exports every function the
headers declare, and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

import simd2_amd

HEADERS = [os.path.join(ROOT, "include", "simd2", h)
           for h in ("srgemm.h", "converge.h", "tensor_emu.h", "knn.h", "batch.h", "schedule.h",
                     "learned.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|bool|char|void|size_t)\s*\*?\s*(\w+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", simd2_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(simd2_amd.LIB_PATH):
        simd2_amd.build()
    return simd2_amd.lib()


def test_headers_declare_expected_surface():
    names = declared_functions()
    for n in simd2_amd.EXPORTS:
        assert n in names, n
    assert len(names) >= 19


def test_library_exports_every_declared_symbol(lib):
    missing = declared_functions() - exported_symbols()
    assert not missing, missing


def test_status_strings(lib):
    assert simd2_amd.status_string(0) == "success"
    assert "overlaps" in simd2_amd.status_string(simd2_amd.ERROR_ALIASING)
    assert simd2_amd.status_string(999) == "unknown status"


# Fake, never-dereferenced device addresses: every call below must be
# rejected by the host-side checks, before any kernel launch.
P1, P2, P3, P4 = 1 << 20, 2 << 20, 3 << 20, 4 << 20


@pytest.mark.parametrize("name", list(simd2_amd.DROPIN.values()))
def test_dropin_argument_errors(lib, name):
    fn = getattr(lib, name)
    assert fn(-1, 4, 4, P1, 4, P2, 4, P3, 4, P4, True, None) == simd2_amd.ERROR_INVALID_PROBLEM
    assert fn(4, 4, 4, P1, 3, P2, 4, P3, 4, P4, True, None) == simd2_amd.ERROR_INVALID_PROBLEM
    assert fn(4, 4, 4, P1, 4, P2, 4, P3, 4, None, True, None) == simd2_amd.ERROR_NULL_POINTER
    assert fn(4, 4, 4, None, 4, P2, 4, P3, 4, P4, True, None) == simd2_amd.ERROR_NULL_POINTER
    assert fn(4, 4, 4, P1, 4, P2, 4, None, 4, P4, True, None) == simd2_amd.ERROR_NULL_POINTER
    # D overlapping A (or B) is rejected; C == D is allowed
    assert fn(4, 4, 4, P1, 4, P2, 4, P3, 4, P1 + 8, True, None) == simd2_amd.ERROR_ALIASING
    assert fn(4, 4, 4, P1, 4, P2, 4, P3, 4, P2, True, None) == simd2_amd.ERROR_ALIASING
    # empty outputs are a successful no-op
    assert fn(0, 4, 4, P1, 4, P2, 4, P3, 4, P4, True, None) == simd2_amd.SUCCESS
    assert fn(4, 0, 4, P1, 4, P2, 4, P3, 4, P4, True, None) == simd2_amd.SUCCESS


def test_generic_entry_point_errors(lib):
    f = lib.simd2_srgemm
    assert f(9, 4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, None) == simd2_amd.ERROR_NOT_SUPPORTED
    assert f(-1, 4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, None) == simd2_amd.ERROR_NOT_SUPPORTED
    assert f(0, 4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 3, 1, None) == simd2_amd.ERROR_INVALID_PROBLEM
    g = lib.simd2_orand_packed
    assert g(4, 40, 4, P1, 1, P2, 1, P3, 2, P4, 2, 1, None) == simd2_amd.ERROR_INVALID_PROBLEM
    assert g(4, 40, 4, P1, 1, P2, 2, P3, 2, None, 2, 1, None) == simd2_amd.ERROR_NULL_POINTER
    assert g(4, 40, 40, P1, 2, P2, 2, P3, 2, P1, 2, 1, None) == simd2_amd.ERROR_ALIASING


def test_minus_square_exact_workspace_and_argument_errors(lib):
    """The workspace size (int8 panels padded to 256 X rows / 128 Y columns
    and to a multiple of 64 in k, norms, tile gates, each 256-byte aligned)
    and the host-side refusals, none of which reaches the GPU."""
    f = lib.simd2_minus_square_workspace
    a256 = lambda b: (b + 255) // 256 * 256

    def want(M, N, K):  # column-major M x N: the row-major N x M problem
        R, Cn, Kp = N, M, (K + 63) // 64 * 64
        Rq, Cq = (R + 255) // 256 * 256, (Cn + 127) // 128 * 128
        return (a256(Rq * Kp) + a256(Cq * Kp) + a256(4 * R) + a256(4 * Cn) +
                a256(4 * (Rq // 64) * (Cq // 64)))
    for M, N, K in ((300, 200, 500), (1, 1, 1), (128, 129, 64), (4096, 4096, 4096),
                    (5, 7, 65536)):
        assert f(M, N, K) == want(M, N, K), (M, N, K)
    assert f(300, 200, 65537) == 0 and f(0, 4, 4) == 0 and f(4, 4, 0) == 0
    g = lib.simd2_minus_square_exact
    need = f(4, 4, 4)
    W = 16 << 20
    assert g(4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, None, need, None) == \
        simd2_amd.ERROR_WORKSPACE
    assert g(4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, W, need - 1, None) == \
        simd2_amd.ERROR_WORKSPACE
    assert g(4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, P4 - 8, need, None) == \
        simd2_amd.ERROR_ALIASING
    assert g(4, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, P1 + 4, need, None) == \
        simd2_amd.ERROR_ALIASING
    assert g(-1, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, W, need, None) == \
        simd2_amd.ERROR_INVALID_PROBLEM
    assert g(4, 4, 4, None, 4, P2, 4, P3, 4, P4, 4, 1, W, need, None) == \
        simd2_amd.ERROR_NULL_POINTER
    assert g(0, 4, 4, P1, 4, P2, 4, P3, 4, P4, 4, 1, None, 0, None) == simd2_amd.SUCCESS
    # K past the int8 limit: all tiles are direct, and the count needs no GPU
    mf, va = ctypes.c_int(-1), ctypes.c_int(-1)
    assert lib.simd2_minus_square_tiles(None, 100, 130, 70000, None, ctypes.byref(mf),
                                        ctypes.byref(va)) == simd2_amd.SUCCESS
    assert (mf.value, va.value) == (0, 2 * 3)


def test_converge_argument_errors(lib):
    itr = ctypes.c_int(-5)
    f = lib.simd2_converge_f32
    assert f(1, 8, P1, 8, P2, 8, P3, 1, 0, 10, 1, None, ctypes.byref(itr)) == \
        simd2_amd.ERROR_NOT_SUPPORTED          # plus_mul has no fixpoint loop
    assert itr.value == 0
    assert f(0, 8, P1, 8, P2, 8, P3, 1, 7, 10, 1, None, ctypes.byref(itr)) == \
        simd2_amd.ERROR_NOT_SUPPORTED          # unknown check mode
    assert f(0, 8, P1, 7, P2, 8, P3, 1, 0, 10, 1, None, ctypes.byref(itr)) == \
        simd2_amd.ERROR_INVALID_PROBLEM
    assert f(0, 8, P1, 8, P1 + 16, 8, P3, 1, 0, 10, 1, None, ctypes.byref(itr)) == \
        simd2_amd.ERROR_ALIASING
    assert f(0, 8, None, 8, P2, 8, P3, 1, 0, 10, 1, None, ctypes.byref(itr)) == \
        simd2_amd.ERROR_NULL_POINTER
    assert f(0, 0, None, 0, None, 0, None, 1, 0, 10, 1, None, ctypes.byref(itr)) == \
        simd2_amd.SUCCESS


def test_knn_argument_errors(lib):
    import simd2_amd.knn as K
    L = K._bind()
    sel = L.simd2_knn_select
    need = L.simd2_knn_select_workspace_size(100, 50, 33)
    assert need > 0
    assert sel(P1, 100, 50, 0, P2, P3, 0, P4, need, None) == simd2_amd.ERROR_INVALID_PROBLEM
    assert sel(P1, 100, 50, 51, P2, P3, 0, P4, need, None) == simd2_amd.ERROR_INVALID_PROBLEM
    assert sel(P1, 100, 10, 16, P2, P3, 0, P4, need, None) == simd2_amd.ERROR_INVALID_PROBLEM
    assert sel(None, 100, 50, 16, P2, P3, 0, P4, need, None) == simd2_amd.ERROR_NULL_POINTER
    # any k <= ref_nb is accepted; the workspace is checked before any launch
    assert sel(P1, 100, 50, 33, P2, P3, 0, None, 0, None) == simd2_amd.ERROR_WORKSPACE
    assert sel(P1, 100, 50, 50, P2, P3, 0, P4,
               L.simd2_knn_select_workspace_size(100, 50, 50) - 1, None) == simd2_amd.ERROR_WORKSPACE
    # queries are batched: the workspace stays bounded at the benchmark's size and for huge k
    assert L.simd2_knn_select_workspace_size(16384, 16384, 16) < (300 << 20)
    assert L.simd2_knn_select_workspace_size(16384, 16384, 16384) < (300 << 20)
    assert L.simd2_knn_select_workspace_size(10, 20, 21) == 0
    ws = L.simd2_knn_bf16_workspace_size(1000, 2000, 128)
    assert ws > 1000 * 128 * 2 + 2000 * 128 * 2
    args = (P1, 128, P2, 128, 1000, 2000, 128, 16, P3, P4, 0)
    assert L.simd2_knn_bf16(*args, P4 + (1 << 20), ws - 1, None, None) == simd2_amd.ERROR_WORKSPACE
    assert L.simd2_knn_bf16(P1, 128, P2, 128, 1000, 2000, 128, 257, P3, P4, 0, P4, ws, None,
                            None) == simd2_amd.ERROR_INVALID_PROBLEM   # k > 256
    assert L.simd2_knn_bf16(P1, 128, P2, 128, 1000, 20, 128, 21, P3, P4, 0, P4, ws, None,
                            None) == simd2_amd.ERROR_INVALID_PROBLEM   # k > ref_nb
    assert L.simd2_knn_bf16(P1, 100, P2, 128, 1000, 2000, 128, 16, P3, P4, 0, P4, ws, None,
                            None) == simd2_amd.ERROR_INVALID_PROBLEM   # ldq < dim
    # any dim: the workspace grows with it
    assert L.simd2_knn_bf16_workspace_size(1000, 2000, 4096) > 3 * ws


def test_batch_argument_errors(lib):
    import simd2_amd.batch as B
    L = B._bind()
    h = ctypes.c_void_p()

    def create(specs, max_rounds=0):
        arr = (B._Graph * len(specs))(*[B._Graph(*s) for s in specs])
        return L.simd2_batch_create(arr, len(specs), max_rounds, ctypes.byref(h))

    assert L.simd2_batch_create(None, 1, 0, ctypes.byref(h)) == simd2_amd.ERROR_NULL_POINTER
    assert L.simd2_batch_create(None, 0, 0, None) == simd2_amd.ERROR_NULL_POINTER
    assert L.simd2_batch_create(None, -1, 0, ctypes.byref(h)) == simd2_amd.ERROR_INVALID_PROBLEM
    assert create([(8, 0, 1, 1)], -1) == simd2_amd.ERROR_INVALID_PROBLEM
    assert create([(-1, 0, 1, 1)]) == simd2_amd.ERROR_INVALID_PROBLEM
    # plus_mul and minus_square have no fixpoint loop; unknown ops and modes
    for op in (1, 8, 9, -1):
        assert create([(8, 0, 1, 1), (8, op, 1, 1)]) == simd2_amd.ERROR_NOT_SUPPORTED
    assert create([(8, 0, 1, 3)]) == simd2_amd.ERROR_NOT_SUPPORTED
    assert h.value is None
    # every other entry point rejects a null batch before touching the GPU
    assert L.simd2_batch_solve(None, None) == simd2_amd.ERROR_NULL_POINTER
    assert L.simd2_batch_autotune(None, None) == simd2_amd.ERROR_NULL_POINTER
    assert L.simd2_batch_upload(None, 0, P1, 8, None) == simd2_amd.ERROR_NULL_POINTER
    assert L.simd2_batch_iterations(None, None, None, None) == simd2_amd.ERROR_NULL_POINTER
    assert L.simd2_batch_destroy(None) == simd2_amd.SUCCESS


def test_schedule_and_learned_argument_errors(lib):
    import simd2_amd.learned as LG
    import simd2_amd.schedule as S
    L = LG._bind()
    itr = ctypes.c_int(-1)
    sc = S.Schedule(-1, 0)
    f = L.simd2_converge_scheduled
    ws = L.simd2_converge_scheduled_workspace(300, -1, 16)
    assert ws > 2 * 384 * 384 * 4
    assert L.simd2_converge_scheduled_workspace(0, -1, 16) == 0
    assert f(1, 300, P1, 300, 1, 1, 16, ctypes.byref(sc), None, P2, ws, None, ctypes.byref(itr),
             None) == simd2_amd.ERROR_NOT_SUPPORTED            # plus_mul
    assert f(0, 300, P1, 299, 1, 1, 16, None, None, P2, ws, None, None, None) == \
        simd2_amd.ERROR_INVALID_PROBLEM
    assert f(0, 300, P1, 300, 1, 1, 16, None, None, P2, ws - 1, None, None, None) == \
        simd2_amd.ERROR_WORKSPACE
    assert f(0, 300, None, 300, 1, 1, 16, None, None, P2, ws, None, None, None) == \
        simd2_amd.ERROR_NULL_POINTER
    bad = S.Schedule(7, 0)
    assert f(0, 300, P1, 300, 1, 1, 16, ctypes.byref(bad), None, P2, ws, None, None, None) == \
        simd2_amd.ERROR_INVALID_PROBLEM
    assert L.simd2_graph_features_workspace(40000) == 0
    fw = L.simd2_graph_features_workspace(100)
    assert L.simd2_graph_features(0, 100, P1, 100, 1, P2, P3, fw - 1, None) == \
        simd2_amd.ERROR_WORKSPACE
    assert L.simd2_graph_features(7, 100, P1, 100, 1, P2, P3, fw, None) == \
        simd2_amd.ERROR_NOT_SUPPORTED
    mw = L.simd2_mlp_workspace(64)
    assert L.simd2_mlp_gradient(P1, P2, P3, P4, 64, None, None, P1, mw - 1, None) == \
        simd2_amd.ERROR_WORKSPACE
    assert L.simd2_mlp_train_step(P1, P2, P3, 0, 1e-3, P4, P1, P2, 64, None, P3, mw, None) == \
        simd2_amd.ERROR_INVALID_PROBLEM                         # steps count from 1
    lw = L.simd2_converge_learned_workspace(500, 32)
    assert lw > ws
    assert L.simd2_converge_learned(0, 500, P1, 500, 1, 1, 32, None, P2, lw, None, None, None,
                                    None) == simd2_amd.ERROR_NULL_POINTER
