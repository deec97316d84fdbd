"""simd2_amd -- semiring GEMM on AMD Instinct MI355X (gfx950).

Python host mirror of the reference's operator surface (``kernel/srgemm.cuh``
``cuasr_*_srsgemm``, ``kernel/converge.cuh`` ``comp_update*``) over the C ABI
of ``libsimd2.so`` (declared in ``include/simd2/*.h``).  Device memory and
streams come from PyTorch tensors; every call goes through ctypes to the HIP
kernels, and there is no CPU fallback: without the built library or a GPU
the calls raise.

Conventions follow the reference exactly:

* ``cuasr_<op>_srsgemm(M, N, K, A, lda, B, ldb, C, ldc, D, do_epilogue)`` is
  column-major and returns the integer status (0 = success), like
  ``kernel/srgemm.cuh``;
* ``srgemm(op, A, B, C, epilogue)`` is the row-major convenience form the
  reference's tests use (``tests/cuAsr_kernel_test.cu:15-21,60``: a row-major
  ``A @ B`` is the column-major call ``(N, M, K, B, N, A, K, C, N, D)``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimd2.so")

# The reference's ``enum op`` order (kernel/tensor_srgemm.cuh:18-26).
OPS = ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul",
       "min_max", "or_and", "minus_square"]
OP = {name: i for i, name in enumerate(OPS)}

# Function names of the drop-ins (kernel/srgemm.cuh), by op.
DROPIN = {
    "min_plus": "cuasr_minplus_srsgemm",
    "plus_mul": "cuasr_plusmul_srsgemm",
    "max_plus": "cuasr_maxplus_srsgemm",
    "max_min": "cuasr_maxmin_srsgemm",
    "max_mul": "cuasr_maxmul_srsgemm",
    "min_mul": "cuasr_minmul_srsgemm",
    "min_max": "cuasr_minmax_srsgemm",
    "or_and": "cuasr_orand_srsgemm",
    "minus_square": "cuasr_plusmiussquare_srsgemm",
}

# Status codes (include/simd2/srgemm.h).
SUCCESS = 0
ERROR_MISALIGNED_OPERAND = 1
ERROR_INVALID_PROBLEM = 4
ERROR_NOT_SUPPORTED = 5
ERROR_INTERNAL = 7
ERROR_NULL_POINTER = 11
ERROR_ALIASING = 12
ERROR_NO_DEVICE = 13
ERROR_NOT_CONVERGED = 14
ERROR_WORKSPACE = 15

CHECK = {"exact": 0, "threshold": 1, "aplp": 2}
CONVERGE_REFERENCE = 0
CONVERGE_FUSED = 1

# Symbols every build must export (the header declarations).
EXPORTS = list(DROPIN.values()) + [
    "simd2_srgemm", "simd2_orand_packed", "simd2_status_string", "simd2_device_info",
    "comp_update", "comp_update2", "comp_update_int",
    "simd2_converge_f32", "simd2_converge_i32", "simd2_converge_orand_packed",
    "tensor_srgemm", "simd2_tensor_srgemm", "f2h_device", "h2f_device",
    "simd2_srgemm_ex", "simd2_tile_config_count", "simd2_pick_tile_config", "simd2_autotune",
    "simd2_minus_square_workspace", "simd2_minus_square_exact", "simd2_minus_square_tiles",
    "simd2_minus_square_exact_noalloc",
    "simd2_knn_select_workspace_size", "simd2_knn_select", "knn_mxu", "simd2_knn_bf16_workspace_size", "simd2_knn_bf16",
    "simd2_batch_create", "simd2_batch_destroy", "simd2_batch_upload", "simd2_batch_download",
    "simd2_batch_buffer", "simd2_batch_autotune", "simd2_batch_solve",
    "simd2_batch_iterations", "simd2_batch_schedule",
    "simd2_converge_scheduled_workspace", "simd2_converge_scheduled",
    "simd2_graph_features_workspace", "simd2_graph_features", "simd2_graph_order",
    "simd2_mlp_forward", "simd2_mlp_workspace", "simd2_mlp_gradient", "simd2_mlp_train_step",
    "simd2_converge_learned_workspace", "simd2_converge_learned",
]


class Simd2Error(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: status {status} ({status_string(status)})")


def build(jobs: int = 8) -> None:
    """Compile libsimd2.so for gfx950 (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded libsimd2.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run simd2_amd.build() first")
        L = ctypes.CDLL(LIB_PATH)
        i, p, b = ctypes.c_int, ctypes.c_void_p, ctypes.c_bool
        for name in DROPIN.values():
            fn = getattr(L, name)
            fn.argtypes = [i, i, i, p, i, p, i, p, i, p, b, p]
            fn.restype = i
        L.simd2_srgemm.argtypes = [i, i, i, i, p, i, p, i, p, i, p, i, i, p]
        L.simd2_srgemm.restype = i
        L.simd2_srgemm_ex.argtypes = [i, i, i, i, p, i, p, i, p, i, p, i, i, i, p]
        L.simd2_srgemm_ex.restype = i
        L.simd2_tile_config_count.argtypes = []
        L.simd2_tile_config_count.restype = i
        L.simd2_pick_tile_config.argtypes = [i, i, i, i]
        L.simd2_pick_tile_config.restype = i
        L.simd2_autotune.argtypes = [i, i, i, i, p, i, p, i, p, i, p, i, i, p,
                                     ctypes.POINTER(i), ctypes.POINTER(ctypes.c_float)]
        L.simd2_autotune.restype = i
        L.simd2_orand_packed.argtypes = [i, i, i, p, i, p, i, p, i, p, i, i, p]
        L.simd2_orand_packed.restype = i
        L.simd2_status_string.argtypes = [i]
        L.simd2_status_string.restype = ctypes.c_char_p
        L.simd2_device_info.argtypes = [ctypes.POINTER(i), ctypes.POINTER(i), ctypes.c_char_p, i]
        L.simd2_device_info.restype = i
        L.simd2_minus_square_workspace.argtypes = [i, i, i]
        L.simd2_minus_square_workspace.restype = ctypes.c_size_t
        L.simd2_minus_square_exact.argtypes = [i, i, i, p, i, p, i, p, i, p, i, i, p,
                                               ctypes.c_size_t, p]
        L.simd2_minus_square_exact.restype = i
        L.simd2_minus_square_tiles.argtypes = [p, i, i, i, p, ctypes.POINTER(i),
                                               ctypes.POINTER(i)]
        L.simd2_minus_square_tiles.restype = i
        for name in ("comp_update", "comp_update2", "comp_update_int"):
            fn = getattr(L, name)
            fn.argtypes = [p, p, p, p, i, i]
            fn.restype = b
        L.simd2_converge_f32.argtypes = [i, i, p, i, p, i, p, i, i, i, i, p, ctypes.POINTER(i)]
        L.simd2_converge_f32.restype = i
        L.simd2_converge_i32.argtypes = [i, p, i, p, i, p, i, i, p, ctypes.POINTER(i)]
        L.simd2_converge_i32.restype = i
        L.simd2_converge_orand_packed.argtypes = [i, p, i, p, i, p, i, p, ctypes.POINTER(i)]
        L.simd2_converge_orand_packed.restype = i
        L.tensor_srgemm.argtypes = [p, p, p, p, i, i, i, b, i]
        L.tensor_srgemm.restype = i
        L.simd2_tensor_srgemm.argtypes = [p, p, p, p, i, i, i, i, i, p]
        L.simd2_tensor_srgemm.restype = i
        L.f2h_device.argtypes = [p, p, i]
        L.f2h_device.restype = None
        L.h2f_device.argtypes = [p, p, i]
        L.h2f_device.restype = None
        _lib = L
    return _lib


def status_string(status: int) -> str:
    try:
        return lib().simd2_status_string(int(status)).decode()
    except RuntimeError:
        return "library not loaded"


def _torch():
    import torch
    return torch


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    return t.data_ptr()


def _stream(stream) -> Optional[int]:
    torch = _torch()
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _check(status: int, where: str) -> None:
    if status != SUCCESS:
        raise Simd2Error(status, where)


# ---- the reference's operator surface (column-major, status int) ----------

def cuasr_srsgemm(op, M, N, K, A, lda, B, ldb, C, ldc, D, do_epilogue, stream=None) -> int:
    """``cuasr_<op>_srsgemm`` of kernel/srgemm.cuh with torch device tensors."""
    name = DROPIN[op if isinstance(op, str) else OPS[op]]
    return getattr(lib(), name)(M, N, K, _ptr(A), lda, _ptr(B), ldb, _ptr(C), ldc, _ptr(D),
                                bool(do_epilogue), _stream(stream))


def _make_dropin(opname):
    def fn(M, N, K, A, lda, B, ldb, C, ldc, D, do_epilogue_min, stream=None):
        return cuasr_srsgemm(opname, M, N, K, A, lda, B, ldb, C, ldc, D, do_epilogue_min, stream)
    fn.__name__ = DROPIN[opname]
    fn.__doc__ = f"Drop-in for {DROPIN[opname]} (kernel/srgemm.cuh); returns the status."
    return fn


for _opname, _fname in DROPIN.items():
    globals()[_fname] = _make_dropin(_opname)


def simd2_srgemm(op, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, do_epilogue, stream=None) -> int:
    opc = OP[op] if isinstance(op, str) else int(op)
    return lib().simd2_srgemm(opc, M, N, K, _ptr(A), lda, _ptr(B), ldb, _ptr(C), ldc, _ptr(D),
                              ldd, int(bool(do_epilogue)), _stream(stream))


def srgemm(op, A, B, C=None, epilogue=False, out=None, stream=None, tile_config=None):
    """Row-major ``D = [C (+)] A (x) B`` for torch device tensors (float32, or
    int32 for or_and), through the column-major drop-in the way the
    reference's tests call it (``tile_config`` selects the VALU tile shape
    through simd2_srgemm_ex).  Returns D."""
    torch = _torch()
    opname = op if isinstance(op, str) else OPS[op]
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError("inner dimensions differ")
    dt = torch.int32 if opname == "or_and" else torch.float32
    if out is None:
        out = torch.empty((M, N), dtype=dt, device=A.device)
    if tile_config is None:
        st = cuasr_srsgemm(opname, N, M, K, B, max(N, 1), A, max(K, 1),
                           C if epilogue else None, max(N, 1), out, bool(epilogue), stream)
    else:
        st = lib().simd2_srgemm_ex(OP[opname], N, M, K, _ptr(B), max(N, 1), _ptr(A), max(K, 1),
                                   _ptr(C) if epilogue else None, max(N, 1), _ptr(out),
                                   max(N, 1), int(bool(epilogue)), int(tile_config),
                                   _stream(stream))
    _check(st, DROPIN[opname])
    return out


def autotune(op, A, B, C=None, epilogue=False, out=None, stream=None):
    """Autotune the VALU tile configuration for the row-major problem
    D = [C (+)] A (x) B (simd2_autotune).  Returns (D, best_config, ms_per_config)."""
    torch = _torch()
    opname = op if isinstance(op, str) else OPS[op]
    M, K = A.shape
    N = B.shape[1]
    if out is None:
        out = torch.empty((M, N), dtype=torch.float32, device=A.device)
    best = ctypes.c_int(-1)
    ms = (ctypes.c_float * 8)()
    st = lib().simd2_autotune(OP[opname], N, M, K, _ptr(B), max(N, 1), _ptr(A), max(K, 1),
                              _ptr(C) if epilogue else None, max(N, 1), _ptr(out), max(N, 1),
                              int(bool(epilogue)), _stream(stream), ctypes.byref(best), ms)
    _check(st, "simd2_autotune")
    n = lib().simd2_tile_config_count()
    return out, best.value, [ms[i] for i in range(n)]


def orand_packed(A, B, N, C=None, epilogue=False, out=None, stream=None):
    """Bit-packed or_and: A (M x ceil(K/32) words), B (K x ceil(N/32) words),
    both int32 tensors holding uint32 bit rows.  Returns D (M x ceil(N/32))."""
    torch = _torch()
    M, kw = A.shape
    K = B.shape[0]
    nw = (N + 31) // 32
    if out is None:
        out = torch.empty((M, nw), dtype=torch.int32, device=A.device)
    st = lib().simd2_orand_packed(M, N, K, _ptr(A), kw, _ptr(B), B.shape[1],
                                  _ptr(C) if epilogue else None, C.shape[1] if epilogue else nw,
                                  _ptr(out), out.shape[1], int(bool(epilogue)), _stream(stream))
    _check(st, "simd2_orand_packed")
    return out


# ---- convergence ----------------------------------------------------------

def comp_update(prev, cur, check, mode="threshold"):
    """Drop-in comp_update / comp_update2 / comp_update(int) (kernel/converge.cuh).
    Synchronous; returns whether anything changed."""
    h = ctypes.c_int(0)
    M, N = prev.shape[0], prev.numel() // max(prev.shape[0], 1)
    L = lib()
    if prev.dtype == _torch().int32:
        fn = L.comp_update_int
    elif mode == "aplp":
        fn = L.comp_update2
    else:
        fn = L.comp_update
    return bool(fn(_ptr(prev), _ptr(cur), _ptr(check), ctypes.addressof(h), M, N))


def converge(op, X, epilogue=True, check="exact", max_itr=500, flags=CONVERGE_FUSED,
             work=None, flag=None, stream=None):
    """Repeated squaring of the n x n row-major X in place until unchanged.
    Returns (X, num_itr, status)."""
    torch = _torch()
    n = X.shape[0]
    opname = op if isinstance(op, str) else OPS[op]
    if work is None:
        work = torch.empty_like(X)
    if flag is None:
        flag = torch.zeros(1, dtype=torch.int32, device=X.device)
    itr = ctypes.c_int(0)
    L = lib()
    if opname == "or_and" and X.dtype == torch.int32 and X.shape[1] == n:
        st = L.simd2_converge_i32(n, _ptr(X), X.stride(0), _ptr(work), work.stride(0), _ptr(flag),
                                  max_itr, flags, _stream(stream), ctypes.byref(itr))
    elif opname == "or_and":
        st = L.simd2_converge_orand_packed(n, _ptr(X), X.stride(0), _ptr(work), work.stride(0),
                                           _ptr(flag), max_itr, _stream(stream), ctypes.byref(itr))
    else:
        st = L.simd2_converge_f32(OP[opname], n, _ptr(X), X.stride(0), _ptr(work), work.stride(0),
                                  _ptr(flag), int(bool(epilogue)), CHECK[check], max_itr, flags,
                                  _stream(stream), ctypes.byref(itr))
    if st not in (SUCCESS, ERROR_NOT_CONVERGED):
        raise Simd2Error(st, "converge")
    return X, itr.value, st


def tensor_srgemm(op, A16, B16, C=None, epilogue=False, out=None, stream=None):
    """FP16-input emulation GEMM (kernel/tensor_srgemm.cuh semantics), row-major
    torch tensors: A16 (m x k) and B16 (k x n) float16, C and the result float32."""
    torch = _torch()
    m, k = A16.shape
    n = B16.shape[1]
    if out is None:
        out = torch.empty((m, n), dtype=torch.float32, device=A16.device)
    opc = OP[op] if isinstance(op, str) else int(op)
    st = lib().simd2_tensor_srgemm(_ptr(A16), _ptr(B16), _ptr(C) if epilogue else None,
                                   _ptr(out), m, n, k, int(bool(epilogue)), opc, _stream(stream))
    _check(st, "tensor_srgemm")
    return out


def f2h(x):
    """f2h_device (kernel/precision.cuh): float32 -> float16, synchronous."""
    torch = _torch()
    out = torch.empty(x.shape, dtype=torch.float16, device=x.device)
    lib().f2h_device(_ptr(out), _ptr(x), x.numel())
    return out


def h2f(x):
    """h2f_device (kernel/precision.cuh): float16 -> float32, synchronous."""
    torch = _torch()
    out = torch.empty(x.shape, dtype=torch.float32, device=x.device)
    lib().h2f_device(_ptr(out), _ptr(x), x.numel())
    return out


def minus_square_workspace(M, N, K):
    """Workspace bytes of minus_square_exact for row-major A (M x K), B (K x N)."""
    return lib().simd2_minus_square_workspace(N, M, K)


def minus_square_exact(A, B, C=None, epilogue=False, out=None, workspace=None, stream=None):
    """Row-major ``D = [C +] sum_k (A[i,k] - B[k,j])^2`` through
    simd2_minus_square_exact: tiles of small-integer operands on the int8
    matrix cores (bit-identical to the direct kernel), the rest direct.
    ``workspace`` (a uint8 device tensor) is allocated when not given.
    Returns (D, workspace)."""
    torch = _torch()
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError("inner dimensions differ")
    if out is None:
        out = torch.empty((M, N), dtype=torch.float32, device=A.device)
    need = minus_square_workspace(M, N, K)
    if workspace is None:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=A.device)
    st = lib().simd2_minus_square_exact(N, M, K, _ptr(B), max(N, 1), _ptr(A), max(K, 1),
                                        _ptr(C) if epilogue else None, max(N, 1), _ptr(out),
                                        max(N, 1), int(bool(epilogue)), _ptr(workspace),
                                        workspace.numel(), _stream(stream))
    _check(st, "simd2_minus_square_exact")
    return out, workspace


def minus_square_tiles(workspace, M, N, K, stream=None):
    """(64x64 tiles on the int8 matrix cores, tiles on the direct kernel) of
    the last minus_square_exact on ``workspace`` for this row-major problem."""
    mf, va = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().simd2_minus_square_tiles(_ptr(workspace), N, M, K, _stream(stream),
                                          ctypes.byref(mf), ctypes.byref(va)),
           "simd2_minus_square_tiles")
    return mf.value, va.value


def device_info():
    cu, clk = ctypes.c_int(0), ctypes.c_int(0)
    name = ctypes.create_string_buffer(64)
    st = lib().simd2_device_info(ctypes.byref(cu), ctypes.byref(clk), name, 64)
    _check(st, "simd2_device_info")
    return {"cu_count": cu.value, "clock_khz": clk.value, "arch": name.value.decode()}


# ---- bit packing helpers (torch, on the tensor's device) --------------------

def pack_bits(M):
    """Pack a 0/1 matrix (rows x cols) into int32 words, bit j%32 of word j//32."""
    torch = _torch()
    rows, cols = M.shape
    nw = (cols + 31) // 32
    pad = torch.zeros((rows, nw * 32), dtype=torch.int64, device=M.device)
    pad[:, :cols] = (M != 0).to(torch.int64)
    w = pad.view(rows, nw, 32) << torch.arange(32, device=M.device, dtype=torch.int64)
    packed = w.sum(dim=2)
    return (packed - (packed >= 2 ** 31).to(torch.int64) * 2 ** 32).to(torch.int32)


def unpack_bits(P, cols):
    torch = _torch()
    rows, nw = P.shape
    w = P.to(torch.int64) & 0xFFFFFFFF
    bits = (w.unsqueeze(2) >> torch.arange(32, device=P.device, dtype=torch.int64)) & 1
    return bits.view(rows, nw * 32)[:, :cols].to(torch.int32)
