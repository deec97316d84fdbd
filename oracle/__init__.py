"""CPU oracle for the semiring-GEMM path -- TEST INFRASTRUCTURE ONLY.

Python access (numpy in, numpy out) to ``liboracle.so``, the plain-C
restatement of the reference's ``kernel/host_srgemm.h:16-85`` and
``kernel/converge.cuh:6-85`` (see ``oracle/host_srgemm.c``), and to
``_ref/libref_host_srgemm.so``, the reference's own header compiled by
``oracle/Makefile`` when the reference tree is present.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker or the timed CPU baseline.
The product path (``simd2_amd``) never does.

Pinning (what checks this restatement):
  * ``srgemm_f32`` / ``srgemm_i32``: bit-exact against the reference's own
    ``host_srgemm`` (``_ref``) on the reference test's inputs
    (``tests/golden/``), plus closed forms (identity / annihilator matrices,
    path-graph prefix sums), brute-force numpy on tiny inputs and FP64
    tolerance for plus_mul / minus_square (``tests/test_oracle.py``).
  * ``comp_update*`` / ``converge_*``: iteration counts on the reference's
    path-graph fixtures (ceil(log2(n-1)) + 1 squarings), scipy Dijkstra rows
    for APSP, and brute-force Floyd-Warshall on tiny graphs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_host_srgemm.so")

# Operator codes: the reference's ``enum op`` (kernel/tensor_srgemm.cuh:18-26).
OPS = ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul",
       "min_max", "or_and", "minus_square"]
OP = {name: i for i, name in enumerate(OPS)}

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference tree exists)."""
    # make's own messages go to stderr so callers' stdout (bench.py's single
    # JSON line) stays clean
    subprocess.run(["make", "-s", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        i, ll = ctypes.c_int, ctypes.c_int64
        lib.oracle_srgemm_f32.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, i]
        lib.oracle_srgemm_f32_omp.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, i, i]
        lib.oracle_srgemm_f32_rows.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, i, i, i, i]
        lib.oracle_srgemm_i32.argtypes = [i, i, i, i, _i32p, _i32p, _i32p, _i32p, i]
        lib.oracle_add_identity_f32.argtypes = [i]
        lib.oracle_add_identity_f32.restype = ctypes.c_float
        lib.oracle_comp_update_f32.argtypes = [_f32p, _f32p, ll]
        lib.oracle_comp_update2_f32.argtypes = [_f32p, _f32p, ll]
        lib.oracle_comp_update_i32.argtypes = [_i32p, _i32p, ll]
        lib.oracle_comp_update_exact_f32.argtypes = [_f32p, _f32p, ll]
        lib.oracle_converge_f32.argtypes = [i, i, _f32p, i, i, i, i]
        lib.oracle_converge_i32.argtypes = [i, _i32p, i]
        lib.oracle_knn_insertion_sort.argtypes = [_f32p, _i32p, i, i, i]
        lib.oracle_knn_insertion_sort.restype = None
        lib.oracle_knn_sqrt.argtypes = [_f32p, i, i]
        lib.oracle_knn_sqrt.restype = None
        _lib = lib
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def _load_ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(_REF_PATH)
        i = ctypes.c_int
        lib.ref_host_srgemm_f32.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, i]
        lib.ref_host_srgemm_i32.argtypes = [i, i, i, _i32p, _i32p, _i32p, _i32p, i]
        lib.ref_rng_init_matrix.argtypes = [_f32p, i, i]
        lib.ref_test_sequence.argtypes = [i, _i32p]
        _ref = lib
    return _ref


def _op(op) -> int:
    return OP[op] if isinstance(op, str) else int(op)


def add_identity(op) -> np.float32:
    return np.float32(_load().oracle_add_identity_f32(_op(op)))


def _c(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def srgemm_f32(op, A, B, C=None, epilogue=False, D=None, threads=1):
    """Row-major ``D = [C (+)] (+)_k A[i,k] (x) B[k,j]`` as host_srgemm does it.

    ``D`` is the prior content that the reference accumulates into; by
    default it is the op's additive identity.  ``threads>1`` uses the OpenMP
    row split, which is bit-identical."""
    op = _op(op)
    A, B = _c(A, np.float32), _c(B, np.float32)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    C = _c(np.zeros((M, N), np.float32) if C is None else C, np.float32)
    if D is None:
        D = np.full((M, N), add_identity(op), np.float32)
    else:
        D = _c(D, np.float32).copy()
    lib = _load()
    if threads == 1:
        rc = lib.oracle_srgemm_f32(op, M, N, K, A, B, C, D, int(bool(epilogue)))
    else:
        rc = lib.oracle_srgemm_f32_omp(op, M, N, K, A, B, C, D, int(bool(epilogue)), threads)
    if rc:
        raise ValueError(f"oracle rejected op {op}")
    return D


def srgemm_f32_rows(op, A, B, C, D, epilogue, r0, r1, threads=0):
    """Compute rows [r0, r1) of D in place (row-sampled CPU timing)."""
    M, K = A.shape
    N = B.shape[1]
    rc = _load().oracle_srgemm_f32_rows(_op(op), M, N, K, A, B, C, D,
                                        int(bool(epilogue)), r0, r1, threads)
    if rc:
        raise ValueError("bad arguments")
    return D


def srgemm_i32(A, B, C=None, epilogue=False, D=None):
    """host_srgemm<int> with bin_or / bin_and (or_and on 0/1 ints)."""
    A, B = _c(A, np.int32), _c(B, np.int32)
    M, K = A.shape
    N = B.shape[1]
    C = _c(np.zeros((M, N), np.int32) if C is None else C, np.int32)
    D = np.zeros((M, N), np.int32) if D is None else _c(D, np.int32).copy()
    if _load().oracle_srgemm_i32(OP["or_and"], M, N, K, A, B, C, D, int(bool(epilogue))):
        raise ValueError("bad arguments")
    return D


def comp_update(prev, cur, mode="threshold"):
    """Returns (changed, new_prev, new_cur) per kernel/converge.cuh."""
    prev = prev.copy()
    cur = cur.copy()
    lib = _load()
    if prev.dtype == np.int32:
        ch = lib.oracle_comp_update_i32(prev, cur, prev.size)
    elif mode == "aplp":
        ch = lib.oracle_comp_update2_f32(prev, cur, prev.size)
    elif mode == "exact":
        ch = lib.oracle_comp_update_exact_f32(prev, cur, prev.size)
    else:
        ch = lib.oracle_comp_update_f32(prev, cur, prev.size)
    return bool(ch), prev, cur


CHECK_MODES = {"threshold": 0, "aplp": 1, "exact": 2}


def converge_f32(op, X, epilogue=True, check="threshold", max_itr=500, threads=0):
    """Repeated squaring until unchanged (apps/apsp/cuASR/apsp_cuASR.cu:37-54).
    Returns (X_final, num_itr)."""
    X = _c(X, np.float32).copy()
    n = X.shape[0]
    itr = _load().oracle_converge_f32(_op(op), n, X, int(bool(epilogue)),
                                      CHECK_MODES[check], max_itr, threads)
    if itr < 0:
        raise ValueError("bad arguments")
    return X, itr


def converge_i32(X, max_itr=500):
    """Transitive closure by or_and squaring (apps/gtc/cuASR/gtc_cuASR.cu:40-51)."""
    X = _c(X, np.int32).copy()
    itr = _load().oracle_converge_i32(X.shape[0], X, max_itr)
    return X, itr


def knn_insertion_sort(dist_rq, k, do_sqrt=True):
    """modified_insertion_sort_linear + compute_sqrt_linear
    (apps/pld/cuASR/knncuda.cu:170-237) on a distance matrix in the
    reference layout: dist_rq has shape (R, Q) (row r = distances of all
    queries to reference r).  Returns (dist k x Q, index k x Q)."""
    D = _c(dist_rq, np.float32).copy()
    R, Q = D.shape
    idx = np.zeros((max(R, k), Q), np.int32)
    _load().oracle_knn_insertion_sort(D, idx, Q, R, k)
    out = np.ascontiguousarray(D[:k])
    if do_sqrt:
        _load().oracle_knn_sqrt(out, Q, k)
    return out, np.ascontiguousarray(idx[:k])


# ---- the reference's own host_srgemm (only where the reference tree exists) ----

def ref_srgemm_f32(op, A, B, C, D, epilogue):
    D = _c(D, np.float32).copy()
    rc = _load_ref().ref_host_srgemm_f32(_op(op), A.shape[0], B.shape[1], A.shape[1],
                                         _c(A, np.float32), _c(B, np.float32),
                                         _c(C, np.float32), D, int(bool(epilogue)))
    if rc:
        raise ValueError("bad op")
    return D


def ref_srgemm_i32(A, B, C, D, epilogue):
    D = _c(D, np.int32).copy()
    _load_ref().ref_host_srgemm_i32(A.shape[0], B.shape[1], A.shape[1], _c(A, np.int32),
                                    _c(B, np.int32), _c(C, np.int32), D, int(bool(epilogue)))
    return D


def ref_rng_init_matrix(n, seed):
    buf = np.empty(n, np.float32)
    _load_ref().ref_rng_init_matrix(buf, n, seed)
    return buf


def ref_test_sequence(n_tests=25):
    out = np.zeros(n_tests * 5, np.int32)
    _load_ref().ref_test_sequence(n_tests, out)
    return out.reshape(n_tests, 5)
