"""k-nearest neighbours over the C ABI (include/simd2/knn.h): the
reference's PLD post-processing (knn_select, knn_mxu) and the BF16-MFMA
precision form (knn_bf16)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _ptr, _stream, _torch, lib

SQRT = 1


def _bind():
    L = lib()
    if getattr(L, "_knn_bound", False):
        return L
    i, p = ctypes.c_int, ctypes.c_void_p
    L.simd2_knn_select_workspace_size.argtypes = [i, i, i]
    L.simd2_knn_select_workspace_size.restype = ctypes.c_size_t
    L.simd2_knn_select.argtypes = [p, i, i, i, p, p, i, p, ctypes.c_size_t, p]
    L.simd2_knn_select.restype = i
    L.simd2_knn_bf16_workspace_size.argtypes = [i, i, i]
    L.simd2_knn_bf16_workspace_size.restype = ctypes.c_size_t
    L.simd2_knn_bf16.argtypes = [p, i, p, i, i, i, i, i, p, p, i, p, ctypes.c_size_t, p, p]
    L.simd2_knn_bf16.restype = i
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.knn_mxu.argtypes = [f32p, i, f32p, i, i, i, f32p, i32p, ctypes.POINTER(ctypes.c_double)]
    L.knn_mxu.restype = ctypes.c_bool
    L._knn_bound = True
    return L


def select_workspace_bytes(Q, R, k):
    return int(_bind().simd2_knn_select_workspace_size(Q, R, k))


def knn_select(dist_rq, k, sqrt=True, stream=None, workspace=None):
    """Top-k per query of a distance matrix in the reference layout
    (dist_rq: R x Q device tensor, row r = distances of all queries to
    reference r), any 1 <= k <= R.  ``workspace`` (uint8 device tensor) is
    allocated when not given.  Returns (dist k x Q, index k x Q)."""
    torch = _torch()
    R, Q = dist_rq.shape
    out_d = torch.empty((k, Q), dtype=torch.float32, device=dist_rq.device)
    out_i = torch.empty((k, Q), dtype=torch.int32, device=dist_rq.device)
    need = select_workspace_bytes(Q, R, k)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=dist_rq.device)
    st = _bind().simd2_knn_select(_ptr(dist_rq), Q, R, k, _ptr(out_d), _ptr(out_i),
                                  SQRT if sqrt else 0, _ptr(workspace), workspace.numel(),
                                  _stream(stream))
    _check(st, "simd2_knn_select")
    if stream is not None:
        s = stream if isinstance(stream, torch.cuda.Stream) else torch.cuda.ExternalStream(stream)
        for t in (workspace, out_d, out_i):
            t.record_stream(s)
    return out_d, out_i


def knn_mxu(ref, query, k):
    """The reference's knn_mxu on host arrays: ref (R x dim), query
    (dim x Q, dimension-major).  Returns (dist k x Q, index k x Q, seconds)."""
    ref = np.ascontiguousarray(ref, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    R, dim = ref.shape
    Q = query.shape[1]
    out_d = np.zeros((k, Q), np.float32)
    out_i = np.zeros((k, Q), np.int32)
    t = ctypes.c_double(0.0)
    ok = _bind().knn_mxu(ref, R, query, Q, dim, k, out_d, out_i, ctypes.byref(t))
    if not ok:
        raise RuntimeError("knn_mxu failed")
    return out_d, out_i, t.value


class KnnBF16:
    """Reusable precision-form kNN: workspace allocated once per shape."""

    def __init__(self, Q, R, dim, device="cuda"):
        torch = _torch()
        self.Q, self.R, self.dim = Q, R, dim
        nbytes = _bind().simd2_knn_bf16_workspace_size(Q, R, dim)
        self.ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
        self.stats = torch.zeros(3, dtype=torch.int32, device=device)

    def __call__(self, query, ref, k, sqrt=False, out=None, stream=None):
        torch = _torch()
        Q, dim = query.shape
        R = ref.shape[0]
        assert (Q, R, dim) == (self.Q, self.R, self.dim)
        if out is None:
            out = (torch.empty((Q, k), dtype=torch.float32, device=query.device),
                   torch.empty((Q, k), dtype=torch.int32, device=query.device))
        st = _bind().simd2_knn_bf16(_ptr(query), query.stride(0), _ptr(ref), ref.stride(0), Q, R,
                                    dim, k, _ptr(out[0]), _ptr(out[1]), SQRT if sqrt else 0,
                                    _ptr(self.ws), self.ws.numel(), _ptr(self.stats),
                                    _stream(stream))
        _check(st, "simd2_knn_bf16")
        return out


def knn_bf16(query, ref, k, sqrt=False, stream=None):
    """One-shot precision-form kNN (any dim, 1 <= k <= 256).  Returns
    (dist Q x k, index Q x k, stats) where stats = [queries that used FP64,
    candidates recomputed in FP64, queries whose candidates overflowed and
    were found by the exact scan of every reference]."""
    torch = _torch()
    runner = KnnBF16(query.shape[0], ref.shape[0], query.shape[1], device=query.device)
    d, i = runner(query, ref, k, sqrt=sqrt, stream=stream)
    if stream is not None:
        # The workspace and outputs were allocated on the current stream but
        # are used by kernels on `stream`: keep the caching allocator from
        # reusing them before that stream's work is done.
        s = stream if isinstance(stream, torch.cuda.Stream) else torch.cuda.ExternalStream(stream)
        for t in (runner.ws, runner.stats, d, i):
            t.record_stream(s)
    return d, i, runner.stats


def _main(argv=None):
    """Time the precision form on Q = R points from 64 Gaussian clusters (the
    PLD app's data): one untimed call per shape, then the mean over --reps
    event-timed calls on the current stream."""
    import argparse
    from .datasets import gaussian_clusters
    ap = argparse.ArgumentParser(prog="python -m simd2_amd.knn")
    ap.add_argument("--sizes", type=int, nargs="+", default=[1024, 2048, 4096, 8192, 16384])
    ap.add_argument("--k", type=int, nargs="+", default=[16])
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    torch = _torch()
    for n in args.sizes:
        ref, centers = gaussian_clusters(torch, n, args.dim, 64, seed=100)
        qry, _ = gaussian_clusters(torch, n, args.dim, 64, seed=101, centers=centers)
        runner = KnnBF16(n, n, args.dim)
        for k in args.k:
            runner.stats.zero_()
            runner(qry, ref, k)
            torch.cuda.synchronize()
            stats = runner.stats.cpu().tolist()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(args.reps):
                runner(qry, ref, k)
            e1.record()
            e1.synchronize()
            print(f"Q=R={n} d={args.dim} k={k}: {e0.elapsed_time(e1) / args.reps:.4f} ms per call;"
                  f" one call: {stats[0]} queries used FP64 ({stats[1]} candidates),"
                  f" {stats[2]} took the exact scan", flush=True)


if __name__ == "__main__":
    _main()
