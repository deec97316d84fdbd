"""Batched repeated squaring over the C ABI (include/simd2/batch.h): many
independent graphs of uneven size solved in one hipGraph launch."""
from __future__ import annotations

import ctypes

import numpy as np

from . import CHECK, OP, OPS, Simd2Error, SUCCESS, _check, _ptr, _stream, _torch, lib


class _Graph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("op", ctypes.c_int), ("do_epilogue", ctypes.c_int),
                ("check_mode", ctypes.c_int)]


def _bind():
    L = lib()
    if getattr(L, "_batch_bound", False):
        return L
    i, p, ll = ctypes.c_int, ctypes.c_void_p, ctypes.c_int64
    L.simd2_batch_create.argtypes = [ctypes.POINTER(_Graph), i, i, ctypes.POINTER(p)]
    L.simd2_batch_destroy.argtypes = [p]
    L.simd2_batch_upload.argtypes = [p, i, p, ll, p]
    L.simd2_batch_download.argtypes = [p, i, p, ll, p]
    L.simd2_batch_buffer.argtypes = [p, i, ctypes.POINTER(p), ctypes.POINTER(ll)]
    L.simd2_batch_autotune.argtypes = [p, p]
    L.simd2_batch_solve.argtypes = [p, p]
    L.simd2_batch_iterations.argtypes = [p, p, ctypes.POINTER(i), p]
    L.simd2_batch_schedule.argtypes = [p, p, ctypes.POINTER(i)]
    for name in ("create", "destroy", "upload", "download", "buffer", "autotune", "solve",
                 "iterations", "schedule"):
        getattr(L, "simd2_batch_" + name).restype = i
    L._batch_bound = True
    return L


class Batch:
    """A batch of graphs.  ``specs`` is a list of (n, op, do_epilogue,
    check) with op a name or code and check "exact" / "threshold" / "aplp".
    FP32 graphs are n x n; or_and graphs are bit-packed (int32 words,
    ceil(n/32) per row)."""

    def __init__(self, specs, max_rounds=0):
        L = _bind()
        arr = (_Graph * len(specs))()
        self.specs = []
        for k, (n, op, epi, check) in enumerate(specs):
            code = OP[op] if isinstance(op, str) else int(op)
            arr[k] = _Graph(int(n), code, int(bool(epi)), CHECK[check])
            self.specs.append((int(n), code, bool(epi), check))
        h = ctypes.c_void_p()
        _check(L.simd2_batch_create(arr, len(specs), int(max_rounds), ctypes.byref(h)),
               "simd2_batch_create")
        self._h = h
        self.count = len(specs)

    def close(self):
        if getattr(self, "_h", None):
            lib().simd2_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _cols(self, g):
        n, op, _, _ = self.specs[g]
        return (n + 31) // 32 if OPS[op] == "or_and" else n

    def upload(self, g, X, stream=None):
        """Copy graph g in from a device tensor (n x n float32, or n x
        ceil(n/32) int32 words for or_and)."""
        if self.specs[g][0] == 0:
            return
        assert X.shape == (self.specs[g][0], self._cols(g)), (X.shape, self.specs[g])
        _check(_bind().simd2_batch_upload(self._h, g, _ptr(X), X.stride(0), _stream(stream)),
               "simd2_batch_upload")

    def download(self, g, out=None, stream=None):
        torch = _torch()
        n, op, _, _ = self.specs[g]
        dt = torch.int32 if OPS[op] == "or_and" else torch.float32
        if out is None:
            out = torch.empty((n, self._cols(g)), dtype=dt, device="cuda")
        if n:
            _check(_bind().simd2_batch_download(self._h, g, _ptr(out), out.stride(0),
                                                _stream(stream)), "simd2_batch_download")
        return out

    def autotune(self, stream=None):
        _check(_bind().simd2_batch_autotune(self._h, _stream(stream)), "simd2_batch_autotune")

    def solve(self, stream=None):
        _check(_bind().simd2_batch_solve(self._h, _stream(stream)), "simd2_batch_solve")

    def iterations(self, stream=None):
        """(iterations per graph as a numpy array, number not converged)."""
        it = np.zeros(max(self.count, 1), np.int32)
        bad = ctypes.c_int(0)
        _check(_bind().simd2_batch_iterations(self._h, it.ctypes.data, ctypes.byref(bad),
                                              _stream(stream)), "simd2_batch_iterations")
        return it[:self.count], bad.value

    def schedule(self):
        """(tile configuration per graph, launches per round)."""
        cfg = np.zeros(max(self.count, 1), np.int32)
        nl = ctypes.c_int(0)
        _check(_bind().simd2_batch_schedule(self._h, cfg.ctypes.data, ctypes.byref(nl)),
               "simd2_batch_schedule")
        return cfg[:self.count], nl.value


__all__ = ["Batch", "Simd2Error", "SUCCESS"]


def _main(argv=None):
    """python -m simd2_amd.batch [--graphs 4096] [--solves 3] [--tune]: the
    batched workload (n in [32, 512], p in [0.01, 0.2], 40/20/20/20 mix of
    min_plus / max_min / min_max / packed or_and), solved --solves times."""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=4096)
    ap.add_argument("--solves", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tune", action="store_true")
    args = ap.parse_args(argv)
    torch = _torch()
    from .graphs import app_graph, batch_specs
    specs = batch_specs(args.graphs, args.seed)
    mats = [app_graph(torch, op, n, p, s) for n, op, epi, check, p, s in specs]
    b = Batch([s[:4] for s in specs])
    if args.tune:
        for g, X in enumerate(mats):
            b.upload(g, X)
        b.autotune()
    times = []
    for _ in range(args.solves):
        for g, X in enumerate(mats):
            b.upload(g, X)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.solve()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    itr, bad = b.iterations()
    cfg, launches = b.schedule()
    print(json.dumps({"graphs": args.graphs, "solve_ms": [round(t, 3) for t in times],
                      "squarings": int(itr.sum()), "max_squarings": int(itr.max()),
                      "not_converged": bad, "launches_per_round": launches}), flush=True)


if __name__ == "__main__":
    _main()
