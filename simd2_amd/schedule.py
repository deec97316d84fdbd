"""Scheduled repeated squaring over the C ABI (include/simd2/schedule.h)."""
from __future__ import annotations

import ctypes

from . import CHECK, OP, OPS, Simd2Error, SUCCESS, ERROR_NOT_CONVERGED, _ptr, _stream, _torch, lib


class Schedule(ctypes.Structure):
    _fields_ = [("tile_config", ctypes.c_int), ("predicted_itr", ctypes.c_int),
                ("row_sparse", ctypes.c_int), ("changed_sparse", ctypes.c_int)]


class Report(ctypes.Structure):
    _fields_ = [("host_syncs", ctypes.c_int), ("squarings_enqueued", ctypes.c_int),
                ("tiles_total", ctypes.c_ulonglong), ("tiles_computed", ctypes.c_ulonglong),
                ("slab_tiles_total", ctypes.c_ulonglong),
                ("slab_tiles_computed", ctypes.c_ulonglong),
                ("rows_sparse", ctypes.c_ulonglong), ("row_entries", ctypes.c_ulonglong),
                ("changed_rows", ctypes.c_ulonglong), ("changed_entries", ctypes.c_ulonglong)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _bind():
    L = lib()
    if getattr(L, "_sched_bound", False):
        return L
    i, p = ctypes.c_int, ctypes.c_void_p
    L.simd2_converge_scheduled_workspace.argtypes = [i, i, i]
    L.simd2_converge_scheduled_workspace.restype = ctypes.c_size_t
    L.simd2_converge_scheduled.argtypes = [i, i, p, i, i, i, i, ctypes.POINTER(Schedule), p, p,
                                           ctypes.c_size_t, p, ctypes.POINTER(i),
                                           ctypes.POINTER(Report)]
    L.simd2_converge_scheduled.restype = i
    L._sched_bound = True
    return L


def workspace_bytes(n, tile_config=-1, max_itr=64):
    return int(_bind().simd2_converge_scheduled_workspace(n, tile_config, max_itr))


def converge_scheduled(op, X, epilogue=True, check="threshold", max_itr=64, tile_config=-1,
                       predicted_itr=0, perm=None, workspace=None, stream=None, row_sparse=-1,
                       changed_sparse=-1):
    """Scheduled converge loop on the n x n float32 device tensor X, in place.
    Returns (X, num_itr, status, report dict)."""
    torch = _torch()
    L = _bind()
    n = X.shape[0]
    opname = op if isinstance(op, str) else OPS[op]
    need = L.simd2_converge_scheduled_workspace(n, tile_config, max_itr)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(max(int(need), 1), dtype=torch.uint8, device=X.device)
    if perm is not None:
        perm = perm.to(device=X.device, dtype=torch.int32).contiguous()
    itr = ctypes.c_int(0)
    rep = Report()
    sc = Schedule(int(tile_config), int(predicted_itr), int(row_sparse), int(changed_sparse))
    st = L.simd2_converge_scheduled(OP[opname], n, _ptr(X), X.stride(0), int(bool(epilogue)),
                                    CHECK[check], max_itr, ctypes.byref(sc), _ptr(perm),
                                    _ptr(workspace), workspace.numel(), _stream(stream),
                                    ctypes.byref(itr), ctypes.byref(rep))
    if st not in (SUCCESS, ERROR_NOT_CONVERGED):
        raise Simd2Error(st, "simd2_converge_scheduled")
    return X, itr.value, st, rep.as_dict()
