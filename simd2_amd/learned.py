"""The learned schedule over the C ABI (include/simd2/learned.h): graph
features, vertex order, the MLP's kernels and the learned converge loop."""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import (CHECK, OP, OPS, ERROR_NOT_CONVERGED, SUCCESS, Simd2Error, _check, _ptr, _stream,
               _torch, lib)
from .schedule import Report, Schedule, _bind as _bind_sched

NF, NH, NO = 24, 64, 2
NPARAM = NH * NF + NH + NH * NH + NH + NO * NH + NO
MODEL_FLOATS = NPARAM + 2 * NF + 2 * NO
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedule_model.json")
# tile class (ordered by tile size) -> simd2_srgemm_ex tile configuration
TILE_OF_CLASS = [1, 2, 0]
CLASS_OF_TILE = {1: 0, 2: 1, 0: 2}


def _bind():
    L = lib()
    if getattr(L, "_learned_bound", False):
        return L
    _bind_sched()
    i, p, sz, f = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float
    L.simd2_graph_features_workspace.argtypes = [i]
    L.simd2_graph_features_workspace.restype = sz
    L.simd2_graph_features.argtypes = [i, i, p, i, i, p, p, sz, p]
    L.simd2_graph_order.argtypes = [i, i, p, i, p, p, sz, p]
    L.simd2_mlp_forward.argtypes = [p, p, i, p, p]
    L.simd2_mlp_workspace.argtypes = [i]
    L.simd2_mlp_workspace.restype = sz
    L.simd2_mlp_gradient.argtypes = [p, p, p, p, i, p, p, p, sz, p]
    L.simd2_mlp_train_step.argtypes = [p, p, p, i, f, p, p, p, i, p, p, sz, p]
    L.simd2_converge_learned_workspace.argtypes = [i, i]
    L.simd2_converge_learned_workspace.restype = sz
    L.simd2_converge_learned.argtypes = [i, i, p, i, i, i, i, p, p, sz, p, ctypes.POINTER(i),
                                         ctypes.POINTER(Schedule), ctypes.POINTER(Report)]
    for name in ("graph_features", "graph_order", "mlp_forward", "mlp_gradient",
                 "mlp_train_step", "converge_learned"):
        getattr(L, "simd2_" + name).restype = i
    L._learned_bound = True
    return L


def _ws(torch, nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def graph_features(op, X, epilogue=True, workspace=None, stream=None):
    """The 24 features of the n x n float32 adjacency X (device tensor)."""
    torch = _torch()
    L = _bind()
    n = X.shape[0]
    need = L.simd2_graph_features_workspace(n)
    if workspace is None or workspace.numel() < need:
        workspace = _ws(torch, need, X.device)
    f = torch.empty(NF, dtype=torch.float32, device=X.device)
    _check(L.simd2_graph_features(OP[op], n, _ptr(X), X.stride(0), int(bool(epilogue)), _ptr(f),
                                  _ptr(workspace), workspace.numel(), _stream(stream)),
           "simd2_graph_features")
    return f


def graph_order(op, X, stream=None):
    torch = _torch()
    L = _bind()
    n = X.shape[0]
    ws = _ws(torch, L.simd2_graph_features_workspace(n), X.device)
    perm = torch.empty(n, dtype=torch.int32, device=X.device)
    _check(L.simd2_graph_order(OP[op], n, _ptr(X), X.stride(0), _ptr(perm), _ptr(ws), ws.numel(),
                               _stream(stream)), "simd2_graph_order")
    return perm


def mlp_forward(params, x, stream=None):
    torch = _torch()
    B = x.shape[0]
    y = torch.empty((B, NO), dtype=torch.float32, device=x.device)
    _check(_bind().simd2_mlp_forward(_ptr(params), _ptr(x), B, _ptr(y), _stream(stream)),
           "simd2_mlp_forward")
    return y


def mlp_gradient(params, x, t, w, stream=None):
    """(loss, gradient) of the weighted squared error, on the GPU."""
    torch = _torch()
    L = _bind()
    B = x.shape[0]
    ws = _ws(torch, L.simd2_mlp_workspace(B), x.device)
    grad = torch.empty(NPARAM, dtype=torch.float32, device=x.device)
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    _check(L.simd2_mlp_gradient(_ptr(params), _ptr(x), _ptr(t), _ptr(w), B, _ptr(grad),
                                _ptr(loss), _ptr(ws), ws.numel(), _stream(stream)),
           "simd2_mlp_gradient")
    return loss, grad


class Trainer:
    """Adam on the GPU through simd2_mlp_train_step."""

    def __init__(self, params, lr=1e-3, max_batch=4096):
        torch = _torch()
        self.P = params
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.lr = lr
        self.step_no = 0
        self.ws = _ws(torch, _bind().simd2_mlp_workspace(max_batch), params.device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=params.device)

    def step(self, x, t, w, stream=None):
        self.step_no += 1
        _check(_bind().simd2_mlp_train_step(_ptr(self.P), _ptr(self.m), _ptr(self.v),
                                            self.step_no, float(self.lr), _ptr(x), _ptr(t),
                                            _ptr(w), x.shape[0], _ptr(self.loss), _ptr(self.ws),
                                            self.ws.numel(), _stream(stream)),
               "simd2_mlp_train_step")
        return self.loss


def load_model(path=MODEL_PATH, device="cuda"):
    """Model vector (params, feature mean / scale, output scale / offset) on
    the device, from the JSON file the trainer writes."""
    torch = _torch()
    with open(path) as f:
        m = json.load(f)
    vec = np.concatenate([np.asarray(m["params"], np.float32),
                          np.asarray(m["feature_mean"], np.float32),
                          np.asarray(m["feature_scale"], np.float32),
                          np.asarray(m["output_scale"], np.float32),
                          np.asarray(m["output_offset"], np.float32)])
    assert vec.size == MODEL_FLOATS
    return torch.from_numpy(vec).to(device), m


def converge_learned(op, X, model, epilogue=True, check="threshold", max_itr=64,
                     workspace=None, stream=None):
    """Learned-schedule converge loop on X in place.  Returns (X, num_itr,
    status, schedule (tile_config, predicted_itr), report dict)."""
    torch = _torch()
    L = _bind()
    n = X.shape[0]
    need = L.simd2_converge_learned_workspace(n, max_itr)
    if workspace is None or workspace.numel() < need:
        workspace = _ws(torch, need, X.device)
    itr = ctypes.c_int(0)
    used = Schedule()
    rep = Report()
    st = L.simd2_converge_learned(OP[op], n, _ptr(X), X.stride(0), int(bool(epilogue)),
                                  CHECK[check], max_itr, _ptr(model), _ptr(workspace),
                                  workspace.numel(), _stream(stream), ctypes.byref(itr),
                                  ctypes.byref(used), ctypes.byref(rep))
    if st not in (SUCCESS, ERROR_NOT_CONVERGED):
        raise Simd2Error(st, "simd2_converge_learned")
    return X, itr.value, st, (used.tile_config, used.predicted_itr), rep.as_dict()
