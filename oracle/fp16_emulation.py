"""CPU restatement of the reference's FP16 "tensor core emulation" GEMM --
TEST INFRASTRUCTURE ONLY.

Follows mmo_gemm_compute_simple (kernel/tensor_srgemm.cuh:37-195): for every
output, k in order 0..K-1, the (x) is evaluated in half precision (IEEE
binary16 round-to-nearest-even; numpy's float16 arithmetic rounds each
operation once, exactly), converted to float, and folded into a float
accumulator that starts at __float2half(FLT_MAX) = +inf for min-based ops and
__float2half(FLT_MIN) = 0 for max-based ops (:52-86); the epilogue folds C
(:152-190).

Pinned by tests/test_oracle.py: with small-integer inputs every half
operation is exact and the result equals the FP32 host_srgemm restatement;
the max-based accumulator quirk; and a sum that half precision must round.
"""
from __future__ import annotations

import numpy as np

OPS = ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul",
       "min_max", "or_and", "minus_square"]


def tensor_srgemm(a16, b16, c, op, epilogue):
    op = OPS[op] if not isinstance(op, str) else op
    a = np.asarray(a16, np.float16)
    b = np.asarray(b16, np.float16)
    c = np.asarray(c, np.float32)
    m, k = a.shape
    n = b.shape[1]
    if op in ("min_plus", "min_mul", "min_max"):
        acc = np.full((m, n), np.inf, np.float32)
    else:
        acc = np.zeros((m, n), np.float32)
    for kk in range(k):
        av = a[:, kk][:, None]
        bv = b[kk][None, :]
        with np.errstate(over="ignore", invalid="ignore"):
            if op == "min_plus":
                acc = np.fmin(acc, (av + bv).astype(np.float32))
            elif op == "plus_mul":
                acc = acc + (av * bv).astype(np.float32)
            elif op == "max_plus":
                acc = np.fmax(acc, (av + bv).astype(np.float32))
            elif op == "max_min":
                acc = np.fmax(acc, np.minimum(av, bv).astype(np.float32))
            elif op == "max_mul":
                acc = np.fmax(acc, (av * bv).astype(np.float32))
            elif op == "min_mul":
                acc = np.fmin(acc, (av * bv).astype(np.float32))
            elif op == "min_max":
                acc = np.fmin(acc, np.maximum(av, bv).astype(np.float32))
            elif op == "or_and":
                acc = ((acc > 0) | ((av > 0) & (bv > 0))).astype(np.float32)
            else:
                d = (av - bv).astype(np.float16)
                acc = acc + (d * d).astype(np.float32)
    if not epilogue:
        return acc
    if op in ("min_plus", "min_mul", "min_max"):
        return np.fmin(acc, c)
    if op in ("max_plus", "max_min", "max_mul"):
        return np.fmax(acc, c)
    if op == "or_and":
        return ((np.trunc(acc) != 0) | (np.trunc(c) != 0)).astype(np.float32)
    return (acc + c).astype(np.float32)
