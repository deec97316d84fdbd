"""Benchmark of the semiring-GEMM path on one MI355X.

Workload (one "step"): the single-srgemm suite M = N = K = 4096 for all nine
operators (float inputs integer-valued U{1..100} with 10% +inf / -inf holes for
min- / max-based ops, U(0,1] for max_mul / min_mul, no holes for plus_mul /
minus_square; or_and on bit-packed Bernoulli(0.02) matrices), inputs resident
in HBM before timing.  `value` is the step's aggregate semiring throughput,
sum over ops of 2*M*N*K / step time, in GOPS.

Also reported: per-op GOPS and fraction of the FP32 vector peak, min_plus at
8192^3, the n=1024 APSP converge loop (reference loop structure and fused
check), the roofline of the dominant kernel (min_plus, timed with HIP events on
the stream it is launched on), and the CPU oracle timed on this host on a
row-sampled slice of the same min_plus workload.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_VECTOR_TFLOPS = 157.3   # MI355X published FP32 vector peak (FMA, packed)
# Issue bound of the semiring inner step with packed FP32 disabled: 256 CUs x
# 4 SIMD x 16 lanes x 2.4 GHz = 39.3 T lane-instr/s, 3 instructions per 4 ops.
ISA_BOUND_TOPS = 39.3216 * 4.0 / 3.0
OPS = ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul", "min_max",
       "or_and", "minus_square"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(torch, op, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if op in ("max_mul", "min_mul"):
        A = 1.0 - torch.rand((n, n), generator=g, device="cuda")
        B = 1.0 - torch.rand((n, n), generator=g, device="cuda")
        C = 1.0 - torch.rand((n, n), generator=g, device="cuda")
        return A, B, C
    A = torch.randint(1, 101, (n, n), generator=g, device="cuda").float()
    B = torch.randint(1, 101, (n, n), generator=g, device="cuda").float()
    C = torch.randint(1, 101, (n, n), generator=g, device="cuda").float()
    if op not in ("plus_mul", "minus_square"):
        hole = float("inf") if op.startswith("min") else float("-inf")
        A[torch.rand((n, n), generator=g, device="cuda") < 0.1] = hole
        B[torch.rand((n, n), generator=g, device="cuda") < 0.1] = hole
    return A, B, C


def make_call(torch, simd2_amd, op, n, seed):
    """Returns a zero-argument launcher for one srgemm of `op` at n^3."""
    stream = torch.cuda.current_stream().cuda_stream
    L = simd2_amd.lib()
    if op == "or_and":
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        bits = lambda: (torch.rand((n, n), generator=g, device="cuda") < 0.02).to(torch.int32)
        A, B, C = (simd2_amd.pack_bits(bits()) for _ in range(3))
        D = torch.empty_like(C)
        nw = A.shape[1]
        args = (n, n, n, A.data_ptr(), nw, B.data_ptr(), nw, C.data_ptr(), nw, D.data_ptr(), nw,
                1, stream)
        keep = (A, B, C, D)

        def call():
            st = L.simd2_orand_packed(*args)
            if st:
                raise RuntimeError(f"or_and status {st}")
        return call, keep
    A, B, C = make_inputs(torch, op, n, seed)
    D = torch.empty_like(C)
    fn = getattr(L, simd2_amd.DROPIN[op])
    # row-major A @ B as the column-major call (N, M, K, B, N, A, K, C, N, D)
    args = (n, n, n, B.data_ptr(), n, A.data_ptr(), n, C.data_ptr(), n, D.data_ptr(), True,
            stream)

    def call():
        st = fn(*args)
        if st:
            raise RuntimeError(f"{op} status {st}")
    return call, (A, B, C, D)


def time_launches(torch, call, reps):
    """Average device time per launch, HIP events on the launching stream."""
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        call()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def cpu_baseline(n=4096):
    """The CPU oracle on this host: min_plus rows of the 4096^3 workload."""
    import oracle
    oracle.build()
    rng = np.random.default_rng(0)
    A = rng.integers(1, 101, (n, n)).astype(np.float32)
    B = rng.integers(1, 101, (n, n)).astype(np.float32)
    A[rng.random((n, n)) < 0.1] = np.inf
    B[rng.random((n, n)) < 0.1] = np.inf
    C = rng.integers(1, 101, (n, n)).astype(np.float32)
    cores = len(os.sched_getaffinity(0))
    out = {}
    # faithful: single thread, function-pointer ops, i-j-k order
    rows1 = 16
    if oracle.ref_available():
        D = np.full((rows1, n), np.inf, np.float32)
        t0 = time.perf_counter()
        oracle.ref_srgemm_f32("min_plus", A[:rows1], B, C[:rows1], D, True)
        t1 = time.perf_counter() - t0
        kind = "reference"
    else:
        D = np.full((n, n), np.inf, np.float32)
        t0 = time.perf_counter()
        oracle.srgemm_f32_rows("min_plus", A, B, C, D, True, 0, rows1, 1)
        t1 = time.perf_counter() - t0
        kind = "port"
    out["single"] = {"value": 2.0 * rows1 * n * n / t1 / 1e9, "unit": "GOPS", "cores": 1,
                     "kind": kind, "sample": f"min_plus rows 0..{rows1 - 1} of 4096^3 "
                     f"(row-sampled, {t1:.1f} s)"}
    rows = 16 * cores
    D = np.full((n, n), np.inf, np.float32)
    t0 = time.perf_counter()
    oracle.srgemm_f32_rows("min_plus", A, B, C, D, True, 0, rows, cores)
    t2 = time.perf_counter() - t0
    out["openmp"] = {"value": 2.0 * rows * n * n / t2 / 1e9, "unit": "GOPS", "cores": cores,
                     "kind": "port", "sample": f"min_plus rows 0..{rows - 1} of 4096^3 "
                     f"(row-sampled, OpenMP over rows, bit-identical, {t2:.1f} s)"}
    return out


def apsp_1024(torch, simd2_amd):
    """n=1024 ER p=0.05 seed 0 APSP: converge-loop wall time per solve."""
    from simd2_amd.graphs import er_graph
    X0 = er_graph(torch, 1024, 0.05, seed=0)
    res = {}
    for name, flags in (("reference_loop", simd2_amd.CONVERGE_REFERENCE),
                        ("fused_check", simd2_amd.CONVERGE_FUSED)):
        times, itr = [], 0
        for _ in range(4):
            X = X0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, itr, st = simd2_amd.converge("min_plus", X, epilogue=True, check="threshold",
                                            flags=flags)
            times.append((time.perf_counter() - t0) * 1e3)
        res[name] = {"ms": float(np.median(times[1:])), "squarings": itr}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip 8192^3 and APSP extras")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the semiring GEMM path runs on a single "
              f"MI355X (asked for --gpus {args.gpus}); multi-GPU sharding is not built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import simd2_amd
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    simd2_amd.lib()
    n = args.n
    info = simd2_amd.device_info()
    log(f"device {info}")

    calls = {}
    for i, op in enumerate(OPS):
        calls[op] = make_call(torch, simd2_amd, op, n, seed=1000 + i)
    # Traffic calibration for the rocprofv3 FETCH_SIZE / WRITE_SIZE passes: one
    # plain streaming kernel (torch.neg) of exactly 48 MiB in and 48 MiB out.
    cal = torch.ones(12 * 1024 * 1024, device="cuda")
    torch.neg(cal, out=torch.empty_like(cal))
    del cal
    torch.cuda.synchronize()

    def step():
        for op in OPS:
            calls[op][0]()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.record()
    for _ in range(args.steps):
        step()
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = s.elapsed_time(e) / args.steps
    ops_per_step = len(OPS) * 2.0 * n ** 3
    value = ops_per_step / (ms_step * 1e-3) / 1e9
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step, wall {wall:.2f} s")

    per_op = {}
    for op in OPS:
        ms = time_launches(torch, calls[op][0], max(5, args.steps))
        tops = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
        per_op[op] = {"ms": round(ms, 4), "GOPS": round(tops * 1e3, 1),
                      "frac_fp32_vector_peak": round(tops / PEAK_FP32_VECTOR_TFLOPS, 4)}
        log(f"{op}: {ms:.3f} ms  {tops:.2f} T ops/s")

    mp = per_op["min_plus"]
    mp_tops = mp["GOPS"] / 1e3
    roofline = {"kernel": "srgemm_valu_kernel<MinPlus,128x128x16>", "bound": "valu",
                "achieved": round(mp_tops, 3), "peak": PEAK_FP32_VECTOR_TFLOPS,
                "unit": "TFLOP/s", "frac": round(mp_tops / PEAK_FP32_VECTOR_TFLOPS, 4),
                "isa_bound": round(ISA_BOUND_TOPS, 2),
                "frac_isa_bound": round(mp_tops / ISA_BOUND_TOPS, 4),
                "traffic": None,
                "note": "achieved = 2*4096^3 semiring ops (one (x) and one (+) per i,j,k) / "
                        "mean launch time; unit is T semiring ops/s"}
    calls.clear()
    torch.cuda.empty_cache()

    extras = {}
    if not args.no_extras:
        call, keep = make_call(torch, simd2_amd, "min_plus", 8192, seed=77)
        call()
        torch.cuda.synchronize()
        ms = time_launches(torch, call, 5)
        tops = 2.0 * 8192 ** 3 / (ms * 1e-3) / 1e12
        extras["min_plus_8192"] = {"ms": round(ms, 3), "GOPS": round(tops * 1e3, 1),
                                   "frac_fp32_vector_peak": round(tops / PEAK_FP32_VECTOR_TFLOPS, 4),
                                   "frac_isa_bound": round(tops / ISA_BOUND_TOPS, 4)}
        log(f"min_plus 8192^3: {ms:.3f} ms {tops:.2f} T ops/s")
        del keep, call
        torch.cuda.empty_cache()
        extras["apsp_er1024"] = apsp_1024(torch, simd2_amd)
        log(f"apsp n=1024: {extras['apsp_er1024']}")

    cpu = None
    if not args.no_cpu:
        base = cpu_baseline()
        cpu = dict(base["single"])
        cpu["openmp"] = base["openmp"]
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "semiring GOPS & % VALU peak; APSP n=8192 ms learned vs plain (bit-exact)",
        "value": round(value, 1),
        "unit": "GOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"srgemm M=N=K={n}, all nine semirings (one launch each per step)",
                   "M": n, "N": n, "K": n, "epilogue": True,
                   "or_and": "bit-packed uint32, density 0.02"},
        "per_op": per_op,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
