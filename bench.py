"""Benchmark of the semiring-GEMM path on one MI355X.

Workload (one "step"): the single-srgemm suite M = N = K = 4096 for all nine
operators (float inputs integer-valued U{1..100} with 10% +inf / -inf holes for
min- / max-based ops, U(0,1] for max_mul / min_mul, no holes for plus_mul /
minus_square; or_and on bit-packed Bernoulli(0.02) matrices), inputs resident
in HBM before timing.  `value` is the step's aggregate semiring throughput,
sum over ops of 2*M*N*K / step time, in GOPS.

minus_square runs through simd2_minus_square_exact (workspace allocated before
timing): on this integer data the int8 matrix cores compute |a|^2 + |b|^2 -
2 a.b, which is exact and bit-identical to the k-ordered FP32 sum wherever the
sums stay at or below 2^24 (all of them here); its per-op entry also gives the
drop-in's (direct FP32 kernel) time and the fraction of the int8 MFMA peak.

Also reported: per-op GOPS and fraction of the FP32 vector peak, min_plus at
8192^3, the n=1024 APSP converge loop (reference loop structure and fused
check), the roofline of the dominant kernel (min_plus, timed with HIP events on
the stream it is launched on; per-op times are the median of 20 back-to-back
launches), and the CPU oracle timed on this host on a row-sampled slice of the
same min_plus workload, plus its full n=1024 APSP loop on all host cores,
checked bit for bit against the GPU's.

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
PEAK_INT8_MFMA_TOPS = 5000.0      # MI355X published dense int8 MFMA peak
# VALU issue ceiling of the min_plus inner step (2 x v_add_f32 + 1 x v_min3_f32
# per 4 semiring ops; packed FP32 is not used).  gfx950 dual-issues the adds but
# not v_min3: a microbenchmark of that mix issues at most 1.60 wave-instructions
# per CU per cycle, so the ceiling is 1.60 x 64 lanes x 4/3 ops x 256 CUs x 2.4 GHz.
ISSUE_CEILING_TOPS = 1.60 * 64 * 4.0 / 3.0 * 256 * 2.4e9 / 1e12
# The same measured rates for the other VALU mixes: min/max-only (2 x v_min +
# v_max3 per 4 ops) issues at most 0.90 wave-instructions per CU per cycle;
# the direct minus_square (v_sub + v_fmac per term, 2 ops) at 1.60.
ISSUE_CEILING_BY_OP = {
    "min_plus": ISSUE_CEILING_TOPS, "max_plus": ISSUE_CEILING_TOPS,
    "min_mul": ISSUE_CEILING_TOPS, "max_mul": ISSUE_CEILING_TOPS,
    "max_min": 0.90 * 64 * 4.0 / 3.0 * 256 * 2.4e9 / 1e12,
    "min_max": 0.90 * 64 * 4.0 / 3.0 * 256 * 2.4e9 / 1e12,
}
# Bytes moved past L2 per launch of the min_plus kernel at n = 4096 (128x64
# tiles): rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, the factors a 48 MiB
# streaming copy reads on the same counters (profiles/srgemm_4096_counters.md):
# 1199 MiB fetched + 64 MiB written.
TRAFFIC = {4096: int((1199 + 64) * 2 ** 20)}
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
    if op != "minus_square":
        return call, (A, B, C, D)
    # minus_square: the exact int8 entry point, its workspace allocated here,
    # outside any timed region; the drop-in (the direct FP32 kernel) into its
    # own output beside it
    ws = torch.empty(simd2_amd.minus_square_workspace(n, n, n), dtype=torch.uint8,
                     device="cuda")
    D2 = torch.empty_like(C)
    args_x = (n, n, n, B.data_ptr(), n, A.data_ptr(), n, C.data_ptr(), n, D.data_ptr(), n, 1,
              ws.data_ptr(), ws.numel(), stream)
    args_d = (n, n, n, B.data_ptr(), n, A.data_ptr(), n, C.data_ptr(), n, D2.data_ptr(), True,
              stream)

    def exact():
        st = L.simd2_minus_square_exact(*args_x)
        if st:
            raise RuntimeError(f"{op} exact status {st}")

    def direct():
        st = fn(*args_d)
        if st:
            raise RuntimeError(f"{op} direct status {st}")
    return exact, (A, B, C, D, D2, ws), direct


def time_launches_median(torch, call, reps=20):
    """Median device time of `reps` back-to-back launches, each bracketed by
    HIP events on the launching stream (torch's current stream)."""
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record()
    for i in range(reps):
        call()
        ev[i + 1].record()
    ev[-1].synchronize()
    return float(np.median([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_launches(torch, call, reps):
    """Average device time per launch, HIP events on the launching stream."""
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        call()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def cpu_baseline(n=4096, apsp=None):
    """The CPU oracle on this host: min_plus rows of the 4096^3 workload, and
    (when `apsp` = (X0, X_gpu, squarings) of the n=1024 APSP run is given)
    the full converge loop of that graph on all host cores, checked against
    the GPU's matrix bit for bit."""
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
    out["host_cpu"] = host_cpu_model()
    if apsp is not None:
        X0, Xg, itr_g = apsp
        t0 = time.perf_counter()
        Xc, itr_c = oracle.converge_f32("min_plus", X0, epilogue=True, check="threshold",
                                        threads=cores)
        t3 = time.perf_counter() - t0
        out["apsp_er1024_loop"] = {
            "ms": round(t3 * 1e3, 1), "cores": cores, "kind": "port", "squarings": itr_c,
            "identical_to_gpu": bool(itr_c == itr_g and np.array_equal(
                Xc.view(np.uint32), np.ascontiguousarray(Xg).view(np.uint32)))}
    return out


def apsp_1024(torch, simd2_amd, keep=None):
    """n=1024 ER p=0.05 seed 0 APSP: converge-loop wall time per solve.
    `keep` (a list) receives (X0, final matrix, squarings) on the host."""
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
    if keep is not None:
        keep.append((X0.cpu().numpy(), X.cpu().numpy(), itr))
    return res


def batched_graphs(torch, simd2_amd, count=4096, seed=0):
    """Batched workload: `count` Erdos-Renyi graphs, n uniform in [32, 512],
    p uniform in [0.01, 0.2], integer weights in [1, 100]; 40% min_plus APSP,
    20% max_min MCP, 20% min_max MST (no epilogue), 20% or_and transitive
    closure on packed bits.  Times one hipGraph solve of the whole batch
    (default schedule, then autotuned) against the per-graph converge loop,
    and checks every graph bit for bit against that loop."""
    from simd2_amd.batch import Batch
    from simd2_amd.graphs import app_graph, batch_specs
    specs = batch_specs(count, seed)
    mats = [app_graph(torch, op, n, p, s) for n, op, epi, check, p, s in specs]
    torch.cuda.synchronize()
    b = Batch([s[:4] for s in specs])
    stream = torch.cuda.current_stream()

    def upload():
        for g, X in enumerate(mats):
            b.upload(g, X)

    def timed_solve():
        upload()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        b.solve()
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e)

    timed_solve()  # first launch of the graph
    default_ms = min(timed_solve() for _ in range(2))
    upload()
    t0 = time.perf_counter()
    b.autotune()
    tune_s = time.perf_counter() - t0
    tuned = [timed_solve() for _ in range(3)]
    itr, bad = b.iterations()
    cfg, launches = b.schedule()
    outs = [b.download(g) for g in range(count)]

    # the per-graph loop (fused change test, one flag read per squaring)
    work = {}
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    singles, same_itr = [], 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (n, op, epi, check, p, s), X in zip(specs, mats):
        Y = X.clone()
        key = (tuple(Y.shape), Y.dtype)
        if key not in work:
            work[key] = torch.empty_like(Y)
        _, it1, st = simd2_amd.converge(op, Y, epilogue=epi, check=check, work=work[key],
                                        flag=flag)
        singles.append((Y, it1))
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) * 1e3
    identical = 0
    for g, (Y, it1) in enumerate(singles):
        same_itr += int(itr[g] == it1)
        identical += int(torch.equal(outs[g].view(torch.int32), Y.view(torch.int32)))
    ops = 0.0
    for (n, op, epi, check, p, s), it in zip(specs, itr):
        if op != "or_and":
            ops += 2.0 * n ** 3 * int(it)
    ms = float(np.median(tuned))
    hist = {}
    for (n, op, *_), c in zip(specs, cfg):
        if c >= 0:
            name = {0: "128x128", 1: "64x64", 2: "128x64"}[int(c)]
            hist[name] = hist.get(name, 0) + 1
    b.close()
    return {"graphs": count, "squarings": int(itr.sum()), "max_squarings": int(itr.max()),
            "not_converged": bad, "ms_default_schedule": round(default_ms, 3),
            "ms_autotuned": round(ms, 3), "autotune_s": round(tune_s, 2),
            "ms_per_graph_loop": round(loop_ms, 1),
            "speedup_vs_loop": round(loop_ms / ms, 1),
            "fp32_GOPS": round(ops / (ms * 1e-3) / 1e9, 1),
            "frac_issue_ceiling": round(ops / (ms * 1e-3) / 1e12 / ISSUE_CEILING_TOPS, 4),
            "launches_per_round": launches, "tile_configs": hist,
            "identical_to_loop": identical, "same_iterations": same_itr}


def learned_schedule(torch, count, n=8192):
    """Learned-schedule APSP / MCP at n = 8192 on the first `count` held-out
    graphs (ER / grid / geometric in turn): plain converge loop vs the learned
    loop vs the scheduled loop without the network, every result checked bit
    for bit against the plain loop."""
    import os
    from simd2_amd.eval_schedule import evaluate, summarize
    from simd2_amd.learned import MODEL_PATH
    if not os.path.exists(MODEL_PATH):
        return {"error": f"no trained model at {MODEL_PATH}"}
    rows = evaluate(torch, count, n)
    s = summarize(rows)
    s["n"] = n
    return s


def pld_knn(torch, simd2_amd, Q=16384, R=16384, d=128, k=16, sample=512):
    """PLD/kNN workload: 64 Gaussian clusters (centres ~N(0, 10^2), sigma 1),
    Q queries x R references, d = 128, k = 16.  Times the plain FP32 direct path
    (minus_square srgemm into the Q x R matrix + the reference top-k) and the
    precision form (BF16-MFMA candidates, FP32 re-rank, FP64 near ties), and
    checks a sample of queries against an FP64 host computation."""
    from simd2_amd.datasets import gaussian_clusters
    from simd2_amd.knn import KnnBF16, knn_select
    ref, centers = gaussian_clusters(torch, R, d, 64, seed=100)
    qry, _ = gaussian_clusters(torch, Q, d, 64, seed=101, centers=centers)
    qT = qry.t().contiguous()
    D = torch.empty((R, Q), device="cuda")
    runner = KnnBF16(Q, R, d)

    def direct():
        simd2_amd.srgemm("minus_square", ref, qT, out=D)
        return knn_select(D, k, sqrt=False)

    def precise():
        return runner(qry, ref, k)

    res = {}
    outs = {}
    for name, fn in (("direct_fp32", direct), ("bf16_mfma_rerank", precise)):
        outs[name] = fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            fn()
        e1.record()
        e1.synchronize()
        res[name] = {"ms": round(e0.elapsed_time(e1) / 5, 4)}
    # FP64 truth on a sample of queries: shortlist by the FP64 expansion, then
    # the direct FP64 sum, ordered by (distance, index).
    rows = np.linspace(0, Q - 1, sample).astype(np.int64)
    Q64 = qry[torch.from_numpy(rows).cuda()].double().cpu().numpy()
    R64 = ref.double().cpu().numpy()
    Dx = (Q64 ** 2).sum(1)[:, None] + (R64 ** 2).sum(1)[None, :] - 2.0 * (Q64 @ R64.T)
    cand = np.argpartition(Dx, k + 8, axis=1)[:, :k + 8]
    want_i = np.zeros((sample, k), np.int64)
    want_d = np.zeros((sample, k))
    for r, c in enumerate(cand):
        dd = ((R64[c] - Q64[r]) ** 2).sum(1)
        o = np.lexsort((c, dd))[:k]
        want_i[r], want_d[r] = c[o], dd[o]
    for name in outs:
        dist, idx = outs[name]
        if name == "direct_fp32":
            dist, idx = dist.t(), idx.t()
        gd = dist.cpu().numpy()[rows].astype(np.float64)
        gi = idx.cpu().numpy()[rows]
        res[name]["exact_topk_agreement"] = float(np.mean(np.all(gi == want_i, axis=1)))
        res[name]["max_abs_err"] = float(np.max(np.abs(gd - want_d)))
        res[name]["max_rel_err"] = float(np.max(np.abs(gd - want_d) / want_d))
    st = runner.stats.cpu().tolist()
    res["bf16_mfma_rerank"]["fp64_queries"] = st[0]
    res["bf16_mfma_rerank"]["fp64_candidates"] = st[1]
    res["config"] = {"Q": Q, "R": R, "d": d, "k": k, "clusters": 64,
                     "accuracy_sample_queries": sample}
    res["speedup_precision_form"] = round(res["direct_fp32"]["ms"] /
                                          res["bf16_mfma_rerank"]["ms"], 2)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip 8192^3 and APSP extras")
    ap.add_argument("--learned-graphs", type=int, default=12,
                    help="held-out n=8192 graphs for the learned-schedule extra")
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
        ms = time_launches_median(torch, calls[op][0], 20)
        tops = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
        per_op[op] = {"ms": round(ms, 4), "GOPS": round(tops * 1e3, 1),
                      "frac_fp32_vector_peak": round(tops / PEAK_FP32_VECTOR_TFLOPS, 4)}
        log(f"{op}: {ms:.3f} ms  {tops:.2f} T ops/s")
    # minus_square: which tiles took the int8 path, and the direct kernel beside it
    mf, va = simd2_amd.minus_square_tiles(calls["minus_square"][1][-1], n, n, n)
    ms_dir = time_launches(torch, calls["minus_square"][2], max(3, args.steps // 2))
    msq = per_op["minus_square"]
    msq.update({"path": "simd2_minus_square_exact: int8 MFMA, exact (|a|^2+|b|^2-2a.b, "
                        "sums <= 2^24), caller's workspace",
                "int8_tiles": mf, "direct_tiles": va,
                "frac_int8_mfma_peak": round(msq["GOPS"] / 1e3 / PEAK_INT8_MFMA_TOPS, 4),
                "dropin_direct_fp32_ms": round(ms_dir, 4),
                "speedup_vs_direct": round(ms_dir / msq["ms"], 2)})
    log(f"minus_square int8 tiles {mf}, direct tiles {va}, direct kernel {ms_dir:.3f} ms")

    mp = per_op["min_plus"]
    mp_tops = mp["GOPS"] / 1e3
    tile = {0: "128x128", 1: "64x64", 2: "128x64"}[
        simd2_amd.lib().simd2_pick_tile_config(0, n, n, n)]
    roofline = {"kernel": f"srgemm_valu_kernel<MinPlus,{tile}x16>", "bound": "valu",
                "achieved": round(mp_tops, 3), "peak": PEAK_FP32_VECTOR_TFLOPS,
                "unit": "TFLOP/s", "frac": round(mp_tops / PEAK_FP32_VECTOR_TFLOPS, 4),
                "issue_ceiling": round(ISSUE_CEILING_TOPS, 2),
                "frac_issue_ceiling": round(mp_tops / ISSUE_CEILING_TOPS, 4),
                "traffic": TRAFFIC.get(n),
                "note": "achieved = 2*n^3 semiring ops (one (x) and one (+) per i,j,k) / "
                        "median launch time on the launching stream, in T semiring ops/s; "
                        "issue_ceiling = measured VALU issue rate of the kernel's "
                        "instruction mix (2 dual-issuable + 1 single-issue per 4 ops)"}
    # the longest kernel of the step (min_plus is the one the target names)
    lop = max(per_op, key=lambda o: per_op[o]["ms"])
    ltops = per_op[lop]["GOPS"] / 1e3
    roofline["longest_kernel"] = {
        "op": lop, "ms": per_op[lop]["ms"], "achieved": round(ltops, 3),
        "frac": round(ltops / PEAK_FP32_VECTOR_TFLOPS, 4),
        "issue_ceiling": round(ISSUE_CEILING_BY_OP[lop], 2) if lop in ISSUE_CEILING_BY_OP
        else None,
        "frac_issue_ceiling": round(ltops / ISSUE_CEILING_BY_OP[lop], 4)
        if lop in ISSUE_CEILING_BY_OP else None}
    calls.clear()
    torch.cuda.empty_cache()

    extras = {}
    apsp_keep = []
    if not args.no_extras:
        call, keep = make_call(torch, simd2_amd, "min_plus", 8192, seed=77)
        call()
        torch.cuda.synchronize()
        ms = time_launches(torch, call, 5)
        tops = 2.0 * 8192 ** 3 / (ms * 1e-3) / 1e12
        extras["min_plus_8192"] = {"ms": round(ms, 3), "GOPS": round(tops * 1e3, 1),
                                   "frac_fp32_vector_peak": round(tops / PEAK_FP32_VECTOR_TFLOPS, 4),
                                   "frac_issue_ceiling": round(tops / ISSUE_CEILING_TOPS, 4)}
        log(f"min_plus 8192^3: {ms:.3f} ms {tops:.2f} T ops/s")
        del keep, call
        torch.cuda.empty_cache()
        extras["apsp_er1024"] = apsp_1024(torch, simd2_amd, apsp_keep)
        log(f"apsp n=1024: {extras['apsp_er1024']}")
        extras["pld_knn"] = pld_knn(torch, simd2_amd)
        log(f"pld knn: {extras['pld_knn']}")
        torch.cuda.empty_cache()
        extras["batched_4096"] = batched_graphs(torch, simd2_amd)
        log(f"batched 4096 graphs: {extras['batched_4096']}")
        torch.cuda.empty_cache()
        extras["learned_8192"] = learned_schedule(torch, args.learned_graphs)
        log(f"learned schedule n=8192: {extras['learned_8192']}")
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        base = cpu_baseline(apsp=apsp_keep[0] if apsp_keep else None)
        cpu = dict(base["single"])
        cpu["openmp"] = base["openmp"]
        cpu["host_cpu"] = base["host_cpu"]
        if "apsp_er1024_loop" in base:
            cpu["apsp_er1024_loop"] = base["apsp_er1024_loop"]
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
        "config": {"workload": f"srgemm M=N=K={n}, all nine semirings, one call each per step "
                               "(the drop-ins; or_and through simd2_orand_packed, minus_square "
                               "through simd2_minus_square_exact with its workspace)",
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
