"""Benchmark of the semiring-GEMM path on one MI355X.

Workload (one "step"): the single-srgemm suite M = N = K = 4096 for all nine
operators, each through its drop-in entry point (cuasr_*_srsgemm; or_and on
the configuration's bit-packed uint32 matrices through simd2_orand_packed),
inputs resident in HBM before timing: float inputs integer-valued U{1..100}
with 10% +inf / -inf holes for min- / max-based ops, U(0,1] for max_mul /
min_mul, no holes for plus_mul / minus_square; or_and on Bernoulli(0.02)
bits.  `value` is the step's aggregate semiring throughput, 9 x 2*M*N*K
semiring operations / step time, in GOPS (one (x) and one (+) per (i, j, k)
for every operator, or_and included).

Beside it, per operator: the median time of 20 event-bracketed launches,
GOPS, the fraction of the FP32 vector peak and of the VALU issue ceiling of
the operator's instruction mix (measured at run time by
simd2_amd/simd2_issue_bench), and -- in the CPU leg, which is the checker --
whether sampled rows of the timed output equal the oracle's bit for bit
(plus_mul: within 1e-5 relative of FP64).  minus_square's int8 form
(simd2_minus_square_exact, exact on this integer data) and or_and's int32
drop-in are timed beside their step entries.  Extras: min_plus at 8192^3
(checked the same way), the n=1024 APSP converge loop, the PLD/kNN workload,
the batched solver, and the learned schedule at n = 8192 on the 200 held-out
graphs.

Roofline: min_plus at 4096^3 (the kernel the target names), achieved = 2 n^3
/ median launch time on the launching stream; traffic = HBM bytes per launch
from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over a child run of
this script, scaled by what the same counters read for a 48 MiB streaming
copy (null when rocprofv3 is unavailable).  CPU baseline: the reference's
host_srgemm (built from its sources) single-threaded and the oracle's OpenMP
row split on all host cores, on row samples of the same min_plus workload,
plus the full n=1024 APSP loop, checked against the GPU's matrix.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_VECTOR_TFLOPS = 157.3   # MI355X published FP32 vector peak (FMA, packed)
PEAK_INT8_MFMA_TOPS = 5000.0      # MI355X published dense int8 MFMA peak
OPS = ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul", "min_mul", "min_max",
       "or_and", "minus_square"]
# Instruction mix of each VALU operator's inner step (simd2_issue_bench names)
# and the semiring operations one wave-instruction lane carries in it:
# 2 products + one three-input fold per 4 ops, or sub + fmac per 2 ops.
ISSUE_MIX = {"min_plus": ("min_plus", 4 / 3), "max_plus": ("min_plus", 4 / 3),
             "min_mul": ("max_mul", 4 / 3), "max_mul": ("max_mul", 4 / 3),
             "max_min": ("max_min", 4 / 3), "min_max": ("max_min", 4 / 3),
             "minus_square": ("sub_fmac", 1.0)}
CAL_FLOATS = 12 * 1024 * 1024     # traffic calibration copy: 48 MiB in, 48 MiB out
SAMPLE_ROWS = 8


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


class OpCase:
    """One operator of the suite: the step's launcher (`call`), the tensors it
    uses, and the form timed beside it (`side`: minus_square's int8 entry
    point, or_and's int32 drop-in), if any."""

    def __init__(self, op, call, tensors, side=None, side_name=None):
        self.op, self.call, self.tensors = op, call, tensors
        self.side, self.side_name = side, side_name

    def sample(self, rows):
        """Host copies of what a row check needs: rows of A, C and D, all of B
        (or_and: the 0/1 matrices unpacked)."""
        import torch
        idx = torch.from_numpy(rows).cuda()
        if self.op == "or_and":
            import simd2_amd
            A, B, C, D = self.tensors[:4]
            n = B.shape[0]
            un = lambda t: simd2_amd.unpack_bits(t, n)
            return {"A": un(A[idx]).cpu().numpy(), "B": un(B).cpu().numpy(),
                    "C": un(C[idx]).cpu().numpy(), "D": un(D[idx]).cpu().numpy()}
        A, B, C, D = self.tensors[:4]
        return {"A": A[idx].cpu().numpy(), "B": B.cpu().numpy(), "C": C[idx].cpu().numpy(),
                "D": D[idx].cpu().numpy()}


def make_call(torch, simd2_amd, op, n, seed):
    """The suite's launcher for `op` at n^3 (row-major A @ B as the
    column-major call (N, M, K, B, N, A, K, C, N, D), as the reference's
    tests make it)."""
    stream = torch.cuda.current_stream().cuda_stream
    L = simd2_amd.lib()
    if op == "or_and":
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        bits = [(torch.rand((n, n), generator=g, device="cuda") < 0.02).to(torch.int32)
                for _ in range(3)]
        A, B, C = (simd2_amd.pack_bits(b) for b in bits)
        D = torch.empty_like(C)
        nw = A.shape[1]
        # packed rows are row-major: the packed entry point takes (M, N, K, A, B)
        args = (n, n, n, A.data_ptr(), nw, B.data_ptr(), nw, C.data_ptr(), nw, D.data_ptr(), nw,
                1, stream)
        Di = torch.empty_like(bits[2])
        fn = getattr(L, simd2_amd.DROPIN["or_and"])
        args_i = (n, n, n, bits[1].data_ptr(), n, bits[0].data_ptr(), n, bits[2].data_ptr(), n,
                  Di.data_ptr(), True, stream)

        def call():
            st = L.simd2_orand_packed(*args)
            if st:
                raise RuntimeError(f"or_and status {st}")

        def dropin():
            st = fn(*args_i)
            if st:
                raise RuntimeError(f"or_and int32 drop-in status {st}")
        return OpCase(op, call, (A, B, C, D, bits, Di), dropin, "int32_dropin")
    A, B, C = make_inputs(torch, op, n, seed)
    D = torch.empty_like(C)
    fn = getattr(L, simd2_amd.DROPIN[op])
    args = (n, n, n, B.data_ptr(), n, A.data_ptr(), n, C.data_ptr(), n, D.data_ptr(), True,
            stream)

    def call():
        st = fn(*args)
        if st:
            raise RuntimeError(f"{op} status {st}")
    if op != "minus_square":
        return OpCase(op, call, (A, B, C, D))
    # minus_square's exact int8 entry point beside the drop-in, its workspace
    # allocated here, outside any timed region
    ws = torch.empty(simd2_amd.minus_square_workspace(n, n, n), dtype=torch.uint8,
                     device="cuda")
    D2 = torch.empty_like(C)
    args_x = (n, n, n, B.data_ptr(), n, A.data_ptr(), n, C.data_ptr(), n, D2.data_ptr(), n, 1,
              ws.data_ptr(), ws.numel(), stream)

    def exact():
        st = L.simd2_minus_square_exact(*args_x)
        if st:
            raise RuntimeError(f"{op} exact status {st}")
    return OpCase(op, call, (A, B, C, D, D2, ws), exact, "int8_exact")


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


def host_cpu():
    """The host's CPU model string (lscpu) and online / usable core counts."""
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                   timeout=10).stdout.strip())
    except (OSError, subprocess.SubprocessError, ValueError):
        nproc = len(os.sched_getaffinity(0))
    return {"lscpu_model": model, "nproc": nproc, "online_cpus": os.cpu_count()}


def issue_rates():
    """VALU issue rates (wave-instructions per CU per cycle) of the kernels'
    instruction mixes, measured now by simd2_issue_bench: {mix: max over
    1/2/4/8 waves per SIMD}.  None when the binary is missing or fails."""
    exe = os.path.join(ROOT, "simd2_amd", "simd2_issue_bench")
    if not os.path.exists(exe):
        return None
    try:
        out = subprocess.run([exe, "--json"], capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.SubprocessError):
        return None
    if out.returncode:
        return None
    rates = {}
    for line in out.stdout.splitlines():
        r = json.loads(line)
        rates.setdefault(r["mix"], {})[int(r["waves_per_simd"])] = float(r["rate"])
    return {mix: {"rate": max(by_w.values()), "by_waves_per_simd": by_w}
            for mix, by_w in rates.items()}


def issue_ceiling(op, rates, info):
    """T semiring ops/s at the measured issue rate of op's mix and the
    device's clock, or None."""
    if not rates or op not in ISSUE_MIX or ISSUE_MIX[op][0] not in rates:
        return None
    mix, ops_per_lane_instr = ISSUE_MIX[op]
    return (rates[mix]["rate"] * 64 * ops_per_lane_instr * info["cu_count"] *
            info["clock_khz"] * 1e3 / 1e12)


def pmc_child(n):
    """The program the traffic passes profile: the 48 MiB calibration copy,
    then one min_plus launch of the suite at n^3."""
    import torch
    import simd2_amd
    simd2_amd.lib()
    cal = torch.ones(CAL_FLOATS, device="cuda")
    torch.neg(cal, out=torch.empty_like(cal))
    case = make_call(torch, simd2_amd, "min_plus", n, seed=1000 + OPS.index("min_plus"))
    case.call()
    torch.cuda.synchronize()


def measure_traffic(n):
    """HBM bytes per launch of the min_plus kernel at n^3: rocprofv3 passes
    with FETCH_SIZE and then WRITE_SIZE (KiB per dispatch) over
    `bench.py --pmc-child`, each scaled by the factor that makes the 48 MiB
    calibration copy read 48 MiB on that counter."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    got = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="simd2_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--pmc-child", "--n", str(n)]
        p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                             start_new_session=True, env=env, cwd="/tmp")
        try:
            p.wait(timeout=150)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            shutil.rmtree(d, ignore_errors=True)
            return None
        vals = {"min_plus": [], "cal": []}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row.get("Kernel_Name", "")
                    if "srgemm_valu" in name and "MinPlus" in name:
                        vals["min_plus"].append(float(row["Counter_Value"]))
                    elif "neg_kernel" in name:
                        vals["cal"].append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if len(vals["min_plus"]) != 1 or len(vals["cal"]) != 1 or vals["cal"][0] <= 0:
            return None
        got[counter] = (vals["min_plus"][0], vals["cal"][0])
    cal_bytes = CAL_FLOATS * 4
    f_fetch = cal_bytes / (got["FETCH_SIZE"][1] * 1024)
    f_write = cal_bytes / (got["WRITE_SIZE"][1] * 1024)
    fetch = got["FETCH_SIZE"][0] * 1024 * f_fetch
    write = got["WRITE_SIZE"][0] * 1024 * f_write
    return {"bytes": int(fetch + write), "fetch_bytes": int(fetch), "write_bytes": int(write),
            "fetch_KiB_raw": got["FETCH_SIZE"][0], "write_KiB_raw": got["WRITE_SIZE"][0],
            "scale_fetch": round(f_fetch, 4), "scale_write": round(f_write, 4)}


def time_launches(torch, call, reps):
    """Average device time per launch, HIP events on the launching stream."""
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        call()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def check_samples(samples):
    """The checker: each op's sampled output rows against the oracle's rows
    (bit for bit; plus_mul within 1e-5 relative of FP64; or_and against the
    boolean product).  Returns {name: True/False/relative error}."""
    import oracle
    threads = len(os.sched_getaffinity(0))
    out = {}
    for name, (op, smp) in samples.items():
        a, b, c, d = smp["A"], smp["B"], smp["C"], smp["D"]
        if op == "or_and":
            want = ((a.astype(np.int64) @ b.astype(np.int64)) > 0) | (c > 0)
            out[name] = {"bit_exact": bool(np.array_equal(d > 0, want))}
        elif op == "plus_mul":
            want = a.astype(np.float64) @ b.astype(np.float64) + c.astype(np.float64)
            rel = float(np.max(np.abs(d - want) / np.maximum(np.abs(want), 1e-30)))
            out[name] = {"within_1e-5_rel_fp64": bool(rel <= 1e-5), "max_rel_err": rel}
        else:
            want = oracle.srgemm_f32(op, a, b, c, True, threads=threads)
            out[name] = {"bit_exact": bool(np.array_equal(
                np.ascontiguousarray(d).view(np.uint32), want.view(np.uint32)))}
        out[name]["rows"] = int(a.shape[0])
    return out


def cpu_baseline(n=4096, apsp=None, samples=None):
    """The CPU leg on this host: the reference's host_srgemm (built from its
    sources; the oracle's port when it is absent) single-threaded and the
    oracle's OpenMP row split on all usable cores, on row samples of the
    min_plus workload; when `apsp` = (X0, X_gpu, squarings) of the n=1024
    APSP run is given, the full converge loop of that graph, checked against
    the GPU's matrix bit for bit; and the parity checks of `samples`."""
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
                     "kind": kind, "sample": f"min_plus rows 0..{rows1 - 1} of {n}^3 "
                     f"(row-sampled, {t1:.1f} s)"}
    rows = 16 * cores
    D = np.full((n, n), np.inf, np.float32)
    t0 = time.perf_counter()
    oracle.srgemm_f32_rows("min_plus", A, B, C, D, True, 0, rows, cores)
    t2 = time.perf_counter() - t0
    out["openmp"] = {"value": 2.0 * rows * n * n / t2 / 1e9, "unit": "GOPS", "cores": cores,
                     "kind": "port", "sample": f"min_plus rows 0..{rows - 1} of {n}^3 "
                     f"(row-sampled, OpenMP over rows, bit-identical, {t2:.1f} s)"}
    out["host_cpu"] = host_cpu()
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
    if samples:
        out["parity"] = check_samples(samples)
    return out


def off_grid(torch, simd2_amd, samples, rates, info):
    """plus_mul (matrix cores) and min_plus (VALU) through the drop-ins on
    shapes off every tile grid: 4000^3, and 1000 x 777 x 300 (row-major
    M x K times K x N).  Median of 20 launches each; sampled rows go to the
    checker under 'off_grid_<op>_<shape>'."""
    stream = torch.cuda.current_stream().cuda_stream
    L = simd2_amd.lib()
    out = {}
    for (M, N, K) in ((4000, 4000, 4000), (1000, 777, 300)):
        for op in ("plus_mul", "min_plus"):
            g = torch.Generator(device="cuda")
            g.manual_seed(M + N + K)
            A = torch.randint(1, 101, (M, K), generator=g, device="cuda").float()
            B = torch.randint(1, 101, (K, N), generator=g, device="cuda").float()
            C = torch.randint(1, 101, (M, N), generator=g, device="cuda").float()
            if op == "min_plus":
                A[torch.rand((M, K), generator=g, device="cuda") < 0.1] = float("inf")
            D = torch.empty_like(C)
            fn = getattr(L, simd2_amd.DROPIN[op])
            args = (N, M, K, B.data_ptr(), N, A.data_ptr(), K, C.data_ptr(), N, D.data_ptr(),
                    True, stream)

            def call():
                st = fn(*args)
                if st:
                    raise RuntimeError(f"{op} {M}x{N}x{K} status {st}")
            call()
            torch.cuda.synchronize()
            ms = time_launches_median(torch, call)
            tops = 2.0 * M * N * K / (ms * 1e-3) / 1e12
            key = f"{op}_{M}x{N}x{K}"
            rec = {"ms": round(ms, 4), "GOPS": round(tops * 1e3, 1)}
            if op == "plus_mul":
                rec["frac_fp32_mfma_peak"] = round(tops / PEAK_FP32_VECTOR_TFLOPS, 4)
            else:
                ceil = issue_ceiling(op, rates, info)
                rec["frac_fp32_vector_peak"] = round(tops / PEAK_FP32_VECTOR_TFLOPS, 4)
                rec["frac_issue_ceiling"] = round(tops / ceil, 4) if ceil else None
            out[key] = rec
            rows = np.linspace(0, M - 1, SAMPLE_ROWS).astype(np.int64)
            idx = torch.from_numpy(rows).cuda()
            samples["off_grid_" + key] = (op, {
                "A": A[idx].cpu().numpy(), "B": B.cpu().numpy(), "C": C[idx].cpu().numpy(),
                "D": D[idx].cpu().numpy()})
            del A, B, C, D
    torch.cuda.empty_cache()
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


def batched_graphs(torch, simd2_amd, count=4096, seed=0, ceiling=None):
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
    default_ms = float(np.median([timed_solve() for _ in range(3)]))  # as the tuned figure
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
            "frac_min_plus_issue_ceiling": round(ops / (ms * 1e-3) / 1e12 / ceiling, 4)
            if ceiling else None,
            "launches_per_round": launches, "tile_configs": hist,
            "identical_to_loop": identical, "same_iterations": same_itr}


def learned_schedule(torch, count, n=8192):
    """Learned-schedule APSP / MCP at n (8192, or 1024 where a squaring takes
    tens of microseconds and the features' cost shows) on the first `count`
    held-out graphs (ER / grid / geometric in turn): plain converge loop vs
    the learned loop vs the scheduled loop without the network, every result
    checked bit for bit against the plain loop."""
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
    from simd2_amd.knn import KnnBF16, knn_select, select_workspace_bytes
    ref, centers = gaussian_clusters(torch, R, d, 64, seed=100)
    qry, _ = gaussian_clusters(torch, Q, d, 64, seed=101, centers=centers)
    qT = qry.t().contiguous()
    D = torch.empty((R, Q), device="cuda")
    runner = KnnBF16(Q, R, d)
    ws = torch.empty(select_workspace_bytes(Q, R, k), dtype=torch.uint8, device="cuda")

    def direct():
        simd2_amd.srgemm("minus_square", ref, qT, out=D)
        return knn_select(D, k, sqrt=False, workspace=ws)

    def select_only():
        return knn_select(D, k, sqrt=False, workspace=ws)

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
    # the reference-layout top-k alone, on the materialised 1 GiB matrix
    direct()
    torch.cuda.synchronize()
    ms_sel = time_launches_median(torch, select_only, 10)
    res["direct_fp32"]["select_ms"] = round(ms_sel, 4)
    res["direct_fp32"]["select_GBps"] = round(4.0 * Q * R / (ms_sel * 1e-3) / 1e9, 1)
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
    ap.add_argument("--no-extras", action="store_true", help="skip 8192^3 and the workloads")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 passes")
    ap.add_argument("--learned-graphs", type=int, default=200,
                    help="held-out n=8192 graphs for the learned-schedule extra")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--only-extra", choices=["pld_knn", "batched_4096", "learned_8192",
                                             "learned_1024", "apsp_er1024"],
                    help="run one extra workload alone and print its result (profiling)")
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.n)
        return
    if args.only_extra:
        import torch
        import simd2_amd
        simd2_amd.lib()
        fn = {"pld_knn": lambda: pld_knn(torch, simd2_amd),
              "batched_4096": lambda: batched_graphs(torch, simd2_amd),
              "learned_8192": lambda: learned_schedule(torch, args.learned_graphs),
              "learned_1024": lambda: learned_schedule(torch, 200, n=1024),
              "apsp_er1024": lambda: apsp_1024(torch, simd2_amd)}[args.only_extra]
        print(json.dumps({args.only_extra: fn()}), flush=True)
        return
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

    cases = {op: make_call(torch, simd2_amd, op, n, seed=1000 + i) for i, op in enumerate(OPS)}
    torch.cuda.synchronize()

    def step():
        for op in OPS:
            cases[op].call()

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

    rates = issue_rates()
    log(f"issue rates: { {m: r['rate'] for m, r in rates.items()} if rates else None}")
    per_op = {}
    rows = np.linspace(0, n - 1, SAMPLE_ROWS).astype(np.int64)
    samples = {}
    for op in OPS:
        c = cases[op]
        ms = time_launches_median(torch, c.call, 20)
        tops = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
        ceil = issue_ceiling(op, rates, info)
        per_op[op] = {"ms": round(ms, 4), "GOPS": round(tops * 1e3, 1),
                      "frac_fp32_vector_peak": round(tops / PEAK_FP32_VECTOR_TFLOPS, 4),
                      "issue_ceiling_TOPS": round(ceil, 2) if ceil else None,
                      "frac_issue_ceiling": round(tops / ceil, 4) if ceil else None}
        if op == "or_and":
            per_op[op]["entry"] = "simd2_orand_packed (bit-packed, the configuration's format)"
        else:
            per_op[op]["entry"] = simd2_amd.DROPIN[op]
        torch.cuda.synchronize()
        samples[op] = (op, c.sample(rows))
        if c.side is not None:
            ms2 = time_launches_median(torch, c.side, 20)
            t2 = 2.0 * n ** 3 / (ms2 * 1e-3) / 1e12
            side = {"ms": round(ms2, 4), "GOPS": round(t2 * 1e3, 1)}
            if op == "minus_square":
                mf, va = simd2_amd.minus_square_tiles(c.tensors[5], n, n, n)
                side.update({"entry": "simd2_minus_square_exact (int8 MFMA, exact: "
                                      "|a|^2+|b|^2-2a.b where sums <= 2^24)",
                             "int8_tiles": mf, "direct_tiles": va,
                             "frac_int8_mfma_peak": round(t2 / PEAK_INT8_MFMA_TOPS, 4),
                             "identical_to_dropin": bool(torch.equal(
                                 c.tensors[4].view(torch.int32), c.tensors[3].view(torch.int32)))})
            else:
                A, B, C, D, bits01, Di = c.tensors
                side.update({"entry": "cuasr_orand_srsgemm (int32 per boolean)",
                             "identical_to_packed": bool(torch.equal(
                                 Di, simd2_amd.unpack_bits(D, n)))})
            per_op[op][c.side_name] = side
        log(f"{op}: {ms:.3f} ms  {tops:.2f} T ops/s"
            + (f"; {c.side_name} {per_op[op][c.side_name]['ms']:.3f} ms" if c.side else ""))
    # the aggregate with minus_square's exact int8 form in place of the drop-in
    ms_alt = ms_step - per_op["minus_square"]["ms"] + per_op["minus_square"]["int8_exact"]["ms"]

    mp = per_op["min_plus"]
    mp_tops = mp["GOPS"] / 1e3
    tile = {0: "128x128", 1: "64x64", 2: "128x64"}[
        simd2_amd.lib().simd2_pick_tile_config(0, n, n, n)]
    cases.clear()
    torch.cuda.empty_cache()
    traffic = None if args.no_traffic else measure_traffic(n)
    log(f"traffic: {traffic}")
    roofline = {"kernel": f"srgemm_valu_kernel<MinPlus,{tile}x16>", "bound": "valu",
                "achieved": round(mp_tops, 3), "peak": PEAK_FP32_VECTOR_TFLOPS,
                "unit": "TFLOP/s", "frac": round(mp_tops / PEAK_FP32_VECTOR_TFLOPS, 4),
                "traffic": traffic["bytes"] if traffic else None,
                "algorithmic_bytes": 4 * n * n * 4,
                "traffic_detail": traffic,
                "issue_ceiling": mp["issue_ceiling_TOPS"],
                "frac_issue_ceiling": mp["frac_issue_ceiling"],
                "issue_rates": rates,
                "note": "achieved = 2*n^3 semiring ops (one (x) and one (+) per i,j,k) / "
                        "median launch time on the launching stream, in T semiring ops/s; "
                        "algorithmic bytes = A, B, C read and D written once (4 n^2 x 4 B); "
                        "traffic = rocprofv3 FETCH_SIZE + WRITE_SIZE of one launch, scaled "
                        "by a 48 MiB streaming copy; issue_ceiling = the VALU issue rate "
                        "simd2_issue_bench measures for the kernel's instruction mix "
                        "(2 x v_add_f32 + v_min3_f32 per 4 ops) x 64 lanes x 4/3 ops x CUs "
                        "x clock"}
    # the longest kernel of the step
    lop = max(per_op, key=lambda o: per_op[o]["ms"])
    roofline["longest_kernel"] = {
        "op": lop, "ms": per_op[lop]["ms"], "achieved": round(per_op[lop]["GOPS"] / 1e3, 3),
        "frac": per_op[lop]["frac_fp32_vector_peak"],
        "issue_ceiling": per_op[lop]["issue_ceiling_TOPS"],
        "frac_issue_ceiling": per_op[lop]["frac_issue_ceiling"]}

    extras = {}
    apsp_keep = []
    if not args.no_extras:
        c = make_call(torch, simd2_amd, "min_plus", 8192, seed=77)
        c.call()
        torch.cuda.synchronize()
        ms = time_launches(torch, c.call, 5)
        tops = 2.0 * 8192 ** 3 / (ms * 1e-3) / 1e12
        ceil = issue_ceiling("min_plus", rates, info)
        extras["min_plus_8192"] = {
            "ms": round(ms, 3), "GOPS": round(tops * 1e3, 1),
            "tile": {0: "128x128", 1: "64x64", 2: "128x64"}[
                simd2_amd.lib().simd2_pick_tile_config(0, 8192, 8192, 8192)],
            "frac_fp32_vector_peak": round(tops / PEAK_FP32_VECTOR_TFLOPS, 4),
            "frac_issue_ceiling": round(tops / ceil, 4) if ceil else None}
        samples["min_plus_8192"] = ("min_plus", c.sample(
            np.linspace(0, 8191, SAMPLE_ROWS).astype(np.int64)))
        log(f"min_plus 8192^3: {ms:.3f} ms {tops:.2f} T ops/s")
        del c
        torch.cuda.empty_cache()
        extras["off_grid"] = off_grid(torch, simd2_amd, samples, rates, info)
        log(f"off-grid shapes: {extras['off_grid']}")
        extras["apsp_er1024"] = apsp_1024(torch, simd2_amd, apsp_keep)
        log(f"apsp n=1024: {extras['apsp_er1024']}")
        extras["pld_knn"] = pld_knn(torch, simd2_amd)
        log(f"pld knn: {extras['pld_knn']}")
        torch.cuda.empty_cache()
        extras["batched_4096"] = batched_graphs(torch, simd2_amd,
                                                ceiling=issue_ceiling("min_plus", rates, info))
        log(f"batched 4096 graphs: {extras['batched_4096']}")
        torch.cuda.empty_cache()
        extras["learned_8192"] = learned_schedule(torch, args.learned_graphs)
        log(f"learned schedule n=8192: {extras['learned_8192']}")
        extras["learned_1024"] = learned_schedule(torch, 200, n=1024)
        log(f"learned schedule n=1024: {extras['learned_1024']}")
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        base = cpu_baseline(apsp=apsp_keep[0] if apsp_keep else None, samples=samples)
        cpu = dict(base["single"])
        cpu["openmp"] = base["openmp"]
        cpu["host_cpu"] = base["host_cpu"]
        if "apsp_er1024_loop" in base:
            cpu["apsp_er1024_loop"] = base["apsp_er1024_loop"]
        par = base.get("parity", {})
        for op in OPS:
            if op in par:
                per_op[op]["parity"] = par[op]
        if "min_plus_8192" in par:
            extras["min_plus_8192"]["parity"] = par["min_plus_8192"]
        for key in extras.get("off_grid", {}):
            if "off_grid_" + key in par:
                extras["off_grid"][key]["parity"] = par["off_grid_" + key]
        log(f"cpu baseline: {cpu}")
        log(f"parity: {par}")

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
                               "through the drop-in entry points (or_and bit-packed through "
                               "simd2_orand_packed, as the configuration states)",
                   "M": n, "N": n, "K": n, "epilogue": True,
                   "or_and": "bit-packed uint32, density 0.02"},
        "aggregate_with_int8_minus_square_GOPS": round(ops_per_step / (ms_alt * 1e-3) / 1e9, 1),
        "per_op": per_op,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
