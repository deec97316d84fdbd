"""python -m simd2_amd.run_apps: the seven graph apps and the PLD kNN over a
list of sizes, as a table and JSON.

The report shape follows the reference's scripts/run_apps.py:100-361 (one line
per application and size: times, then the iteration count); the columns are
this library's two converge-loop structures, run through the app driver
`simd2_app` (the reference's mains rebuilt on the C ABI): the reference's
loop (squaring, separate change test, a flag read per squaring) and the loop
with the change test fused into the squaring.  Both must give the same
iteration count and the same checksum (utils/check_sum.h:3-18), which the
table shows.  With --reference-inputs the rows are the reference table's own (its input
files that exist in its tree, copied under tests/golden/graphs/reference/;
GTC and PLD from the drivers' generators with the table's arguments).
Otherwise inputs are synthetic, in each app's input format:

  APSP / MCP / MinRP / MaxRP  directed Erdos-Renyi edge lists on stdin
                              (apps/apsp/cuASR/apsp_cuASR.cu:78-91), weights
                              U{1..100} (MinRP U{1..9}, MaxRP U(0.01, 0.99));
  APLP                        a random DAG (u < v) edge list on stdin;
  MST                         a PBBS weighted edge array file
                              (apps/data/from_pbbs.h:10-57);
  GTC                         the driver's generator, rgg_1d_directed, seed 7
                              (apps/data/graph_gen.h:33-81);
  PLD                         knn_mxu (minus_square distances + the reference
                              top-k) and the BF16-MFMA precision form, 64
                              Gaussian clusters, d = 128, k = 16.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
APP_BIN = os.path.join(HERE, "simd2_app")

APPS = ["apsp", "aplp", "mcp", "maxrp", "minrp", "mst", "gtc"]
NAMES = {"apsp": "APSP", "aplp": "APLP", "mcp": "MCP", "maxrp": "MaxRP", "minrp": "MinRP",
         "mst": "MST", "gtc": "GTC"}


def edge_list(app, v, p, seed):
    """The app's stdin edge list ("v e" then "u w v" lines) as text."""
    rng = np.random.default_rng(seed)
    mask = rng.random((v, v)) < p
    np.fill_diagonal(mask, False)
    if app == "aplp":
        mask = np.triu(mask, 1)
    u, w = np.nonzero(mask)
    if app == "maxrp":
        wt = rng.uniform(0.01, 0.99, len(u))
        fmt = "{} {} {:.6f}"
    else:
        wt = rng.integers(1, 10 if app == "minrp" else 101, len(u))
        fmt = "{} {} {}"
    lines = [f"{v} {len(u)}"] + [fmt.format(a, b, c) for a, b, c in zip(u, w, wt)]
    return "\n".join(lines) + "\n"


def pbbs_file(v, p, seed, path):
    """A symmetric PBBS WeightedEdgeArray with v vertices (a spanning path
    keeps every vertex present)."""
    rng = np.random.default_rng(seed)
    mask = np.triu(rng.random((v, v)) < p, 1)
    mask[np.arange(v - 1), np.arange(1, v)] = True
    u, w = np.nonzero(mask)
    wt = rng.integers(1, 101, len(u))
    with open(path, "w") as f:
        f.write("WeightedEdgeArray\n")
        for a, b, c in zip(u, w, wt):
            f.write(f"{a} {b} {c}\n")


def parse(out):
    lines = [ln.strip() for ln in out.splitlines() if ln.strip()]
    ms = float(lines[0])
    iters = int(lines[1].split()[1])
    checksum = float(lines[2].split()[1])
    return ms, iters, checksum


def run_app(app, v, p, loop, reps, seed=1, binary=APP_BIN):
    """(mean ms per solve, iterations, checksum) of one app run."""
    cmd = [binary, app]
    stdin = None
    tmp = None
    if app == "gtc":
        cmd += [str(v), str(p)]
    elif app == "mst":
        fd, tmp = tempfile.mkstemp(suffix=".pbbs")
        os.close(fd)
        pbbs_file(v, p, seed, tmp)
        cmd += ["-f", tmp]
    else:
        cmd += ["-f"]
        stdin = edge_list(app, v, p, seed)
    cmd += ["--reps", str(reps), "--warmup", "1", "--loop", loop]
    try:
        res = subprocess.run(cmd, input=stdin, capture_output=True, text=True, check=True)
    finally:
        if tmp:
            os.unlink(tmp)
    return parse(res.stdout)


REFERENCE_DATA = os.path.join(os.path.dirname(HERE), "tests", "golden", "graphs", "reference")
# The rows of the reference's table (scripts/run_apps.py:100-361) whose input
# files exist in its tree: (app, table name, file); GTC and PLD use the
# drivers' generators with the reference's arguments.
REFERENCE_ROWS = [("aplp", "APLP-Small", "DAG_rmat_4096.txt"),
                  ("aplp", "APLP-Medium", "DAG_rmat_8192.txt"),
                  ("mst", "MST-Small", "mst_rmat_1024.txt"),
                  ("minrp", "MINRP-Small", "minrp_rmat_4096.txt"),
                  ("mcp", "MCP-1024", "parsed_rmat_1024.txt"),
                  ("mcp", "MCP-2048", "parsed_rmat_2048.txt"),
                  ("apsp", "APSP-1024", "parsed_rmat_1024.txt"),
                  ("apsp", "APSP-2048", "parsed_rmat_2048.txt")]
REFERENCE_GTC = [("GTC-SMALL", 1024, 0.0001), ("GTC-Medium", 4096, 0.0001),
                 ("GTC-Large", 8192, 0.0001)]
REFERENCE_PLD = [("KNN-SMALL", 4096, 10), ("KNN-Medium", 8192, 10)]


def run_app_file(app, path, loop, reps, binary=APP_BIN):
    """(mean ms per solve, iterations, checksum) of simd2_app on a reference
    input file (gzip-compressed edge list or PBBS file)."""
    import gzip
    with gzip.open(path, "rt") as f:
        data = f.read()
    cmd = [binary, app, "-f"]
    stdin, tmp = None, None
    if app == "mst":
        fd, tmp = tempfile.mkstemp(suffix=".pbbs")
        with os.fdopen(fd, "w") as f:
            f.write(data)
        cmd.append(tmp)
    else:
        stdin = data
    cmd += ["--reps", str(reps), "--warmup", "1", "--loop", loop]
    try:
        res = subprocess.run(cmd, input=stdin, capture_output=True, text=True, check=True)
    finally:
        if tmp:
            os.unlink(tmp)
    return parse(res.stdout)


def run_pld_uniform(v, k, reps=3):
    """The reference's PLD row: knn_mxu with rn = qn = d = v and k, on
    coordinates uniform in [0, 10) as its test generates them
    (apps/pld/cuASR/test.cpp); returns the distance kernel's ms inside
    knn_mxu (what the reference's kernel_time measures) and the whole call's."""
    import time
    from .knn import knn_mxu
    rng = np.random.default_rng(v)
    ref = (10.0 * rng.random((v, v))).astype(np.float32)
    qry = (10.0 * rng.random((v, v))).astype(np.float32)
    knn_mxu(ref[:256], qry[:, :256], k)  # warm-up
    kern, wall = 0.0, 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, t = knn_mxu(ref, qry, k)
        wall += time.perf_counter() - t0
        kern += t
    return kern / reps * 1e3, wall / reps * 1e3


def reference_table(reps, data=REFERENCE_DATA, pld=True):
    """The reference table's rows on the reference's own inputs."""
    rows = []
    print("Application Datasize reference-loop(ms) fused-loop(ms) iterations same-result",
          flush=True)
    for app, name, fname in REFERENCE_ROWS:
        path = os.path.join(data, fname + ".gz")
        r_ms, r_it, r_cs = run_app_file(app, path, "reference", reps)
        f_ms, f_it, f_cs = run_app_file(app, path, "fused", reps)
        same = r_it == f_it and r_cs == f_cs
        rows.append({"row": name, "app": NAMES[app], "input": fname,
                     "reference_loop_ms": round(r_ms, 3), "fused_loop_ms": round(f_ms, 3),
                     "iterations": r_it, "checksum": r_cs, "same_result": same})
        print(f"{name}: {fname} {r_ms:.3f} {f_ms:.3f} {r_it} {same}", flush=True)
    for name, v, p in REFERENCE_GTC:
        r_ms, r_it, r_cs = run_app("gtc", v, p, "reference", reps)
        f_ms, f_it, f_cs = run_app("gtc", v, p, "fused", reps)
        same = r_it == f_it and r_cs == f_cs
        rows.append({"row": name, "app": "GTC", "input": f"rgg_1d_directed({v}, {p}), seed 7",
                     "reference_loop_ms": round(r_ms, 3), "fused_loop_ms": round(f_ms, 3),
                     "iterations": r_it, "checksum": r_cs, "same_result": same})
        print(f"{name}: {v} {r_ms:.3f} {f_ms:.3f} {r_it} {same}", flush=True)
    if pld:
        for name, v, k in REFERENCE_PLD:
            kern, wall = run_pld_uniform(v, k, reps)
            rows.append({"row": name, "app": "PLD", "input": f"rn = qn = d = {v}, k = {k}",
                         "knn_mxu_distance_kernel_ms": round(kern, 3),
                         "knn_mxu_call_ms": round(wall, 3)})
            print(f"{name}: {v}^3 k={k} distances {kern:.3f} whole call {wall:.3f}", flush=True)
    return rows


def run_pld(v, reps=3):
    """knn_mxu and the precision form for v queries x v references: the
    distance kernel's ms inside knn_mxu (what the reference's kernel_time
    measures), the precision form's whole call in ms, and the fraction of
    queries whose k neighbours agree."""
    import torch
    import simd2_amd
    from simd2_amd.datasets import gaussian_clusters
    from simd2_amd.knn import KnnBF16, knn_mxu
    d, k = 128, 16
    ref, c = gaussian_clusters(torch, v, d, 64, seed=100)
    qry, _ = gaussian_clusters(torch, v, d, 64, seed=101, centers=c)
    ref_h = ref.cpu().numpy()
    qry_h = np.ascontiguousarray(qry.cpu().numpy().T)  # dimension-major, as knn_mxu takes it
    knn_mxu(ref_h, qry_h, k)  # warm-up
    secs = 0.0
    for _ in range(reps):
        _, idx_mxu, t = knn_mxu(ref_h, qry_h, k)
        secs += t  # the distance kernel's time, as the reference's knn_mxu reports it
    runner = KnnBF16(v, v, d)
    runner(qry, ref, k)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _, idx = runner(qry, ref, k)
    e1.record()
    e1.synchronize()
    same = float(np.mean(np.all(idx.cpu().numpy() == np.asarray(idx_mxu).T, axis=1)))
    return secs / reps * 1e3, e0.elapsed_time(e1) / reps, same


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sizes", type=int, nargs="+", default=[1024, 2048])
    ap.add_argument("--density", type=float, default=0.01)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--apps", nargs="+", default=APPS + ["pld"])
    ap.add_argument("--json", default=None, help="also write the rows here")
    ap.add_argument("--reference-inputs", action="store_true",
                    help="the reference table's rows on its own input files instead")
    args = ap.parse_args(argv)
    if not os.path.exists(APP_BIN):
        sys.exit(f"{APP_BIN} is missing: build it with make -C simd2_amd/csrc")
    if args.reference_inputs:
        rows = reference_table(args.reps)
        if args.json:
            with open(args.json, "w") as f:
                json.dump(rows, f, indent=1)
        return rows
    rows = []
    print("Application Datasize reference-loop(ms) fused-loop(ms) iterations same-result",
          flush=True)
    for app in args.apps:
        for v in args.sizes:
            if app == "pld":
                mxu_ms, bf16_ms, same = run_pld(v, args.reps)
                row = {"app": "PLD", "size": v, "knn_mxu_distance_kernel_ms": round(mxu_ms, 3),
                       "bf16_form_call_ms": round(bf16_ms, 3), "same_neighbours": same}
                print(f"PLD: {v} knn_mxu-distances {mxu_ms:.3f} bf16-form {bf16_ms:.3f} "
                      f"k=16 same-neighbours {same:.4f}", flush=True)
            else:
                r_ms, r_it, r_cs = run_app(app, v, args.density, "reference", args.reps)
                f_ms, f_it, f_cs = run_app(app, v, args.density, "fused", args.reps)
                same = r_it == f_it and r_cs == f_cs
                row = {"app": NAMES[app], "size": v, "reference_loop_ms": round(r_ms, 3),
                       "fused_loop_ms": round(f_ms, 3), "iterations": r_it,
                       "checksum": r_cs, "same_result": same}
                print(f"{NAMES[app]}: {v} {r_ms:.3f} {f_ms:.3f} {r_it} {same}", flush=True)
            rows.append(row)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rows, f, indent=1)
    return rows


if __name__ == "__main__":
    main()
