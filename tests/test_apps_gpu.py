"""The graph-app driver (simd2_app) against the oracle: each reference app's
input format, fill convention, operator and convergence check."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from util import FLT_MAX

pytestmark = pytest.mark.gpu
APP = os.path.join(ROOT, "simd2_amd", "simd2_app")


def run_app(args, stdin=None, tmp_path=None, dtype=np.float32, v=None):
    dump = os.path.join(str(tmp_path), "out.bin")
    print("simd2_app", " ".join(args), flush=True)
    r = subprocess.run([APP] + args + ["--reps", "1", "--dump", dump], input=stdin,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    ms = float(lines[0])
    itr = int(lines[1].split()[1])
    out = np.fromfile(dump, dtype=dtype)
    if v is not None:
        out = out.reshape(v, v)
    return ms, itr, out


def edge_text(v, edges):
    # 9 significant digits round-trip every float32 weight exactly
    return f"{v} {len(edges)}\n" + "".join(f"{a} {b} {w:.9g}\n" for a, b, w in edges)


def random_edges(v, p, seed, wlo=1, whi=100, dag=False):
    rng = np.random.default_rng(seed)
    edges = []
    for a in range(v):
        for b in range(v):
            if a != b and rng.random() < p and (not dag or a < b):
                edges.append((a, b, float(rng.integers(wlo, whi + 1))))
    return edges


def test_apsp_path_graph_stdin(gpu, golden_converge, tmp_path):
    text = open(os.path.join(GOLDEN, "graphs", "16_LD.txt")).read()
    for loop in ("reference", "fused"):
        ms, itr, out = run_app(["apsp", "-f", "--loop", loop], text, tmp_path, v=16)
        assert itr == int(golden_converge["16_LD/apsp/itr"][0])
        assert np.array_equal(out, golden_converge["16_LD/apsp/out"])


@pytest.mark.parametrize("app,op,fill,diag,epi,check", [
    ("apsp", "min_plus", FLT_MAX, 0.0, True, "threshold"),
    ("mcp", "max_min", -FLT_MAX, -FLT_MAX, True, "threshold"),
    ("minrp", "min_mul", FLT_MAX, None, True, "threshold"),
    ("maxrp", "max_mul", 0.0, None, True, "threshold"),
    ("aplp", "max_plus", -FLT_MAX, None, True, "aplp"),
])
def test_stdin_apps_vs_oracle(gpu, oracle_mod, tmp_path, app, op, fill, diag, epi, check):
    v = 96
    if app == "maxrp":   # reliabilities in (0, 1)
        edges = [(a, b, w / 128.0) for a, b, w in random_edges(v, 0.05, 3)]
    elif app == "aplp":  # longest paths on a DAG
        edges = random_edges(v, 0.05, 4, dag=True)
    else:
        edges = random_edges(v, 0.05, 5, wlo=1, whi=9 if app == "minrp" else 100)
    X0 = np.full((v, v), fill, np.float32)
    for a, b, w in edges:
        X0[a, b] = w
    if diag is not None:
        np.fill_diagonal(X0, diag)
    ms, itr, out = run_app([app, "-f"], edge_text(v, edges), tmp_path, v=v)
    Xo, itr_o = oracle_mod.converge_f32(op, X0, epilogue=epi, check=check, threads=8)
    assert itr == itr_o
    assert np.array_equal(out.view(np.uint32), Xo.view(np.uint32))


def test_mst_pbbs_file(gpu, oracle_mod, tmp_path):
    v = 80
    edges = [(a, b, w) for a, b, w in random_edges(v, 0.08, 6) if a < b]
    path = os.path.join(str(tmp_path), "g.pbbs")
    with open(path, "w") as f:
        f.write("WeightedEdgeArray\n")
        for a, b, w in edges:
            f.write(f"{a} {b} {w:.9g}\n")
    vmax = max(max(a, b) for a, b, _ in edges) + 1
    X0 = np.zeros((vmax, vmax), np.float32)
    for a, b, w in edges:
        X0[a, b] = X0[b, a] = w
    X0[(X0 == 0) & ~np.eye(vmax, dtype=bool)] = FLT_MAX
    ms, itr, out = run_app(["mst", "-f", path], None, tmp_path, v=vmax)
    Xo, itr_o = oracle_mod.converge_f32("min_max", X0, epilogue=False, check="threshold")
    assert itr == itr_o and np.array_equal(out, Xo)


def glibc_rgg(v, density, weight, seed, directed):
    """The reference's rgg_1d / rgg_1d_directed draw order with glibc rand()."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    RAND_MAX = np.float32(2147483647.0)
    g = np.zeros((v, v), np.float32)
    for i in range(v):
        for j in range(v):
            if i == j:
                continue
            if directed or j > i:
                w = np.float32(libc.rand()) / (RAND_MAX / np.float32(weight))
                has = np.float32(libc.rand()) / RAND_MAX
                g[i, j] = w if has < np.float32(density) else FLT_MAX
                if not directed:
                    g[j, i] = g[i, j]
    return g


def test_generator_mode_gtc_and_apsp(gpu, oracle_mod, tmp_path):
    v = 120
    g = glibc_rgg(v, 0.03, 10.0, 7, directed=True)
    T0 = ((g < FLT_MAX - 100) | np.eye(v, dtype=bool)).astype(np.int32)
    ms, itr, out = run_app(["gtc", str(v), "0.03"], None, tmp_path, dtype=np.int32, v=v)
    To, itr_o = oracle_mod.converge_i32(T0)
    assert itr == itr_o and np.array_equal(out, To)
    g = glibc_rgg(v, 0.05, 10.0, 7, directed=False)
    ms, itr, out = run_app(["apsp", str(v), "0.05", "10"], None, tmp_path, v=v)
    Xo, itr_o = oracle_mod.converge_f32("min_plus", g, check="threshold")
    assert itr == itr_o and np.array_equal(out, Xo)


def test_report_harness_all_apps(gpu, tmp_path):
    """python -m simd2_amd.run_apps at a small size: every app runs under both
    loop structures with the same iteration count and checksum, and the kNN
    precision form finds knn_mxu's neighbours."""
    import json
    from simd2_amd import run_apps
    out = tmp_path / "apps.json"
    rows = run_apps.main(["--sizes", "256", "--reps", "1", "--density", "0.02",
                          "--json", str(out)])
    assert len(rows) == len(run_apps.APPS) + 1
    assert json.loads(out.read_text()) == rows
    for r in rows:
        if r["app"] == "PLD":
            assert r["same_neighbours"] >= 0.99, r
        else:
            assert r["same_result"] and r["iterations"] >= 1, r
