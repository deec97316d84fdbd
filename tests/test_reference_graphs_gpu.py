"""The reference's own graph inputs (tests/golden/graphs/reference/, copied
from apps/data by tests/golden/make_reference_graphs.py), through the app
driver simd2_app and the converge entry points.

* rmat_4096 (apps/data/rmat_data/rmat_4096.txt): the reference records its
  hop diameter 6 and 16471467 reachable (i, j) pairs besides (0, 0)
  (apps/data/dim_calc/4096, computed by apps/data/dim_calc/dim_cacl.cpp with
  Floyd-Warshall).  Unit-weight APSP by min_plus squaring and the transitive
  closure by bit-packed or_and squaring must reproduce both numbers, in
  ceil(log2 6) + 1 = 4 squarings.
* parsed_rmat_1024 / 2048 (APSP, MCP), mst_rmat_1024 (MST, PBBS format):
  simd2_app's final matrix against the oracle's converge loop on the matrix
  the app builds, bit for bit, with the same squaring count.
* DAG_rmat_4096 (APLP) and minrp_rmat_4096 (MinRP): too large for a CPU
  loop here; simd2_app's reference loop and fused loop agree bit for bit and
  in count, and sampled rows of the result are a fixed point of the
  squaring (X = X (+) X (x) X on those rows, by the oracle; for APLP with
  the app's mapping of (-1, 1) to -FLT_MAX); MinRP's entries never exceed
  the input's."""
import gzip
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "graphs", "reference")
APP = os.path.join(ROOT, "simd2_amd", "simd2_app")
FLT_MAX = np.float32(3.4028234663852886e38)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def text(name):
    with gzip.open(os.path.join(DATA, name + ".gz"), "rt") as f:
        return f.read()


def run_app(app, name, tmp_path, loop="reference"):
    """simd2_app on a fixture: (matrix, squarings)."""
    dump = tmp_path / f"{app}_{name}_{loop}.bin"
    cmd = [APP, app, "-f"]
    stdin = None
    if app == "mst":
        src = tmp_path / name
        src.write_text(text(name))
        cmd.append(str(src))
    else:
        stdin = text(name)
    cmd += ["--reps", "1", "--loop", loop, "--dump", str(dump)]
    say(" ".join(cmd))
    r = subprocess.run(cmd, input=stdin, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    iters = int(r.stdout.split("iters")[1].split()[0])
    X = np.fromfile(dump, dtype=np.float32)
    n = int(round(np.sqrt(X.size)))
    return X.reshape(n, n), iters


def app_matrix(app, name):
    """The input matrix the reference's main builds from an edge list
    (fills as apps/<app>/cuASR/<app>_cuASR.cu; PBBS as apps/data/from_pbbs.h)."""
    lines = text(name).splitlines()
    if app == "mst":
        # edges in file order, each written both ways (a repeated pair keeps its last weight)
        e = [ln.split() for ln in lines[1:] if ln.strip()]
        v = max(max(int(t[0]), int(t[1])) for t in e) + 1
        X = np.zeros((v, v), np.float32)
        for t in e:
            a, b, w = int(t[0]), int(t[1]), np.float32(float(t[2]))
            X[a, b] = w
            X[b, a] = w
        off = ~np.eye(v, dtype=bool)
        X[off & (X == 0)] = FLT_MAX
        return X
    v, ne = map(int, lines[0].split())
    fill = -FLT_MAX if app in ("aplp", "mcp") else FLT_MAX
    X = np.full((v, v), fill, np.float32)
    for ln in lines[1:1 + ne]:  # file order: a repeated edge keeps its last weight
        t = ln.split()
        X[int(t[0]), int(t[1])] = np.float32(float(t[2]))
    if app == "apsp":
        np.fill_diagonal(X, 0)
    if app == "mcp":
        np.fill_diagonal(X, -FLT_MAX)
    return X


OP = {"apsp": "min_plus", "mcp": "max_min", "mst": "min_max", "aplp": "max_plus",
      "minrp": "min_mul"}
EPI = {"mst": False}
CHECK = {"aplp": "aplp"}


def test_rmat_4096_hop_diameter_and_path_count(gpu):
    """apps/data/dim_calc/4096: 'Path count: 16471467', 'Diameter: 6'."""
    import simd2_amd
    torch = gpu
    lines = text("rmat_4096.txt").splitlines()
    v, ne = map(int, lines[0].split())
    e = np.array([ln.split()[:2] for ln in lines[1:1 + ne]], dtype=np.int64)
    X = np.full((v, v), np.inf, np.float32)
    X[e[:, 0], e[:, 1]] = 1.0
    np.fill_diagonal(X, 0.0)
    say(f"rmat_4096 unit-weight APSP v={v} e={ne}")
    D, itr, st = simd2_amd.converge("min_plus", torch.from_numpy(X).cuda(), check="exact")
    torch.cuda.synchronize()
    D = D.cpu().numpy()
    finite = np.isfinite(D)
    assert st == 0 and itr == 4, (st, itr)
    assert int(D[finite].max()) == 6
    assert int(finite.sum()) - int(finite[0, 0]) == 16471467
    # the same reachability by bit-packed or_and (GTC's semiring, diagonal set)
    say("rmat_4096 transitive closure, packed or_and")
    A = torch.from_numpy((X < np.inf).astype(np.int32)).cuda()
    P = simd2_amd.pack_bits(A)
    _, itr_c, st = simd2_amd.converge("or_and", P)
    torch.cuda.synchronize()
    assert st == 0 and itr_c == 4, (st, itr_c)
    R = simd2_amd.unpack_bits(P, v).cpu().numpy().astype(bool)
    assert np.array_equal(R, finite)


@pytest.mark.parametrize("app,name", [("apsp", "parsed_rmat_1024.txt"),
                                      ("mcp", "parsed_rmat_1024.txt"),
                                      ("apsp", "parsed_rmat_2048.txt"),
                                      ("mst", "mst_rmat_1024.txt")])
def test_app_on_reference_input_vs_oracle(gpu, oracle_mod, tmp_path, app, name):
    X0 = app_matrix(app, name)
    G, itr = run_app(app, name, tmp_path)
    say(f"oracle converge {app} {name} n={X0.shape[0]}")
    W, itr_o = oracle_mod.converge_f32(OP[app], X0, epilogue=EPI.get(app, True),
                                       check=CHECK.get(app, "threshold"), threads=16)
    assert itr == itr_o, (itr, itr_o)
    assert np.array_equal(G.view(np.uint32), W.view(np.uint32))


@pytest.mark.parametrize("app,name", [("aplp", "DAG_rmat_4096.txt"),
                                      ("minrp", "minrp_rmat_4096.txt")])
def test_app_on_large_reference_input(gpu, oracle_mod, tmp_path, app, name):
    X0 = app_matrix(app, name)
    G, itr = run_app(app, name, tmp_path, "reference")
    F, itr_f = run_app(app, name, tmp_path, "fused")
    assert itr == itr_f and np.array_equal(G.view(np.uint32), F.view(np.uint32))
    n = X0.shape[0]
    rows = np.linspace(0, n - 1, 8).astype(np.int64)
    op = OP[app]
    say(f"fixed point of {app} {name} on {len(rows)} rows")
    sq = oracle_mod.srgemm_f32(op, G[rows], G, G[rows], True, threads=16)
    if app == "aplp":
        # the APLP check maps (-1, 1) to -FLT_MAX after each squaring
        # (kernel/converge.cuh:32-59)
        sq[(sq > -1) & (sq < 1)] = -FLT_MAX
    assert np.array_equal(sq.view(np.uint32), G[rows].view(np.uint32))
    if app == "minrp":  # min-based with the epilogue: entries only decrease
        assert np.all(G[rows] <= X0[rows])
