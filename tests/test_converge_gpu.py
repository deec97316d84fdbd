"""GPU parity of the repeated-squaring drivers and the comp_update drop-ins."""
import hashlib
import sys

import numpy as np
import pytest

from util import APP_OP, FLT_MAX, app_matrix, er_graph, read_edges

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_converge(torch, op, X0, epi, check, flags, max_itr=500):
    import simd2_amd
    X = torch.from_numpy(np.ascontiguousarray(X0)).cuda()
    if op == "or_and" and flags == "packed":
        n = X0.shape[0]
        P = simd2_amd.pack_bits(X)
        P, itr, st = simd2_amd.converge("or_and", P, max_itr=max_itr)
        return simd2_amd.unpack_bits(P, n).cpu().numpy(), itr, st
    fl = simd2_amd.CONVERGE_FUSED if flags == "fused" else simd2_amd.CONVERGE_REFERENCE
    X, itr, st = simd2_amd.converge(op, X, epilogue=epi, check=check, flags=fl,
                                    max_itr=max_itr)
    return X.cpu().numpy(), itr, st


@pytest.mark.parametrize("flags", ["reference", "fused"])
def test_golden_graphs(gpu, golden_converge, flags):
    g = golden_converge
    for gname, fname in (("16_LD", "16_LD.txt"), ("rmat_16", "rmat_16.txt"),
                         ("rmat_256", "rmat_256.txt")):
        v, edges = read_edges(fname)
        for app, (op, epi) in APP_OP.items():
            key = f"{gname}/{app}"
            say(f"converge {key} {flags}")
            X0 = app_matrix(v, edges, app)
            X, itr, st = gpu_converge(gpu, op, X0, epi, "threshold", flags)
            assert st == 0
            assert itr == int(g[f"{key}/itr"][0]), key
            assert hashlib.sha256(X.tobytes()).digest() == g[f"{key}/sha256"].tobytes(), key
            if op == "or_and":
                Xp, itp, _ = gpu_converge(gpu, op, X0, epi, "exact", "packed")
                assert itp == itr and np.array_equal(Xp, X), key


def test_path_graph_1024_closed_form(gpu):
    v, edges = read_edges("1024_LD.txt")
    X0 = app_matrix(v, edges, "apsp")
    w = np.zeros(v)
    for u, t, wt in edges:
        w[t] = wt
    pref = np.cumsum(w)
    i, j = np.indices((v, v))
    want = np.where(i == j, 0.0, np.where(j > i, pref[j] - pref[i], FLT_MAX)).astype(np.float32)
    for flags in ("reference", "fused"):
        say(f"path 1024 {flags}")
        X, itr, st = gpu_converge(gpu, "min_plus", X0, True, "threshold", flags)
        assert st == 0 and itr == 11  # ceil(log2(1023)) + 1
        assert np.array_equal(X, want)


def test_apsp_er1024_vs_dijkstra_and_oracle(gpu, oracle_mod):
    """The n=1024, p=0.05, seed-0 APSP workload: every row equals Dijkstra,
    and the squaring count equals the oracle loop's."""
    from scipy.sparse.csgraph import dijkstra
    X0 = er_graph(1024, 0.05, seed=0)
    say("apsp er1024")
    X, itr, st = gpu_converge(gpu, "min_plus", X0, True, "exact", "fused")
    assert st == 0
    W = np.where(np.isinf(X0), 0, X0).astype(np.float64)
    assert np.array_equal(X.astype(np.float64), dijkstra(W, directed=True))
    Xr, itr_r, _ = gpu_converge(gpu, "min_plus", X0, True, "threshold", "reference")
    assert np.array_equal(Xr, X) and itr_r == itr
    Xo, itr_o = oracle_mod.converge_f32("min_plus", X0, check="threshold", threads=16)
    assert itr_o == itr and np.array_equal(Xo, X)


def test_mcp_mst_aplp_small(gpu, oracle_mod):
    dag = er_graph(150, 0.05, 5, fill=-FLT_MAX, diag=-FLT_MAX)
    dag[np.tril_indices(150, -1)] = -FLT_MAX          # APLP on a DAG (edges i < j)
    rng_graphs = [
        ("max_min", er_graph(200, 0.05, 3, fill=-np.inf, diag=-np.inf), True, "exact"),
        ("min_max", np.minimum(er_graph(200, 0.05, 4), er_graph(200, 0.05, 4).T), False,
         "exact"),
        ("max_plus", dag, True, "aplp"),
        # reliabilities in (0, 1): products stay far above the FP32 subnormal range
        ("max_mul", np.where(np.eye(130) > 0, 0, er_graph(130, 0.05, 6, fill=0) / 128),
         True, "exact"),
        ("min_mul", er_graph(130, 0.05, 7, wlo=1, whi=9, fill=np.inf, diag=1.0), True,
         "exact"),
    ]
    for op, X0, epi, check in rng_graphs:
        X0 = X0.astype(np.float32)
        for flags in ("reference", "fused"):
            say(f"converge {op} {flags}")
            X, itr, st = gpu_converge(gpu, op, X0, epi, check, flags, max_itr=40)
            Xo, itr_o = oracle_mod.converge_f32(op, X0, epilogue=epi, check=check, max_itr=40,
                                                threads=8)
            assert itr == itr_o, (op, flags)
            assert np.array_equal(X.view(np.uint32), Xo.view(np.uint32)), (op, flags)


def test_not_converged_status(gpu):
    import simd2_amd
    X0 = np.array([[0, 1], [1, 0]], np.float32)  # positive cycle under max_plus
    X, itr, st = gpu_converge(gpu, "max_plus", X0, True, "exact", "fused", max_itr=6)
    assert st == simd2_amd.ERROR_NOT_CONVERGED and itr == 6


def test_comp_update_dropins(gpu):
    import torch
    import simd2_amd
    inf = float("inf")
    prev = torch.tensor([[inf, inf], [1.0, 2.0]], device="cuda")
    cur = torch.tensor([[inf, 5.0], [1.00005, 2.0]], device="cuda")
    chk = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert simd2_amd.comp_update(prev, cur, chk)
    assert torch.equal(prev, cur)
    prev = torch.tensor([[inf, 1.0]], device="cuda")
    cur = torch.tensor([[inf, 1.00005]], device="cuda")
    assert not simd2_amd.comp_update(prev, cur, chk)
    prev = torch.tensor([[-FLT_MAX, 3.0]], device="cuda")
    cur = torch.tensor([[0.5, 3.0]], device="cuda")
    assert not simd2_amd.comp_update(prev, cur, chk, mode="aplp")
    assert cur[0, 0].item() == -FLT_MAX
    pi = torch.tensor([[1, 2]], dtype=torch.int32, device="cuda")
    ci = torch.tensor([[1, 3]], dtype=torch.int32, device="cuda")
    assert simd2_amd.comp_update(pi, ci, chk)
    assert torch.equal(pi, ci)


@pytest.mark.parametrize("n", [64, 65, 97, 300, 1000])
def test_closure_i32_any_values_and_stride(gpu, oracle_mod, n):
    """simd2_converge_i32 on a strided int32 matrix whose entries are not only
    0/1 (the reference's bin_or/bin_and treat any nonzero as true): both loop
    structures give the oracle's closure and squaring count, also when
    max_itr stops the loop early (the matrix after max_itr squarings)."""
    import simd2_amd
    rng = np.random.default_rng(n)
    vals = np.array([0, 1, 7, -2], np.int32)
    X0 = np.where(rng.random((n, n)) < 2.0 / n, rng.choice(vals[1:], (n, n)), 0).astype(np.int32)
    X0[rng.integers(0, n, n // 4), rng.integers(0, n, n // 4)] = -2
    Xo, itr_o = oracle_mod.converge_i32(X0)
    assert itr_o >= 3
    for flags in (simd2_amd.CONVERGE_REFERENCE, simd2_amd.CONVERGE_FUSED):
        for max_itr in (1, 2, itr_o - 1, 500):
            say(f"closure n={n} flags={flags} max_itr={max_itr}")
            want, it_w = (Xo, itr_o) if max_itr == 500 else oracle_mod.converge_i32(X0, max_itr)
            wide = gpu.full((n, n + 37), 5, dtype=gpu.int32, device="cuda")
            wide[:, :n] = gpu.from_numpy(X0).cuda()
            X, itr, st = simd2_amd.converge("or_and", wide[:, :n], flags=flags, max_itr=max_itr)
            assert itr == it_w, (flags, max_itr)
            assert st == (0 if max_itr == 500 else simd2_amd.ERROR_NOT_CONVERGED)
            assert np.array_equal(X.cpu().numpy(), want), (flags, max_itr)
            assert bool((wide[:, n:] == 5).all()), "wrote past the row"
