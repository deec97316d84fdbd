"""The benchmark's own configurations, at full size, against the oracle.

Inputs come from bench.py's generators with bench.py's seeds, so the kernels
checked here are the ones the benchmark times, in the tile shapes the
automatic rule picks at those sizes (128x64 at 4096^3, 128x128 at 8192^3):

* the nine operators at M = N = K = 4096 -- sampled rows of every result
  against the oracle's rows (kernel/host_srgemm.h:34-48 restated; bit for bit,
  plus_mul within 1e-5 relative of FP64), every VALU operator also forced
  through each tile configuration and compared whole with the automatic one,
  or_and on the packed and the int32 drop-in forms, minus_square's int8 path
  against its direct kernel;
* min_plus at 8192^3 (128x128 tiles), sampled rows;
* the learned-schedule loop at n = 2048 and 4096, where the sparse squaring
  forms engage, against the plain converge loop on each graph family.

Rows are independent, so a sampled row is checked exactly: the oracle
computes D[r, :] from A[r, :] and all of B."""
import sys

import numpy as np
import pytest

from bench import OPS, make_inputs

pytestmark = pytest.mark.gpu

VALU_OPS = ["min_plus", "max_plus", "max_min", "max_mul", "min_mul", "min_max"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def sample_rows(n, count=8):
    """Rows spread over the matrix, including both sides of a tile edge."""
    rows = set(np.linspace(0, n - 1, count - 2).astype(int).tolist()) | {127, 128}
    return np.array(sorted(rows))


def rows_of(t, rows):
    import torch
    return t[torch.from_numpy(rows).to(t.device)].cpu().numpy()


def check_rows(oracle_mod, op, A, B, C, D, rows):
    """D's sampled rows against the oracle on the same rows (A, B, C, D are
    device tensors of the whole problem)."""
    a, c, d = rows_of(A, rows), rows_of(C, rows), rows_of(D, rows)
    b = B.cpu().numpy()
    if op == "plus_mul":
        want = a.astype(np.float64) @ b.astype(np.float64) + c.astype(np.float64)
        rel = np.max(np.abs(d - want) / np.maximum(np.abs(want), 1e-30))
        assert rel <= 1e-5, (op, rel)
        return
    want = oracle_mod.srgemm_f32(op, a, b, c, True, threads=16)
    assert np.array_equal(bits(d), bits(want)), op


@pytest.mark.parametrize("op", [o for o in OPS if o != "or_and"])
def test_bench_op_4096_against_oracle(gpu, oracle_mod, op):
    import simd2_amd
    torch = gpu
    n = 4096
    seed = 1000 + OPS.index(op)
    A, B, C = make_inputs(torch, op, n, seed)
    if op in VALU_OPS:
        # the configuration the benchmark runs: 128x64 tiles at 4096^3
        assert simd2_amd.lib().simd2_pick_tile_config(OPS.index(op), n, n, n) == 2
    say(f"bench config {op} {n}^3 seed {seed} (drop-in)")
    D = simd2_amd.srgemm(op, A, B, C, True)
    torch.cuda.synchronize()
    check_rows(oracle_mod, op, A, B, C, D, sample_rows(n))
    if op == "minus_square":
        say(f"bench config {op} {n}^3 int8 path vs the direct kernel")
        E, ws = simd2_amd.minus_square_exact(A, B, C, True)
        torch.cuda.synchronize()
        assert torch.equal(E.view(torch.int32), D.view(torch.int32))
        mf, va = simd2_amd.minus_square_tiles(ws, n, n, n)
        assert (mf, va) == ((n // 64) ** 2, 0)
    if op in VALU_OPS or op == "minus_square":
        for cfg in range(simd2_amd.lib().simd2_tile_config_count()):
            say(f"bench config {op} {n}^3 forced tile config {cfg}")
            E = simd2_amd.srgemm(op, A, B, C, True, tile_config=cfg)
            torch.cuda.synchronize()
            assert torch.equal(E.view(torch.int32), D.view(torch.int32)), (op, cfg)


def test_bench_or_and_4096_packed_and_dropin(gpu):
    """or_and on Bernoulli(0.02) bits: the packed kernel's sampled rows
    against a boolean product, and the int32 drop-in equal to it whole."""
    import simd2_amd
    torch = gpu
    n = 4096
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 + OPS.index("or_and"))
    bits01 = [(torch.rand((n, n), generator=g, device="cuda") < 0.02).to(torch.int32)
              for _ in range(3)]
    A, B, C = (simd2_amd.pack_bits(x) for x in bits01)
    D = torch.empty_like(C)
    nw = A.shape[1]
    say(f"bench config or_and {n}^3 packed")
    st = simd2_amd.lib().simd2_orand_packed(n, n, n, A.data_ptr(), nw, B.data_ptr(), nw,
                                            C.data_ptr(), nw, D.data_ptr(), nw, 1,
                                            torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    rows = sample_rows(n)
    got = simd2_amd.unpack_bits(D, n)
    a = rows_of(bits01[0], rows).astype(np.int64)
    b = bits01[1].cpu().numpy().astype(np.int64)
    want = ((a @ b) > 0) | (rows_of(bits01[2], rows) > 0)
    assert np.array_equal(rows_of(got, rows) > 0, want)
    say(f"bench config or_and {n}^3 int32 drop-in")
    Di = simd2_amd.srgemm("or_and", bits01[0], bits01[1], bits01[2], True)
    torch.cuda.synchronize()
    assert torch.equal(Di, got.to(torch.int32))


def test_bench_min_plus_8192_against_oracle(gpu, oracle_mod):
    import simd2_amd
    torch = gpu
    n = 8192
    assert simd2_amd.lib().simd2_pick_tile_config(0, n, n, n) == 0  # 128x128 tiles
    A, B, C = make_inputs(torch, "min_plus", n, 77)
    say(f"bench config min_plus {n}^3 seed 77")
    D = simd2_amd.srgemm("min_plus", A, B, C, True)
    torch.cuda.synchronize()
    check_rows(oracle_mod, "min_plus", A, B, C, D, sample_rows(n))


@pytest.mark.parametrize("n", [2048, 4096])
@pytest.mark.parametrize("fam,param", [("er", 0.002), ("grid", 0.3), ("rgg", 16.0)])
def test_learned_loop_where_sparse_forms_engage(gpu, n, fam, param):
    """simd2_converge_learned with the trained network at sizes where the
    automatic sparse squarings run, against the plain loop bit for bit with
    the same squaring count."""
    import simd2_amd
    from simd2_amd.learned import converge_learned, load_model
    from simd2_amd.train_schedule import make_graph
    torch = gpu
    model, _ = load_model()
    for op in ("min_plus", "max_min"):
        seed = 4242 + n + len(fam)
        X = make_graph(torch, fam, op, n, param, seed)
        say(f"learned loop {fam} {op} n={n} param={param} seed={seed}")
        Y = X.clone()
        _, itr, st = simd2_amd.converge(op, Y, epilogue=True, check="threshold", max_itr=64,
                                        flags=simd2_amd.CONVERGE_REFERENCE)
        Z = X.clone()
        _, itr_l, st_l, sched, rep = converge_learned(op, Z, model, epilogue=True,
                                                      check="threshold", max_itr=64)
        torch.cuda.synchronize()
        assert st == 0 and st_l == 0
        assert itr_l == itr, (itr_l, itr)
        assert torch.equal(Z.view(torch.int32), Y.view(torch.int32))
        say(f"  squarings {itr}, predicted {sched[1]}, sparse rows {rep['rows_sparse']}, "
            f"changed rows {rep['changed_rows']}, tiles {rep['tiles_computed']}/"
            f"{rep['tiles_total']}")
