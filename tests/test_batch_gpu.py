"""Batched solver (include/simd2/batch.h) against the single-graph driver and
the oracle: every graph of a batch must come out bit-identical to
simd2_converge_* on the same graph, with the same iteration count."""
import numpy as np
import pytest

import simd2_amd
from simd2_amd.batch import Batch
from simd2_amd.graphs import app_graph, batch_specs
from util import FLT_MAX

pytestmark = pytest.mark.gpu


def single(torch, op, X, epi, check, max_itr=500):
    """Result and iteration count of the one-graph driver."""
    Y = X.clone()
    if op == "or_and":
        work = torch.zeros_like(Y)
        _, itr, st = simd2_amd.converge("or_and", Y, max_itr=max_itr, work=work)
    else:
        _, itr, st = simd2_amd.converge(op, Y, epilogue=epi, check=check, max_itr=max_itr)
    return Y, itr, st


def run_batch(torch, specs, mats, max_rounds=0, tune=False):
    b = Batch([(n, op, epi, check) for n, op, epi, check in specs], max_rounds=max_rounds)
    for g, X in enumerate(mats):
        b.upload(g, X)
    if tune:
        b.autotune()
        for g, X in enumerate(mats):  # autotuning leaves the inputs unchanged
            assert torch.equal(b.download(g), X)
    b.solve()
    itr, bad = b.iterations()
    outs = [b.download(g) for g in range(len(mats))]
    torch.cuda.synchronize()
    return b, outs, itr, bad


def edge_graph(torch, op, n, seed, p=0.08):
    """Graphs for every batched op, including those outside the app mix."""
    if op in ("min_plus", "max_min", "min_max", "or_and"):
        return app_graph(torch, op, n, p, seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    mask = torch.rand((n, n), generator=g, device="cuda") < p
    if op == "max_plus":       # longest paths on a DAG (APLP)
        mask = torch.triu(mask, 1)
        w = torch.randint(1, 101, (n, n), generator=g, device="cuda").float()
        return torch.where(mask, w, torch.full_like(w, -FLT_MAX))
    if op == "max_mul":        # reliabilities in (0, 1]
        w = torch.randint(1, 129, (n, n), generator=g, device="cuda").float() / 128.0
        X = torch.where(mask, w, torch.zeros_like(w))
        X.fill_diagonal_(1.0)
        return X
    if op == "min_mul":        # products of weights >= 1
        w = torch.randint(1, 10, (n, n), generator=g, device="cuda").float()
        X = torch.where(mask, w, torch.full_like(w, FLT_MAX))
        X.fill_diagonal_(1.0)
        return X
    raise ValueError(op)


EDGE_SIZES = [1, 2, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 200, 257]
EDGE_OPS = [("min_plus", True, "threshold"), ("max_plus", True, "aplp"),
            ("max_min", True, "threshold"), ("max_mul", True, "threshold"),
            ("min_mul", True, "threshold"), ("min_max", False, "threshold"),
            ("or_and", True, "exact"), ("min_plus", True, "exact")]


@pytest.mark.parametrize("tune", [False, True])
def test_mixed_edge_batch_matches_single_driver(gpu, tune):
    torch = gpu
    specs, mats = [], []
    for i, n in enumerate(EDGE_SIZES):
        for j, (op, epi, check) in enumerate(EDGE_OPS):
            if (i + j) % 3:      # a third of the (size, op) grid keeps the test short
                continue
            specs.append((n, op, epi, check))
            mats.append(edge_graph(torch, op, n, 1000 * i + j))
    specs.append((0, "min_plus", True, "threshold"))   # empty graph
    mats.append(torch.empty((0, 0), device="cuda"))
    print(f"batch of {len(specs)} graphs, tune={tune}", flush=True)
    b, outs, itr, bad = run_batch(torch, specs, mats, tune=tune)
    assert bad == 0
    for g, ((n, op, epi, check), X) in enumerate(zip(specs, mats)):
        if n == 0:
            assert itr[g] == 0
            continue
        Y, itr1, st = single(torch, op, X, epi, check)
        assert st == simd2_amd.SUCCESS
        assert itr[g] == itr1, (g, n, op)
        assert torch.equal(outs[g].view(torch.int32), Y.view(torch.int32)), (g, n, op)


def test_batch_vs_oracle(gpu, oracle_mod):
    torch = gpu
    specs, mats = [], []
    for k, (op, epi, check) in enumerate(EDGE_OPS):
        n = [33, 70, 129, 96, 50, 65, 100, 90][k]
        specs.append((n, op, epi, check))
        mats.append(edge_graph(torch, op, n, 77 + k))
    b, outs, itr, bad = run_batch(torch, specs, mats)
    for g, ((n, op, epi, check), X) in enumerate(zip(specs, mats)):
        if op == "or_and":
            T0 = simd2_amd.unpack_bits(X, n).cpu().numpy()
            To, itr_o = oracle_mod.converge_i32(T0)
            got = simd2_amd.unpack_bits(outs[g], n).cpu().numpy()
            assert np.array_equal(got, To) and itr[g] == itr_o
        else:
            Xo, itr_o = oracle_mod.converge_f32(op, X.cpu().numpy(), epilogue=epi, check=check,
                                                threads=8)
            assert itr[g] == itr_o, (op, n)
            assert np.array_equal(outs[g].cpu().numpy().view(np.uint32), Xo.view(np.uint32))


def test_round_cap_and_resolve(gpu):
    torch = gpu
    n = 64
    path = torch.full((n, n), float("inf"), device="cuda")
    idx = torch.arange(n - 1, device="cuda")
    path[idx, idx + 1] = 1.0
    path.fill_diagonal_(0.0)
    specs = [(n, "min_plus", True, "threshold"), (40, "or_and", True, "exact")]
    mats = [path, app_graph(torch, "or_and", 40, 0.05, 5)]
    b, outs, itr, bad = run_batch(torch, specs, mats, max_rounds=3)
    # ceil(log2(63)) + 1 = 7 squarings are needed: the cap stops the path graph
    assert bad >= 1 and itr[0] == 3
    Y, itr1, st = single(torch, "min_plus", path, True, "threshold", max_itr=3)
    assert st == simd2_amd.ERROR_NOT_CONVERGED and itr1 == 3
    assert torch.equal(outs[0], Y)
    # a converged batch solved again: one squaring each, results unchanged
    b2, outs2, itr2, bad2 = run_batch(torch, specs, mats)
    assert bad2 == 0 and itr2[0] == 7
    b2.solve()
    itr3, bad3 = b2.iterations()
    assert list(itr3) == [1, 1] and bad3 == 0
    for g in range(2):
        assert torch.equal(b2.download(g), outs2[g])


def test_config_mix_sample(gpu):
    """A 192-graph sample of the batched workload (n in [32, 512], p in
    [0.01, 0.2], 40/20/20/20 op mix), tuned, against the single-graph driver."""
    torch = gpu
    specs = batch_specs(192, seed=3)
    mats = [app_graph(torch, op, n, p, s) for n, op, epi, check, p, s in specs]
    b, outs, itr, bad = run_batch(torch, [s[:4] for s in specs], mats, tune=True)
    assert bad == 0
    cfg, launches = b.schedule()
    assert 1 <= launches <= 10
    for g, ((n, op, epi, check, p, s), X) in enumerate(zip(specs, mats)):
        Y, itr1, st = single(torch, op, X, epi, check)
        assert itr[g] == itr1, (g, n, op)
        assert torch.equal(outs[g].view(torch.int32), Y.view(torch.int32)), (g, n, op)
