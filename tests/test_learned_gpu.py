"""The learned schedule on the GPU against its oracle (oracle/learned.py):
features, vertex order, the MLP's forward pass, gradient and Adam step; and
the learned converge loop against the plain loop, bit for bit, whatever the
network predicts."""
import os

import numpy as np
import pytest

import simd2_amd
from simd2_amd import learned as LG
from simd2_amd.graphs import er_family, grid_family, rgg_family

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oracle_learned():
    from oracle import learned
    return learned


def graphs(torch, sizes=(70, 301, 1000)):
    out = []
    for k, n in enumerate(sizes):
        out.append(("min_plus", True, er_family(torch, "min_plus", n, 4.0 / n, k)))
        out.append(("max_min", True, rgg_family(torch, "max_min", n, 9, k)))
        out.append(("min_max", False, grid_family(torch, "min_max", n, k)))
    return out


def test_features_and_order_match_oracle(gpu, oracle_learned):
    """Both search forms: adjacency in LDS up to n = 1024 (32 words a row),
    in global memory above."""
    torch = gpu
    for op, epi, X in graphs(torch, (70, 301, 1000, 1024, 1025, 1500)):
        print("features", op, X.shape[0], flush=True)
        f = LG.graph_features(op, X, epi).cpu().numpy()
        fo = oracle_learned.graph_features(op, X.cpu().numpy(), epi)
        np.testing.assert_allclose(f, fo, rtol=2e-6, atol=1e-7)
        perm = LG.graph_order(op, X).cpu().numpy()
        np.testing.assert_array_equal(perm, oracle_learned.order(X.cpu().numpy()))


def test_mlp_forward_and_gradient_match_oracle(gpu, oracle_learned):
    torch = gpu
    rng = np.random.default_rng(5)
    P = oracle_learned.init_params(7).astype(np.float32)
    P[oracle_learned.OFF_B1:oracle_learned.OFF_W2] = rng.normal(0, 0.1, 64)
    for B in (1, 37, 256):
        x = rng.normal(size=(B, 24)).astype(np.float32)
        t = rng.normal(size=(B, 2)).astype(np.float32)
        w = np.array([1.0, 0.5], np.float32)
        Pd, xd, td, wd = (torch.from_numpy(a).cuda() for a in (P, x, t, w))
        y = LG.mlp_forward(Pd, xd).cpu().numpy()
        np.testing.assert_allclose(y, oracle_learned.mlp_forward(P, x), rtol=1e-4, atol=1e-5)
        loss, g = LG.mlp_gradient(Pd, xd, td, wd)
        lo, go = oracle_learned.mlp_loss_grad(P, x, t, w)
        assert abs(float(loss.item()) - lo) <= 1e-4 * abs(lo)
        g = g.cpu().numpy()
        scale = np.abs(go).max()
        np.testing.assert_allclose(g, go, rtol=0, atol=2e-5 * scale)


def test_adam_steps_match_oracle_and_learn(gpu, oracle_learned):
    torch = gpu
    rng = np.random.default_rng(9)
    B = 128
    x = rng.normal(size=(B, 24)).astype(np.float32)
    A = rng.normal(size=(24, 2)) / 5
    t = (x @ A + 0.3).astype(np.float32)
    w = np.array([1.0, 1.0], np.float32)
    P0 = oracle_learned.init_params(11).astype(np.float32)
    Pd = torch.from_numpy(P0.copy()).cuda()
    tr = LG.Trainer(Pd, lr=1e-3, max_batch=B)
    xd, td, wd = (torch.from_numpy(a).cuda() for a in (x, t, w))
    P, m, v = P0.astype(np.float64), np.zeros(P0.size), np.zeros(P0.size)
    for step in range(1, 4):
        tr.step(xd, td, wd)
        _, g = oracle_learned.mlp_loss_grad(P, x, t, w)
        P, m, v = oracle_learned.adam_step(P, m, v, g, step, 1e-3)
    np.testing.assert_allclose(Pd.cpu().numpy(), P, rtol=0, atol=2e-5)
    first = oracle_learned.mlp_loss(P0, x, t, w)
    for _ in range(400):
        loss = tr.step(xd, td, wd)
    assert float(loss.item()) < 0.1 * first


def plain(torch, op, X, epi):
    Y = X.clone()
    _, itr, st = simd2_amd.converge(op, Y, epilogue=epi, check="threshold", max_itr=64)
    return Y, itr


@pytest.mark.parametrize("which", ["random", "trained"])
def test_learned_loop_is_bit_exact(gpu, oracle_learned, which):
    torch = gpu
    if which == "trained":
        if not os.path.exists(LG.MODEL_PATH):
            pytest.skip("no trained model file")
        model, _ = LG.load_model()
    else:  # an untrained network: the schedule is arbitrary, the result is not
        rng = np.random.default_rng(1)
        vec = np.concatenate([oracle_learned.init_params(2), rng.normal(0, 1, 48),
                              np.array([3.0, 1.0, 4.0, 1.0])]).astype(np.float32)
        model = torch.from_numpy(vec).cuda()
    for op, epi, X in graphs(torch):
        Y, itr = plain(torch, op, X, epi)
        Z, itr2, st, sched, rep = LG.converge_learned(op, X.clone(), model, epilogue=epi)
        print(op, X.shape[0], itr, itr2, sched, rep["host_syncs"], flush=True)
        assert st == simd2_amd.SUCCESS and itr2 == itr
        assert torch.equal(Z.view(torch.int32), Y.view(torch.int32))


def test_trained_network_replaces_flag_reads(gpu):
    """What the trained network buys, measured without a clock: on held-out
    graphs (the evaluation's generator and seed stream, n = 768) its predicted
    squaring count lets the loop read the convergence flags once or twice where
    the reference's loop structure reads them after every squaring, and the
    prediction is never below the true count on most graphs; the results stay
    bit-identical."""
    torch = gpu
    if not os.path.exists(LG.MODEL_PATH):
        pytest.skip("no trained model file")
    from simd2_amd.eval_schedule import held_out_specs
    from simd2_amd.train_schedule import make_graph
    model, _ = LG.load_model()
    learned_reads = plain_reads = covered = 0
    specs = held_out_specs(12)
    for fam, op, epi, check, param, seed in specs:
        X = make_graph(torch, fam, op, 768, param, seed)
        Y, itr = plain(torch, op, X, epi)
        Z, itr2, st, (tile, pred), rep = LG.converge_learned(op, X.clone(), model, epilogue=epi)
        print(fam, op, itr, pred, rep["host_syncs"], flush=True)
        assert st == simd2_amd.SUCCESS and itr2 == itr
        assert torch.equal(Z.view(torch.int32), Y.view(torch.int32))
        learned_reads += rep["host_syncs"]
        plain_reads += itr
        covered += int(pred >= itr)
    assert covered >= 0.8 * len(specs), covered
    assert learned_reads <= 0.4 * plain_reads, (learned_reads, plain_reads)
