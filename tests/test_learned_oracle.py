"""CPU checks that pin the learned-schedule oracle (oracle/learned.py):
gradients against finite differences, features against closed forms, and
breadth-first depths against scipy's unweighted shortest paths."""
import math

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

from oracle import learned as L

INF = np.float32(np.inf)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = L.init_params(1)
    P[L.OFF_B1:L.OFF_W2] = rng.normal(0, 0.1, L.NH)  # nonzero biases
    x = rng.normal(size=(13, L.NF))
    t = rng.normal(size=(13, L.NO))
    w = np.array([1.0, 0.5])
    loss, g = L.mlp_loss_grad(P, x, t, w)
    assert math.isclose(loss, L.mlp_loss(P, x, t, w), rel_tol=1e-12)
    idx = np.concatenate([rng.choice(L.NPARAM, 300, replace=False),
                          np.arange(L.OFF_B3, L.NPARAM), np.arange(L.OFF_B1, L.OFF_B1 + 4)])
    fd = L.finite_difference_grad(P, x, t, w, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_gradient_catches_a_dropped_term():
    """A plausible slip -- leaving the ReLU mask off the first layer --
    must fail the finite-difference comparison."""
    rng = np.random.default_rng(2)
    P = L.init_params(3)
    x = rng.normal(size=(9, L.NF))
    t = rng.normal(size=(9, L.NO))
    w = np.array([1.0, 1.0])
    W1, b1, W2, b2, W3, b3 = L.unpack(P)
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    y = h2 @ W3.T + b3
    dy = 2 * w * (y - t) / 9
    dh2 = (dy @ W3) * (h2 > 0)
    wrong = (dh2 @ W2).T @ x  # no (h1 > 0) mask
    idx = np.arange(L.OFF_W1, L.OFF_B1)
    fd = L.finite_difference_grad(P, x, t, w, idx)
    assert not np.allclose(wrong.ravel(), fd, rtol=1e-3, atol=1e-6)


def test_adam_first_step_moves_by_lr():
    # with bias correction the first step is lr * g / |g| per coordinate
    P = np.zeros(5)
    g = np.array([1.0, -2.0, 0.5, 3.0, -0.1])
    P1, m, v = L.adam_step(P, np.zeros(5), np.zeros(5), g, 1, 0.01)
    np.testing.assert_allclose(P1, -0.01 * np.sign(g), rtol=1e-6)


def path_graph(n):
    X = np.full((n, n), INF, np.float32)
    for i in range(n - 1):
        X[i, i + 1] = i + 1
    np.fill_diagonal(X, 0)
    return X


def test_path_graph_features_closed_form():
    n = 10
    f = L.graph_features("min_plus", path_graph(n), True)
    assert f[0] == pytest.approx(math.log2(10) / 16)
    assert f[1] == pytest.approx(9 / 90)
    assert f[3] == pytest.approx(math.log2(1 + 0.9) / 16)
    assert f[6] == pytest.approx(0.1) and f[7] == pytest.approx(0.1)
    assert f[9] == pytest.approx(5 / 9)            # mean weight 5 of max 9
    assert f[11] == pytest.approx(1 / 9)
    assert f[13] == 0.0                             # no reverse edges
    assert f[14] == pytest.approx(0.1) and f[15] == pytest.approx(0.1) and f[16] == 1.0
    # sources k*10//8 = 0,1,2,3,5,6,7,8: depth 9 - s, reached 10 - s
    srcs = [0, 1, 2, 3, 5, 6, 7, 8]
    assert f[17] == pytest.approx(math.log2(1 + np.mean([9 - s for s in srcs])) / 16)
    assert f[18] == pytest.approx(math.log2(10) / 16)
    assert f[19] == pytest.approx(np.mean([(10 - s) / 10 for s in srcs]))
    assert f[21] == 0.0 and f[22] == 1.0
    assert f[23] == pytest.approx((math.ceil(math.log2(9)) + 1) / 16)


def test_complete_and_star_graph_features():
    n = 12
    K = np.ones((n, n), np.float32) * 7
    np.fill_diagonal(K, 0)
    f = L.graph_features("max_min", K, True)
    assert f[1] == 1.0 and f[13] == 1.0 and f[18] == pytest.approx(1 / 16)
    assert f[4] == 0.0 and f[9] == 1.0 and f[10] == 0.0 and f[21] == 1.0
    S = np.full((n, n), -np.inf, np.float32)
    S[0, 1:] = 2
    S[1:, 0] = 2
    f = L.graph_features("max_min", S, False)
    m = 2 * (n - 1)
    assert f[1] == pytest.approx(m / (n * (n - 1)))
    assert f[5] == pytest.approx(math.log2(1 + n - 1) / 16)
    assert f[13] == 1.0 and f[22] == 0.0
    # from the centre depth 1, from a leaf depth 2
    assert f[18] == pytest.approx(math.log2(3) / 16)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bfs_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    n = 60
    E = rng.random((n, n)) < 0.04
    np.fill_diagonal(E, False)
    D = shortest_path(csr_matrix(E.astype(np.float64)), unweighted=True, directed=True)
    for src in (0, 17, 59):
        lv, depth, reached = L.bfs(E, src)
        finite = np.isfinite(D[src])
        assert reached == finite.sum()
        assert depth == int(D[src][finite].max())
        np.testing.assert_array_equal(lv[finite], D[src][finite].astype(np.int64))
        assert (lv[~finite] == -1).all()


def test_order_is_a_level_sorted_permutation():
    rng = np.random.default_rng(4)
    n = 80
    X = np.where(rng.random((n, n)) < 0.05, np.float32(3), INF).astype(np.float32)
    perm = L.order(X)
    assert sorted(perm.tolist()) == list(range(n))
    lv = L.levels(X)
    key = np.where(lv < 0, 10 ** 9, lv)[perm]
    assert (np.diff(key) >= 0).all()
