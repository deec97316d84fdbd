"""Scheduled converge loop (include/simd2/schedule.h) against the plain loop:
same matrix bit for bit and the same iteration count, whatever tile shape,
number of squarings enqueued ahead, or vertex order it is given, while
tiles and slabs are actually skipped."""
import pytest

import simd2_amd
from simd2_amd.graphs import er_family, grid_family, rgg_family
from simd2_amd.schedule import converge_scheduled

pytestmark = pytest.mark.gpu

OPS = [("min_plus", True, "threshold"), ("max_min", True, "threshold"),
       ("min_max", False, "threshold"), ("max_plus", True, "aplp")]


def family(torch, kind, op, n, seed):
    if kind == "er_sparse":
        return er_family(torch, op, n, 3.0 / n, seed)
    if kind == "er_dense":
        return er_family(torch, op, n, 0.05, seed)
    if kind == "grid":
        return grid_family(torch, op, n, seed)
    if kind == "rgg":
        return rgg_family(torch, op, n, 10, seed)
    raise ValueError(kind)


def dag(torch, X):
    """Longest paths need a DAG: keep the upper triangle only."""
    return torch.where(torch.triu(torch.ones_like(X, dtype=torch.bool), 1), X,
                       torch.full_like(X, float("-inf")))


def plain(torch, op, X, epi, check, max_itr=64):
    Y = X.clone()
    _, itr, st = simd2_amd.converge(op, Y, epilogue=epi, check=check, max_itr=max_itr)
    return Y, itr, st


@pytest.mark.parametrize("kind", ["er_sparse", "er_dense", "grid", "rgg"])
@pytest.mark.parametrize("n", [100, 257, 640])
def test_matches_plain_loop(gpu, kind, n):
    torch = gpu
    for k, (op, epi, check) in enumerate(OPS):
        X = family(torch, kind, op, n, 31 * n + k)
        if op == "max_plus":
            X = dag(torch, X)
        Y, itr, st = plain(torch, op, X, epi, check)
        assert st == simd2_amd.SUCCESS
        # (tile shape, squarings enqueued ahead, row-sparse limit, changed-entry
        # limit): the defaults, tiles only, every squaring row by row, every
        # squaring after the first over the changed entries
        changed_ok = epi and check != "aplp"
        for cfg, pred, rs, cs in ((-1, 0, -1, -1), (0, 3, 0, 0), (1, 64, n, -1),
                                  (2, itr, 0, n)):
            print(f"{kind} n={n} {op} cfg={cfg} pred={pred} row_sparse={rs} "
                  f"changed_sparse={cs}", flush=True)
            Z, itr2, st2, rep = converge_scheduled(op, X.clone(), epilogue=epi, check=check,
                                                   tile_config=cfg, predicted_itr=pred,
                                                   row_sparse=rs, changed_sparse=cs)
            assert st2 == simd2_amd.SUCCESS and itr2 == itr, (itr2, itr)
            assert torch.equal(Z.view(torch.int32), Y.view(torch.int32))
            assert rep["tiles_computed"] <= rep["tiles_total"]
            if rs == n:
                assert rep["rows_sparse"] == n * itr and rep["tiles_computed"] == 0
            if rs == 0:
                assert rep["rows_sparse"] == 0
            if cs == 0 or not changed_ok:
                assert rep["changed_rows"] == 0
            if cs == n and rs == 0:
                want = n * (itr - 1) if changed_ok else 0
                assert rep["changed_rows"] == want, rep
                if changed_ok:  # only the first squaring went tile by tile
                    assert rep["tiles_computed"] <= rep["tiles_total"] // itr
            if pred >= itr:
                assert rep["host_syncs"] == 1
            if pred == 0:
                assert rep["host_syncs"] == itr


def test_skipping_happens_on_banded_graphs(gpu):
    torch = gpu
    X = grid_family(torch, "min_plus", 2048, 5)
    Y, itr, _ = plain(torch, "min_plus", X, True, "threshold")
    Z, itr2, _, rep = converge_scheduled("min_plus", X.clone(), tile_config=1, row_sparse=0,
                                         changed_sparse=0)
    assert itr2 == itr and torch.equal(Z, Y)
    # the early squarings of a banded matrix touch few slabs
    assert rep["slab_tiles_computed"] < 0.6 * rep["slab_tiles_total"], rep
    # the automatic choice among tiled, row-sparse and changed-entry forms
    Z, itr3, _, rep = converge_scheduled("min_plus", X.clone(), tile_config=1)
    assert itr3 == itr and torch.equal(Z, Y), rep
    # the sparse forms wherever the list holds them (up to 1024 entries per
    # row): the early squarings row by row, the late ones over the changed
    # entries
    for rs, cs, key, least in ((2048, 0, "rows_sparse", 3), (0, 2048, "changed_rows", 2)):
        Z, itr3, _, rep = converge_scheduled("min_plus", X.clone(), tile_config=1,
                                             row_sparse=rs, changed_sparse=cs)
        assert itr3 == itr and torch.equal(Z, Y), rep
        assert rep[key] % 2048 == 0 and rep[key] >= least * 2048, rep


def test_vertex_order_does_not_change_result(gpu):
    torch = gpu
    n = 300
    X = rgg_family(torch, "min_plus", n, 8, 11)
    Y, itr, _ = plain(torch, "min_plus", X, True, "threshold")
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    perm = torch.randperm(n, generator=g, device="cuda").to(torch.int32)
    Z, itr2, _, _ = converge_scheduled("min_plus", X.clone(), perm=perm)
    assert itr2 == itr and torch.equal(Z, Y)


def test_cap_and_edge_sizes(gpu):
    torch = gpu
    n = 200
    path = torch.full((n, n), float("inf"), device="cuda")
    i = torch.arange(n - 1, device="cuda")
    path[i, i + 1] = 1.0
    path.fill_diagonal_(0.0)
    Y, itr, st = plain(torch, "min_plus", path, True, "threshold", max_itr=4)
    Z, itr2, st2, _ = converge_scheduled("min_plus", path.clone(), max_itr=4, predicted_itr=2)
    assert st == st2 == simd2_amd.ERROR_NOT_CONVERGED and itr == itr2 == 4
    assert torch.equal(Y, Z)
    # a workspace whose base is not 256-byte aligned is refused, X untouched
    need = simd2_amd.lib().simd2_converge_scheduled_workspace(n, -1, 4)
    ws = torch.empty(need + 256, dtype=torch.uint8, device="cuda")
    with pytest.raises(simd2_amd.Simd2Error) as err:
        converge_scheduled("min_plus", path, max_itr=4, workspace=ws[64:])
    assert err.value.status == simd2_amd.ERROR_MISALIGNED_OPERAND
    assert torch.equal(path[0, :3].cpu(), torch.tensor([0.0, 1.0, float("inf")]))
    for n in (1, 2, 15, 16, 17, 129):
        X = er_family(torch, "min_plus", n, 0.3, n)
        Y, itr, _ = plain(torch, "min_plus", X, True, "threshold")
        # automatic forms, then every squaring row by row, then over the
        # changed entries (X symmetrised so the symmetric form runs too)
        Xs = torch.minimum(X, X.t())
        Ys, itrs, _ = plain(torch, "min_plus", Xs, True, "threshold")
        for Xc, Yc, it, rs, cs in ((X, Y, itr, -1, -1), (X, Y, itr, n, 0), (X, Y, itr, 0, n),
                                   (Xs, Ys, itrs, 0, n)):
            Z, itr2, _, _ = converge_scheduled("min_plus", Xc.clone(), predicted_itr=5,
                                               row_sparse=rs, changed_sparse=cs)
            assert itr2 == it and torch.equal(Z, Yc), (n, rs, cs)


def test_sparse_forms_past_one_column_pass(gpu):
    """n above 8192: the sparse kernels take a row in more than one pass of
    8192 columns, the changed-entry scatter its list in more than one chunk,
    and the symmetrising pass ends on a partial tile pair."""
    torch = gpu
    n = 8192 + 300
    for kind, X in (("er", er_family(torch, "min_plus", n, 4.0 / n, 5)),
                    ("grid", grid_family(torch, "min_plus", n, 6))):
        Y, itr, st = plain(torch, "min_plus", X, True, "threshold")
        assert st == simd2_amd.SUCCESS
        for rs, cs, key in ((1024, 0, "rows_sparse"), (0, 1024, "changed_rows")):
            print(f"{kind} n={n} row_sparse={rs} changed_sparse={cs}", flush=True)
            Z, itr2, _, rep = converge_scheduled("min_plus", X.clone(), tile_config=0,
                                                 predicted_itr=itr, row_sparse=rs,
                                                 changed_sparse=cs)
            assert itr2 == itr and torch.equal(Z.view(torch.int32), Y.view(torch.int32)), rep
            assert rep[key] >= n, rep
        del Y, Z


@pytest.mark.parametrize("op", ["min_plus", "max_min"])
def test_exact_check_mode_every_form(gpu, op):
    """The exact change test (bit comparison) with each squaring form forced:
    tiles only, row by row, and over the changed entries (directed and
    symmetric inputs)."""
    torch = gpu
    n = 700
    for kind, X in (("er", er_family(torch, op, n, 4.0 / n, 17)),
                    ("grid", grid_family(torch, op, n, 18))):
        Y, itr, st = plain(torch, op, X, True, "exact")
        assert st == simd2_amd.SUCCESS
        for rs, cs in ((0, 0), (n, 0), (0, n), (-1, -1)):
            print(f"{kind} {op} exact row_sparse={rs} changed_sparse={cs}", flush=True)
            Z, itr2, _, rep = converge_scheduled(op, X.clone(), check="exact", row_sparse=rs,
                                                 changed_sparse=cs, predicted_itr=3)
            assert itr2 == itr and torch.equal(Z.view(torch.int32), Y.view(torch.int32)), rep


@pytest.mark.parametrize("kind", ["er_dense", "grid"])
def test_two_group_split_matches_plain_loop(gpu, kind):
    """n = 1600: 625 tiles of 64x64, between two and four per CU, so each
    scheduled squaring splits a tile's slab list over two thread groups (the
    n <= 640 cases above split over four); same matrix and count as the plain
    loop."""
    torch = gpu
    n = 1600
    for k, (op, epi, check) in enumerate(OPS[:2]):
        X = family(torch, kind, op, n, 7 * n + k)
        Y, itr, st = plain(torch, op, X, epi, check)
        assert st == simd2_amd.SUCCESS
        print(f"{kind} n={n} {op} two-group split", flush=True)
        Z, itr2, st2, rep = converge_scheduled(op, X.clone(), epilogue=epi, check=check,
                                               tile_config=1)
        assert st2 == simd2_amd.SUCCESS and itr2 == itr, (itr2, itr)
        assert torch.equal(Z.view(torch.int32), Y.view(torch.int32))
