"""Train the schedule MLP on runs the method generates itself.

    python -m simd2_amd.train_schedule --graphs 20000 --out simd2_amd/schedule_model.json

For every training graph (Erdos-Renyi with p in [0.0005, 0.05], 2D grids
with random diagonals, random geometric graphs with mean degree 8-64; APSP
min_plus or MCP max_min; n log-uniform in [256, 4096]) it computes the
features on the GPU, runs the scheduled converge loop once to learn the
number of squarings, then times the loop under each tile shape with that
number enqueued ahead; the labels are (squarings, fastest tile class).  The
network is trained on the GPU with the library's own forward / backward /
Adam kernels (simd2_mlp_train_step).  Writes the model as JSON (parameters,
feature normalisation, output scaling, training summary) and the data set as
.npz next to it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

FAMILIES = ("er", "grid", "rgg")
APPS = (("min_plus", True, "threshold"), ("max_min", True, "threshold"))


def graph_spec(rng, nmin, nmax):
    fam = FAMILIES[rng.integers(len(FAMILIES))]
    op, epi, check = APPS[rng.integers(len(APPS))]
    n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
    if fam == "er":
        param = float(math.exp(rng.uniform(math.log(0.0005), math.log(0.05))))
    elif fam == "grid":
        param = float(rng.uniform(0.1, 0.5))
    else:
        param = float(rng.uniform(8, 64))
    return fam, op, epi, check, n, param, int(rng.integers(1 << 30))


def make_graph(torch, fam, op, n, param, seed):
    from .graphs import er_family, grid_family, rgg_family
    if fam == "er":
        return er_family(torch, op, n, param, seed)
    if fam == "grid":
        return grid_family(torch, op, n, seed, diag_p=param)
    return rgg_family(torch, op, n, param, seed)


def label_runs(torch, specs, log_every=500, max_itr=64, checkpoint=None):
    """Features, squarings and per-tile-shape times of every graph
    (checkpoint(feats, itrs, times) is called at every progress line)."""
    from .learned import CLASS_OF_TILE, graph_features
    from .schedule import converge_scheduled
    feats, itrs, times = [], [], []
    ws = None
    t0 = time.time()
    for k, (fam, op, epi, check, n, param, seed) in enumerate(specs):
        X = make_graph(torch, fam, op, n, param, seed)
        feats.append(graph_features(op, X, epi).cpu().numpy())
        from .schedule import workspace_bytes
        need = max(workspace_bytes(n, c, max_itr) for c in range(3))
        if ws is None or ws.numel() < need:
            ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        _, itr, st, _ = converge_scheduled(op, X.clone(), epilogue=epi, check=check,
                                           max_itr=max_itr, workspace=ws)
        per_tile = np.zeros(3)
        for cfg in range(3):
            Y = X.clone()
            torch.cuda.synchronize()
            s = time.perf_counter()
            converge_scheduled(op, Y, epilogue=epi, check=check, max_itr=max_itr,
                               tile_config=cfg, predicted_itr=itr, workspace=ws)
            per_tile[CLASS_OF_TILE[cfg]] = time.perf_counter() - s
        itrs.append(itr)
        times.append(per_tile)
        if (k + 1) % log_every == 0:
            print(f"  labelled {k + 1}/{len(specs)} graphs ({time.time() - t0:.0f} s)",
                  flush=True)
            if checkpoint:
                checkpoint(np.array(feats, np.float32), np.array(itrs, np.int32),
                           np.array(times))
    return np.array(feats, np.float32), np.array(itrs, np.int32), np.array(times)


def init_params(rng):
    """He-uniform weights, zero biases (layout of include/simd2/learned.h)."""
    from .learned import NF, NH, NO, NPARAM
    P = np.zeros(NPARAM, np.float32)
    off = 0
    for fan_in, rows in ((NF, NH), (None, NH), (NH, NH), (None, NH), (NH, NO), (None, NO)):
        if fan_in is None:
            off += rows
            continue
        lim = math.sqrt(6.0 / fan_in)
        P[off:off + rows * fan_in] = rng.uniform(-lim, lim, rows * fan_in)
        off += rows * fan_in
    return P


def train(torch, feats, itrs, tile_cls, epochs=150, batch=256, lr=2e-3, seed=0,
          weights=(1.0, 0.5)):
    """Trains on the GPU; returns (params, normalisation dict, history)."""
    from .learned import Trainer, mlp_forward
    rng = np.random.default_rng(seed)
    n = feats.shape[0]
    idx = rng.permutation(n)
    n_val = max(1, n // 10)
    val, tr = idx[:n_val], idx[n_val:]
    mean = feats[tr].mean(0)
    sd = feats[tr].std(0)
    scale = 1.0 / np.maximum(sd, 1e-3)
    targets = np.stack([itrs.astype(np.float32), tile_cls.astype(np.float32)], 1)
    out_off = targets[tr].mean(0)
    out_scale = np.maximum(targets[tr].std(0), 1e-3)
    xn = ((feats - mean) * scale).astype(np.float32)
    tn = ((targets - out_off) / out_scale).astype(np.float32)
    dev = "cuda"
    X = torch.from_numpy(xn).to(dev)
    T = torch.from_numpy(tn).to(dev)
    W = torch.tensor(weights, dtype=torch.float32, device=dev)
    P = torch.from_numpy(init_params(rng)).to(dev)
    trainer = Trainer(P, lr=lr, max_batch=batch)
    tr_t = torch.from_numpy(tr).to(dev)
    history = []
    for ep in range(epochs):
        order = tr_t[torch.randperm(tr_t.numel(), device=dev)]
        tot = 0.0
        steps = 0
        for b0 in range(0, order.numel() - batch + 1, batch):
            sel = order[b0:b0 + batch]
            loss = trainer.step(X[sel].contiguous(), T[sel].contiguous(), W)
            if b0 == 0 or (b0 // batch) % 16 == 0:
                tot += float(loss.item())
                steps += 1
        if ep % 10 == 0 or ep == epochs - 1:
            y = mlp_forward(P, X).cpu().numpy() * out_scale + out_off
            pi = np.ceil(y[:, 0] + 0.25 - 1e-6)
            exact = float((np.round(y[val, 0]) == itrs[val]).mean())
            cover = float((pi[val] >= itrs[val]).mean())
            tacc = float((np.clip(np.round(y[val, 1]), 0, 2) == tile_cls[val]).mean())
            history.append({"epoch": ep, "train_loss": tot / max(steps, 1),
                            "val_itr_exact": exact, "val_itr_covered": cover,
                            "val_tile_acc": tacc})
            print(f"  epoch {ep}: loss {tot / max(steps, 1):.4f}  val itr exact {exact:.3f} "
                  f"covered {cover:.3f}  tile acc {tacc:.3f}", flush=True)
    norm = {"feature_mean": mean.tolist(), "feature_scale": scale.tolist(),
            "output_scale": out_scale.tolist(), "output_offset": out_off.tolist()}
    return P.cpu().numpy(), norm, history, val


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=20000)
    ap.add_argument("--nmin", type=int, default=256)
    ap.add_argument("--nmax", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "schedule_model.json"))
    ap.add_argument("--data", default=None, help="existing .npz data set to train on")
    ap.add_argument("--data-out", default=None, help="where to write the labelled data set")
    args = ap.parse_args(argv)
    import torch
    from . import lib
    lib()
    if args.data:
        z = np.load(args.data)
        feats, itrs, times = z["features"], z["squarings"], z["tile_times"]
    else:
        rng = np.random.default_rng(args.seed)
        specs = [graph_spec(rng, args.nmin, args.nmax) for _ in range(args.graphs)]
        print(f"labelling {len(specs)} method-generated runs", flush=True)
        t0 = time.time()

        def save(f, i, t):
            if args.data_out:
                np.savez_compressed(args.data_out, features=f, squarings=i, tile_times=t)
        feats, itrs, times = label_runs(torch, specs, checkpoint=save)
        print(f"labelled in {time.time() - t0:.0f} s", flush=True)
        save(feats, itrs, times)
    tile_cls = times.argmin(1).astype(np.int32)
    t0 = time.time()
    P, norm, history, val = train(torch, feats, itrs, tile_cls, epochs=args.epochs,
                                  seed=args.seed)
    print(f"trained in {time.time() - t0:.1f} s", flush=True)
    model = {"params": P.tolist(), **norm,
             "summary": {"graphs": int(feats.shape[0]), "validation_graphs": int(len(val)),
                         "n_range": [args.nmin, args.nmax], "families": list(FAMILIES),
                         "apps": [a[0] for a in APPS], "history": history,
                         "tile_class_counts": np.bincount(tile_cls, minlength=3).tolist(),
                         "squarings_range": [int(itrs.min()), int(itrs.max())]}}
    with open(args.out, "w") as f:
        json.dump(model, f)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
