"""Held-out evaluation of the learned schedule at n = 8192.

    python -m simd2_amd.eval_schedule --graphs 200 --n 8192 --out results.json

Graphs: Erdos-Renyi (p in [0.0005, 0.05]), 2D grids with random diagonals
(weights in [1, 1000]) and random geometric graphs (mean degree 8-64), in
turn; APSP (min_plus) and MCP (max_min) alternate.  Seeds come from a stream
the trainer never draws from.  For each graph it times

  plain    the reference's loop structure (simd2_converge_f32 with
           SIMD2_CONVERGE_REFERENCE: full squaring, full change-test pass,
           one flag read per squaring); plain_fused: the same loop with the
           change test fused into the squaring;
  learned  simd2_converge_learned (features, prediction, scheduled loop);
  method   the scheduled loop without the network: same features and vertex
           order rule, automatic tile shape, a flag read after every squaring;

and checks that learned and method return the plain loop's matrix bit for bit
with the same number of squarings.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

HELD_OUT_SEED = 1_000_003
FAMILIES = ("er", "grid", "rgg")
APPS = (("min_plus", True, "threshold"), ("max_min", True, "threshold"))


def held_out_specs(count, seed=HELD_OUT_SEED):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        fam = FAMILIES[k % 3]
        op, epi, check = APPS[(k // 3) % 2]
        if fam == "er":
            param = float(math.exp(rng.uniform(math.log(0.0005), math.log(0.05))))
        elif fam == "grid":
            param = float(rng.uniform(0.1, 0.5))
        else:
            param = float(rng.uniform(8, 64))
        out.append((fam, op, epi, check, param, int(rng.integers(1 << 30))))
    return out


def reorder_rule(f, n):
    """The vertex-order rule of simd2_converge_learned: relabel by search
    levels when the numbering has no locality, a search reaches most of the
    graph and the mean search depth exceeds log2 n."""
    depth = 2.0 ** (16.0 * float(f[17])) - 1.0
    return bool(f[14] > 0.05 and f[19] > 0.5 and depth > math.log2(n))


def _timed(torch, fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) * 1e3, r


def evaluate(torch, count=200, n=8192, model=None, log=None, max_itr=64):
    import simd2_amd
    from .learned import converge_learned, graph_features, graph_order, load_model
    from .schedule import converge_scheduled
    from .train_schedule import make_graph
    if model is None:
        model, _ = load_model()
    L = simd2_amd.learned._bind()
    ws_l = torch.empty(L.simd2_converge_learned_workspace(n, max_itr), dtype=torch.uint8,
                       device="cuda")
    from .schedule import workspace_bytes
    ws_s = torch.empty(workspace_bytes(n, -1, max_itr), dtype=torch.uint8, device="cuda")
    rows = []
    specs = held_out_specs(count)
    # one untimed call of every loop first (kernel loading, first-use set-up)
    for fam, op, epi, check, param, seed in specs[:2]:
        X = make_graph(torch, fam, op, n, param, seed)
        simd2_amd.converge(op, X.clone(), epilogue=epi, check=check, max_itr=max_itr,
                           flags=simd2_amd.CONVERGE_REFERENCE)
        simd2_amd.converge(op, X.clone(), epilogue=epi, check=check, max_itr=max_itr)
        converge_learned(op, X.clone(), model, epilogue=epi, check=check, max_itr=max_itr,
                         workspace=ws_l)
        converge_scheduled(op, X.clone(), epilogue=epi, check=check, max_itr=max_itr,
                           workspace=ws_s)
        del X
    for k, (fam, op, epi, check, param, seed) in enumerate(specs):
        X = make_graph(torch, fam, op, n, param, seed)
        Y = X.clone()
        t_plain, (_, itr, st) = _timed(torch, lambda: simd2_amd.converge(
            op, Y, epilogue=epi, check=check, max_itr=max_itr,
            flags=simd2_amd.CONVERGE_REFERENCE))
        Yf = X.clone()
        t_fused, (_, itr_f, st_f) = _timed(torch, lambda: simd2_amd.converge(
            op, Yf, epilogue=epi, check=check, max_itr=max_itr))
        Z = X.clone()
        t_learn, (_, itr_l, st_l, sched, rep) = _timed(torch, lambda: converge_learned(
            op, Z, model, epilogue=epi, check=check, max_itr=max_itr, workspace=ws_l))
        W = X.clone()

        def method():
            f = graph_features(op, W, epi).cpu().numpy()
            perm = graph_order(op, W) if reorder_rule(f, n) else None
            return converge_scheduled(op, W, epilogue=epi, check=check, max_itr=max_itr,
                                      perm=perm, workspace=ws_s)
        t_meth, (_, itr_m, st_m, rep_m) = _timed(torch, method)
        exact = bool(torch.equal(Z.view(torch.int32), Y.view(torch.int32)) and
                     torch.equal(W.view(torch.int32), Y.view(torch.int32)) and
                     torch.equal(Yf.view(torch.int32), Y.view(torch.int32)))
        row = {"family": fam, "param": round(param, 5), "op": op, "squarings": itr,
               "plain_ms": round(t_plain, 2), "plain_fused_ms": round(t_fused, 2),
               "learned_ms": round(t_learn, 2),
               "method_ms": round(t_meth, 2), "bit_exact": exact,
               "same_squarings": itr == itr_l == itr_m == itr_f,
               "predicted": sched[1], "tile": sched[0], "learned_syncs": rep["host_syncs"],
               "method_syncs": rep_m["host_syncs"] + 1,
               "slab_tiles_frac": round(rep["slab_tiles_computed"] /
                                        max(rep["slab_tiles_total"], 1), 4)}
        rows.append(row)
        if log:
            log(json.dumps(row))
        del X, Y, Yf, Z, W
    return rows


def summarize(rows):
    def tot(key, sel=lambda r: True):
        return sum(r[key] for r in rows if sel(r))
    out = {"graphs": len(rows), "bit_exact": all(r["bit_exact"] for r in rows),
           "same_squarings": all(r["same_squarings"] for r in rows),
           "plain_ms": round(tot("plain_ms"), 1),
           "plain_fused_ms": round(tot("plain_fused_ms"), 1),
           "learned_ms": round(tot("learned_ms"), 1),
           "method_ms": round(tot("method_ms"), 1)}
    out["speedup_learned_vs_plain"] = round(out["plain_ms"] / out["learned_ms"], 3)
    out["speedup_learned_vs_plain_fused"] = round(out["plain_fused_ms"] / out["learned_ms"], 3)
    out["speedup_method_vs_plain"] = round(out["plain_ms"] / out["method_ms"], 3)
    out["gain_of_network"] = round(out["method_ms"] / out["learned_ms"], 3)
    per = {}
    for fam in FAMILIES:
        for op in ("min_plus", "max_min"):
            sel = [r for r in rows if r["family"] == fam and r["op"] == op]
            if sel:
                p = sum(r["plain_ms"] for r in sel)
                l = sum(r["learned_ms"] for r in sel)
                m = sum(r["method_ms"] for r in sel)
                per[f"{fam}/{op}"] = {"graphs": len(sel), "plain_ms": round(p, 1),
                                      "learned_ms": round(l, 1), "method_ms": round(m, 1),
                                      "speedup": round(p / l, 3)}
    out["per_family"] = per
    out["predicted_exact"] = round(float(np.mean([r["predicted"] == r["squarings"]
                                                  for r in rows])), 3)
    out["predicted_covers"] = round(float(np.mean([r["predicted"] >= r["squarings"]
                                                   for r in rows])), 3)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=200)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    import torch
    import simd2_amd
    simd2_amd.lib()
    rows = evaluate(torch, args.graphs, args.n, log=lambda s: print(s, flush=True))
    summary = summarize(rows)
    print(json.dumps(summary), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"summary": summary, "rows": rows}, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
