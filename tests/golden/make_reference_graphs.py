"""Copies the reference's own graph inputs into tests/golden/graphs/reference/
as gzip files (data fixtures; nothing of the reference is executed).

    python tests/golden/make_reference_graphs.py [/root/reference]

The files are the inputs the reference's report script runs its apps on
(scripts/run_apps.py:100-361) that exist in its tree, plus the graph whose
hop diameter and reachable-pair count the reference records
(apps/data/dim_calc/4096: v = 4096, e = 167772, path count 16471467,
diameter 6, computed by apps/data/dim_calc/dim_cacl.cpp from
apps/data/rmat_data/rmat_4096.txt).
"""
import gzip
import os
import shutil
import sys

FILES = [
    "apps/data/rmat_data/rmat_4096.txt",          # dim_calc/4096's graph (u v per line)
    "apps/data/rmat_data/parsed_rmat_1024.txt",   # APSP / MCP edge lists (v e, u v w)
    "apps/data/rmat_data/parsed_rmat_2048.txt",
    "apps/data/aplp_input/DAG_rmat_4096.txt",     # APLP (negative weights, a DAG)
    "apps/data/aplp_input/DAG_rmat_8192.txt",
    "apps/data/mst_input/mst_rmat_1024.txt",      # MST (PBBS WeightedEdgeArray)
    "apps/data/minrp_input/minrp_rmat_4096.txt",  # MinRP (weights in (0, 1))
]


def main(ref="/root/reference"):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "graphs", "reference")
    os.makedirs(out, exist_ok=True)
    for rel in FILES:
        dst = os.path.join(out, os.path.basename(rel) + ".gz")
        with open(os.path.join(ref, rel), "rb") as f, gzip.GzipFile(dst, "wb", 9, mtime=0) as g:
            shutil.copyfileobj(f, g)
        print(dst, os.path.getsize(dst))


if __name__ == "__main__":
    main(*sys.argv[1:])
