"""Edges-per-second table over the reference's motif library, on synthetic graphs.

The counterpart of the reference's benchmark driver (reproduce.py:7-96, with
geometric_fringe.py:8-66 for the geometric mean): for every motif file of one family
(`vertex`, `edge`, `wedge`, `triangle`, `tails`, `tri_fringe`, `fig3wedge`) and every
graph, preprocess the motif, count it on the GPU and record
  edges / runtime
where `edges` is the graph's directed entry count (what the reference's `info`
prints) and `runtime` the wall time of the count call (the scope the reference times,
fringeCount.cu:2317-2321). The motif edge lists come from the committed fixture of
the reference's `patterns/` files (tests/golden/mo_golden.json); the reference's
downloaded graphs are replaced by synthetic ones of the same kinds.

    python -m fsgc.reproduce edge [--graphs spec,spec,...] [--out edge.csv]
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(os.path.dirname(_HERE))
sys.path.insert(0, os.path.dirname(_HERE))

import fsgc  # noqa: E402

# stand-ins for the reference's graph list (reproduce.py:8): a road-like sparse
# graph, a power-law web/social graph, the R-MAT 16 graph itself, and a denser ER graph
DEFAULT_GRAPHS = [
    "er:262144:524288:3",
    "chunglu:500000:2.1:10:5000:7",
    "rmat:16:16:0.57:0.19:0.19:16",
    "er:65536:1048576:5",
]

FAMILIES = {
    "vertex": "patterns/vertex_",
    "edge": "patterns/edge_",
    "wedge": "patterns/wedge_",
    "triangle": "patterns/triangle_",
    "tails": "patterns/tails/",
    "tri_fringe": "patterns/tri_fringe/",
    "fig3wedge": "patterns/wedge/",
}


def library(family: str):
    """(name, nodes, edges) of every motif file of a family, sorted by name."""
    g = json.load(open(os.path.join(_ROOT, "tests", "golden", "mo_golden.json")))
    pre = FAMILIES[family]
    out = []
    for r in g["motifs"]:
        src = r["source"]
        if src.startswith(pre) and "/" not in src[len(pre):]:
            out.append((os.path.splitext(os.path.basename(src))[0], r["nodes"], [tuple(e) for e in r["edges"]]))
    return sorted(out)


def geomean(xs):
    xs = [x for x in xs if x and x > 0]
    return math.exp(sum(math.log(x) for x in xs) / len(xs)) if xs else 0.0


def run(family: str, graphs, out_csv: str | None = None):
    fsgc.init(0)
    motifs = library(family)
    header = ["Graph File"] + [f"fsgc MI355X {name}" for name, _, _ in motifs] + ["geomean"]
    rows = []
    for spec in graphs:
        g = fsgc.Graph.generate(spec)
        n, m, _ = g.info()
        row = [spec]
        for name, nodes, edges in motifs:
            mo, _ = fsgc.preprocess(nodes, edges)
            g.count(mo)  # per-graph index and automorphism sum, as a warm call
            t0 = time.perf_counter()
            r = g.count(mo)
            dt = time.perf_counter() - t0
            row.append(m / dt)
            print(f"# {spec} {name}: {r.exact} occurrences, {dt * 1e3:.2f} ms", flush=True)
        row.append(geomean(row[1:]))
        rows.append(row)
        g.free()
    if out_csv:
        with open(out_csv, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(header)
            w.writerows(rows)
    return header, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("family", choices=sorted(FAMILIES))
    ap.add_argument("--graphs", default=",".join(DEFAULT_GRAPHS))
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    header, rows = run(a.family, a.graphs.split(","), a.out or f"{a.family}.csv")
    for row in rows:
        print(row[0], f"geomean {row[-1]:.3e} edges/s", flush=True)


if __name__ == "__main__":
    main()
