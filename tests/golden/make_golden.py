"""Writes the committed golden fixtures under tests/golden/.

mo_golden.json     live matching-order fields (and the "including ..." report lines)
                   produced by the reference's own fringePreprocess, built from the
                   reference sources into oracle/_ref/ (oracle/Makefile `ref`), for:
                   the 31 predefined patterns, the 52 pattern files of the reference's
                   patterns/ directory (symmetrised and deduplicated the way its
                   reproduce.py:51-62 does), the 64 random motifs from seed 3, and the
                   motifs of the benchmark configurations.
counts_golden.json non-induced occurrence counts from networkx's VF2 monomorphism
                   enumeration divided by |Aut(H)|, for small graphs and patterns; an
                   independent pin of the brute force and of the counting path.

Run from the repository root:  python tests/golden/make_golden.py
Needs /root/reference (for the pattern files and the reference preprocessor) and
networkx; the fixtures themselves need neither.
"""
import glob
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "fringe-sgc_amd"))
import fsgc  # noqa: E402

REF = "/root/reference"
REF_PRE = os.path.join(ROOT, "oracle", "_ref", "fringePreprocess")

BENCH_MOTIFS = {
    "triangle_2tails_c1": (5, [(0, 1), (1, 2), (0, 2), (0, 3), (0, 4)]),
    "wedge_311_c2": (8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)]),
    "clique4_32_c4": (9, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (0, 4), (0, 5), (0, 6), (1, 7), (1, 8)]),
}


def ref_preprocess(args):
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([REF_PRE] + [str(a) for a in args], cwd=d, capture_output=True, text=True, timeout=60)
        mo_path = os.path.join(d, "motif.mo")
        mo = open(mo_path, "rb").read() if os.path.exists(mo_path) else None
    inc = [l for l in r.stdout.splitlines() if l.startswith("including")]
    return r.returncode, mo, inc


def pattern_files():
    out = []
    for f in sorted(glob.glob(os.path.join(REF, "patterns", "**", "*.txt"), recursive=True)):
        lines = open(f).read().splitlines()
        n = int(re.search(r"Nodes:\s*(\d+)", lines[0]).group(1))
        pairs = set()
        for l in lines[1:]:
            if l.strip():
                a, b = map(int, l.split())
                pairs.add(tuple(sorted((a, b))))
        out.append((os.path.relpath(f, REF), n, sorted(pairs)))
    return out


def predefined_edges():
    tab = open(os.path.join(ROOT, "fringe-sgc_amd", "csrc", "patterns_table.inc")).read()
    pats = {}
    for m in re.finditer(r'\{(\d+), "([^"]*)", (\d+), (\d+), \{([^}]*)\}\}', tab):
        fl = [int(x) for x in m.group(5).split(",")]
        pats[int(m.group(1))] = (int(m.group(3)), [(fl[2 * i], fl[2 * i + 1]) for i in range(len(fl) // 2)])
    return pats


def make_mo_golden():
    rows = []
    pats = predefined_edges()
    for num in range(31):
        rc, mo, inc = ref_preprocess([num])
        assert rc == 0 and mo is not None
        n, e = pats[num]
        rows.append({"source": f"pattern {num}", "number": num, "nodes": n, "edges": e,
                     "live": fsgc.live_fields(mo), "including": inc})
    for name, n, e in pattern_files():
        rc, mo, inc = ref_preprocess([n] + [x for p in e for x in p])
        assert rc == 0 and mo is not None, name
        rows.append({"source": name, "nodes": n, "edges": e, "live": fsgc.live_fields(mo), "including": inc})
    for i, (n, e) in enumerate(fsgc.random_patterns(3, 64)):
        rc, mo, inc = ref_preprocess([n] + [x for p in e for x in p])
        assert rc == 0 and mo is not None
        rows.append({"source": f"random motif {i} (seed 3)", "nodes": n, "edges": e, "live": fsgc.live_fields(mo),
                     "including": inc})
    for name, (n, e) in BENCH_MOTIFS.items():
        rc, mo, inc = ref_preprocess([n] + [x for p in e for x in p])
        assert rc == 0 and mo is not None
        rows.append({"source": name, "nodes": n, "edges": e, "live": fsgc.live_fields(mo), "including": inc})
    # rejected inputs: disconnected motif, more than 6 core vertices
    rejects = []
    for args in ([4, 0, 1, 2, 3], [8, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 0]):
        rc, mo, inc = ref_preprocess(args)
        rejects.append({"args": args, "returncode": rc})
    with open(os.path.join(ROOT, "tests", "golden", "mo_golden.json"), "w") as f:
        json.dump({"motifs": rows, "rejects": rejects}, f, indent=0)
    print("mo_golden.json:", len(rows), "motifs")


def make_counts_golden():
    import networkx as nx
    from networkx.algorithms import isomorphism as iso
    graphs = {}
    for spec in ("er:10:24:101", "er:12:30:102", "er:9:22:103"):
        ni, nl = fsgc.generate_csr(spec)
        graphs[spec] = (ni.tolist(), nl.tolist())
    pats = predefined_edges()
    rows = []
    for spec, (ni, nl) in graphs.items():
        G = nx.Graph()
        G.add_nodes_from(range(len(ni) - 1))
        for u in range(len(ni) - 1):
            for j in range(ni[u], ni[u + 1]):
                G.add_edge(u, nl[j])
        for num in range(1, 31):
            n, e = pats[num]
            H = nx.Graph()
            H.add_nodes_from(range(n))
            H.add_edges_from(e)
            aut = sum(1 for _ in iso.GraphMatcher(H, H).isomorphisms_iter())
            inj = sum(1 for _ in iso.GraphMatcher(G, H).subgraph_monomorphisms_iter())
            assert inj % aut == 0
            rows.append({"graph": spec, "pattern": num, "count": inj // aut})
    with open(os.path.join(ROOT, "tests", "golden", "counts_golden.json"), "w") as f:
        json.dump({"graphs": graphs, "counts": rows}, f, indent=0)
    print("counts_golden.json:", len(rows), "counts")


if __name__ == "__main__":
    make_mo_golden()
    make_counts_golden()
