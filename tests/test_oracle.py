"""The oracle (exact CPU restatement of the reference method) against brute force,
networkx golden counts, closed forms and invariants. CPU only."""
import itertools
import json
import math
import os
import random

import numpy as np
import pytest

import oracle
import fsgc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def small_graphs():
    out = []
    for i, (n, m) in enumerate([(9, 20), (10, 30), (12, 40), (14, 45), (8, 28)]):
        ni, nl = fsgc.generate_csr(f"er:{n}:{m}:{i + 1}")
        out.append((f"er{n}_{m}", ni, nl))
    out.append(("K7",) + complete(7))
    return out


def complete(n):
    ni, nl = [0], []
    for u in range(n):
        nl += [v for v in range(n) if v != u]
        ni.append(len(nl))
    return np.array(ni), np.array(nl)


def predefined():
    return {r["number"]: (r["nodes"], [tuple(e) for e in r["edges"]])
            for r in json.load(open(os.path.join(GOLDEN, "mo_golden.json")))["motifs"] if "number" in r}


@pytest.mark.parametrize("num", range(1, 31))
def test_predefined_vs_bruteforce(num):
    n, e = predefined()[num]
    mo, _ = fsgc.preprocess(n, e)
    for name, ni, nl in small_graphs():
        r = oracle.count(mo, ni, nl)
        assert r["exact"] == oracle.bruteforce(n, e, ni, nl), (num, name)
        assert not (r["flags"] & fsgc.FLAG_AUT_REMAINDER)


def test_random_motifs_vs_bruteforce():
    mos = fsgc.random_motifs(3, 64)
    graphs = [(f"er{n}", *fsgc.generate_csr(f"er:{n}:{m}:{s}")) for n, m, s in ((10, 30, 11), (12, 36, 12), (11, 40, 14))]
    seen = set()
    for mo in mos:
        hn, he = oracle.motif_graph_edges(mo)
        if hn > 10:
            continue
        for name, ni, nl in graphs:
            r = oracle.count(mo, ni, nl)
            seen.add(r["motif_class"])
            assert r["exact"] == oracle.bruteforce(hn, he, ni, nl), (fsgc.live_fields(mo), name)
    assert seen == {0, 1, 2, 3, 4}


def test_general_core_first_level_repeat():
    """Motifs whose last anchor set is the whole anchor universe (the reference repeats
    functionX's first level): the scaled sums still divide to the brute-force count."""
    rng = random.Random(5)
    hit = 0
    ni, nl = fsgc.generate_csr("er:11:38:21")
    for _ in range(200):
        k = rng.randint(4, 5)
        edges = [(i, j) for i in range(k) for j in range(i + 1, k) if rng.random() < 0.7]
        n = k
        el = list(edges)
        for a in range(k):
            el.append((a, n))  # one fringe on every core vertex
        n += 1
        el.append((0, n)); n += 1
        try:
            mo, _ = fsgc.preprocess(n, el)
        except fsgc.FsgcError:
            continue
        f = fsgc.live_fields(mo)
        anchors = 0
        for a in f["anch"]:
            anchors |= a
        if not (f["nodes"] >= 4 and f["anchorsets"] > 1 and f["anch"][-1] == anchors):
            continue
        hit += 1
        hn, he = oracle.motif_graph_edges(mo)
        r = oracle.count(mo, ni, nl)
        assert r["exact"] == oracle.bruteforce(hn, he, ni, nl)
        if hit >= 6:
            break
    assert hit >= 3


def test_networkx_golden_counts():
    g = json.load(open(os.path.join(GOLDEN, "counts_golden.json")))
    pats = predefined()
    for row in g["counts"]:
        ni, nl = g["graphs"][row["graph"]]
        n, e = pats[row["pattern"]]
        mo, _ = fsgc.preprocess(n, e)
        assert oracle.count(mo, ni, nl)["exact"] == row["count"], row
        assert oracle.bruteforce(n, e, ni, nl) == row["count"], row


def test_motif_in_itself_is_one():
    pats = predefined()
    for num in range(1, 31):
        n, e = pats[num]
        mo, _ = fsgc.preprocess(n, e)
        hn, he = oracle.motif_graph_edges(mo)
        ni = [0]
        nl = []
        adj = [[] for _ in range(hn)]
        for a, b in he:
            adj[a].append(b)
            adj[b].append(a)
        for u in range(hn):
            nl += sorted(adj[u])
            ni.append(len(nl))
        assert oracle.count(mo, ni, nl)["exact"] == 1, num


@pytest.mark.parametrize("num", [3, 7, 9, 13, 20, 23, 25, 30])
def test_complete_graph_closed_form(num):
    """count(H in K_n) = n! / ((n - |H|)! |Aut(H)|)."""
    n, e = predefined()[num]
    mo, _ = fsgc.preprocess(n, e)
    perms = sum(1 for p in itertools.permutations(range(n))
                if all((min(p[a], p[b]), max(p[a], p[b])) in set(map(lambda t: (min(t), max(t)), e)) for a, b in e))
    for kn in (n, n + 2, 9):
        ni, nl = complete(kn)
        want = math.factorial(kn) // (math.factorial(kn - n) * perms)
        assert oracle.count(mo, ni, nl)["exact"] == want


def test_relabel_invariance():
    ni, nl = fsgc.generate_csr("rmat:7:4:0.57:0.19:0.19:3")
    n = len(ni) - 1
    perm = np.random.default_rng(1).permutation(n)
    adj = [[] for _ in range(n)]
    for u in range(n):
        for j in range(ni[u], ni[u + 1]):
            adj[perm[u]].append(perm[nl[j]])
    pni, pnl = [0], []
    for u in range(n):
        pnl += sorted(adj[u])
        pni.append(len(pnl))
    for mo in fsgc.random_motifs(3, 24):
        a = oracle.count(mo, ni, nl)
        b = oracle.count(mo, pni, pnl)
        assert a["exact"] == b["exact"] and a["sg"] == b["sg"] and a["aut"] == b["aut"]


def test_disjoint_union_additivity():
    a_ni, a_nl = fsgc.generate_csr("er:30:70:5")
    b_ni, b_nl = fsgc.generate_csr("rmat:6:3:0.57:0.19:0.19:4")
    na = len(a_ni) - 1
    u_ni = np.concatenate([a_ni, b_ni[1:] + a_ni[-1]])
    u_nl = np.concatenate([a_nl, b_nl + na])
    for mo in fsgc.random_motifs(3, 30):
        ra, rb, ru = (oracle.count(mo, x, y) for x, y in ((a_ni, a_nl), (b_ni, b_nl), (u_ni, u_nl)))
        if ru["motif_class"] == 0 and fsgc.live_fields(mo)["anchorsets"] == 0:
            continue
        assert ru["exact"] == ra["exact"] + rb["exact"]


def _fringe_placements(fr, venn, universe_bits):
    """Brute force: ways to choose disjoint F_S (|F_S| = fr[S]) with each element of F_S
    in a region R that contains S. Elements are enumerated explicitly."""
    elems = []
    for R, c in enumerate(venn):
        if R and c > 0:
            elems += [R] * c
    sets = [(S, k) for S, k in enumerate(fr) if S and k > 0]
    total = 0

    def rec(i, used):
        nonlocal total
        if i == len(sets):
            total += 1
            return
        S, k = sets[i]
        cand = [x for x in range(len(elems)) if x not in used and (elems[x] & S) == S]
        for comb in itertools.combinations(cand, k):
            rec(i + 1, used | set(comb))

    rec(0, frozenset())
    return total


def test_fc3_vs_enumeration():
    rng = random.Random(7)
    for _ in range(250):
        fr = [0] + [rng.choice([0, 0, 1, 1, 2, 3]) for _ in range(7)]
        if sum(fr) > 6:
            continue
        venn = [0] + [rng.randint(0, 3) for _ in range(7)]
        assert oracle.fc3(fr, venn) == _fringe_placements(fr, venn, 3), (fr, venn)


def test_functionx_vs_enumeration():
    rng = random.Random(11)
    done = 0
    while done < 120:
        nodes = rng.randint(4, 5)
        full = (1 << nodes) - 1
        masks = sorted(set(rng.randint(1, full) for _ in range(rng.randint(1, 3))))
        fcnt = [rng.randint(1, 2) for _ in masks]
        if sum(fcnt) > 5:
            continue
        # regions exist only over the anchor universe (commonX never sets other bits)
        union = 0
        for m in masks:
            union |= m
        nonanch = full & ~union
        venn = [0] * 64
        for R in range(1, 1 << nodes):
            if R & nonanch == 0:
                venn[R] = rng.randint(0, 2)
        fr = [0] * 64
        for m, c in zip(masks, fcnt):
            fr[m] = c
        assert oracle.functionx(nodes, masks, fcnt, venn) == _fringe_placements(fr, venn, nodes), (masks, fcnt)
        done += 1


def test_ncr_semantics():
    # fringeCount.cu:189-197: k = min(k, n-k); k < 0 -> 0; k == 0 -> 1
    for n in range(-3, 40):
        for k in range(0, 12):
            v, fl = oracle.ncr(n, k)
            want = math.comb(n, k) if n >= 0 else 0
            assert v == want, (n, k)
    v, fl = oracle.ncr(100000, 6)
    assert v == math.comb(100000, 6) and (fl & fsgc.FLAG_REF_NCK_INEXACT)
    v, fl = oracle.ncr(65000, 3)
    assert v == math.comb(65000, 3) and not (fl & fsgc.FLAG_REF_NCK_INEXACT)


def test_slice_partition():
    """Slices by root vertex partition the class sum."""
    ni, nl = fsgc.generate_csr("rmat:7:4:0.57:0.19:0.19:9")
    for mo in fsgc.random_motifs(3, 16):
        full = oracle.count(mo, ni, nl)
        parts = [oracle.count(mo, ni, nl, slice_mod=3, slice_rem=r) for r in range(3)]
        if full["motif_class"] == 0 and fsgc.live_fields(mo)["anchorsets"] <= 1 and (
                fsgc.live_fields(mo)["anchorsets"] == 0 or fsgc.live_fields(mo)["fcnt"][0] == 1):
            continue
        assert sum(p["sg"] for p in parts) == full["sg"]
        assert sum(p["core_matches"] for p in parts) == full["core_matches"]
