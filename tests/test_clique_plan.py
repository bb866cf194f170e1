"""Host side of the clique schedule (no GPU): which 4-clique motifs it takes on which
graphs, and the fringe formula its tables encode, restated here with Python integers
and checked against the oracle's functionX enumeration (FC:408-486) on small graphs."""
import itertools
import math
from collections import Counter

import oracle


def c3_clique_motifs(fsgc):
    mos = fsgc.random_motifs(3, 64)
    return [mo for mo in mos if fsgc.motif_class(mo) == "general" and fsgc.live_fields(mo)["list"] == [1, 3, 7]]


def test_clique_schedule_guard(fsgc_mod):
    fsgc = fsgc_mod
    mos = c3_clique_motifs(fsgc)
    assert len(mos) == 6
    for mo in mos:
        assert fsgc.plan_schedule(mo, 163586) == "clique"  # R-MAT s22's maximum degree
        assert fsgc.plan_schedule(mo, 1 << 21) == "per-root"  # three factors no longer fit 64 bits
    # motif 6 of seed 3 has 7 fringes: maxdeg^7 must stay below 2^128
    six = mos[0]
    assert sum(fsgc.live_fields(six)["fcnt"]) == 7
    r = int(round(2 ** (128 / 7)))
    while r ** 7 >= 1 << 128:
        r -= 1
    while (r + 1) ** 7 < 1 << 128:
        r += 1
    assert fsgc.plan_schedule(six, r) == "clique"  # r^7 < 2^128
    assert fsgc.plan_schedule(six, r + 1) == "per-root"
    # a path core is not a clique core
    path = next(mo for mo in fsgc.random_motifs(3, 64)
                if fsgc.motif_class(mo) == "general" and fsgc.live_fields(mo)["list"] != [1, 3, 7])
    assert fsgc.plan_schedule(path, 1000) == "per-root"


def set_partitions(items):
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for p in set_partitions(rest):
        for i in range(len(p)):
            yield p[:i] + [[first] + p[i]] + p[i + 1:]
        yield [[first]] + p


def clique_sum(mo, fsgc, ni, nl):
    """Sum over 4-cliques and the 24 assignments passing the degree filters of
    (1 / prod fcnt!) * sum over set partitions P of the labelled fringes of
    prod_B (-1)^{|B|-1} (|B|-1)! N[union of the anchor sets in B]."""
    lf = fsgc.live_fields(mo)
    labels = [S for S, f in zip(lf["anch"], lf["fcnt"]) for _ in range(f)]
    terms = Counter()
    for p in set_partitions(list(range(len(labels)))):
        mu, masks = 1, []
        for B in p:
            mu *= (-1) ** (len(B) - 1) * math.factorial(len(B) - 1)
            u = 0
            for j in B:
                u |= labels[j]
            masks.append(u)
        terms[tuple(sorted(masks))] += mu
    den = math.prod(math.factorial(f) for f in lf["fcnt"])
    thr = [lf["startdeg"]] + lf["destdeg"]
    n = len(ni) - 1
    N = [set(nl[ni[v]:ni[v + 1]].tolist()) for v in range(n)]
    total = matches = 0
    for K in itertools.combinations(range(n), 4):
        if not all(K[b] in N[K[a]] for a, b in itertools.combinations(range(4), 2)):
            continue
        for perm in itertools.permutations(K):
            if any(len(N[perm[i]]) < thr[i] for i in range(4)):
                continue
            s = 0
            for masks, c in terms.items():
                t = c
                for m in masks:
                    U = [perm[i] for i in range(4) if m >> i & 1]
                    t *= len(set.intersection(*[N[u] for u in U]) - set(K))
                s += t
            assert s % den == 0
            total += s // den
            matches += 1
    return total, matches


def test_clique_formula_vs_oracle(fsgc_mod):
    fsgc = fsgc_mod
    for spec in ("er:12:30:1", "er:14:50:4"):
        ni, nl = fsgc.generate_csr(spec)
        for mo in c3_clique_motifs(fsgc):
            o = oracle.count(mo, ni, nl)
            assert clique_sum(mo, fsgc, ni, nl) == (o["sg"], o["core_matches"]), spec
