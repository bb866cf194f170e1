"""Device degree relabelling (fsgc_graph_relabel; the reference's relabelByDegree,
fringeCount.cu:1221-1288): the relabelled CSR equals a host renumbering of the same
graph in the reference's order (degree descending, ties by the larger id first), or
the ascending order, and every count on it equals the count on the original graph
(counts are label-invariant). Marked gpu."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

C1_MOTIF = (5, [(0, 1), (1, 2), (0, 2), (0, 3), (0, 4)])
C2_MOTIF = (8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)])
C4_MOTIF = (9, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (0, 4), (0, 5), (0, 6), (1, 7), (1, 8)])
# vertex, edge, wedge, triangle and general cores among the predefined patterns
PATTERNS = [2, 3, 14, 7, 19, 26, 9, 21, 29, 23, 30]


def host_relabel(ni, nl, order):
    """The renumbering on the host: new id = rank of (degree, id) in the order."""
    n = len(ni) - 1
    deg = np.diff(ni).astype(np.int64)
    ids = np.arange(n, dtype=np.int64)
    if order == "desc":
        perm = np.lexsort((-ids, -deg))  # degree descending, ties by the larger id first
    else:
        perm = np.lexsort((ids, deg))
    mp = np.empty(n, dtype=np.int32)
    mp[perm] = np.arange(n, dtype=np.int32)
    new_ni = np.zeros(n + 1, dtype=np.int32)
    new_ni[1:] = np.cumsum(deg[perm])
    new_nl = np.empty(len(nl), dtype=np.int32)
    for i, u in enumerate(perm):
        new_nl[new_ni[i]:new_ni[i + 1]] = np.sort(mp[nl[ni[u]:ni[u + 1]]])
    return mp, new_ni, new_nl


@pytest.mark.parametrize("spec", ["er:12:30:1", "rmat:8:8:0.57:0.19:0.19:3", "er:4096:32768:1",
                                  "rmat:12:16:0.57:0.19:0.19:7"])
@pytest.mark.parametrize("order", ["desc", "asc"])
def test_relabel_matches_host_renumbering(gpu, spec, order):
    print("relabel", spec, order, flush=True)
    ni, nl = gpu.generate_csr(spec)
    g = gpu.Graph.from_csr(ni, nl)
    rg, mp, ms = g.relabel(order)
    exp_mp, exp_ni, exp_nl = host_relabel(ni, nl, order)
    got_ni, got_nl = rg.csr()
    assert np.array_equal(mp, exp_mp)
    assert np.array_equal(got_ni, exp_ni)
    assert np.array_equal(got_nl, exp_nl)
    rg.validate()
    # the original graph is untouched
    o_ni, o_nl = g.csr()
    assert np.array_equal(o_ni, ni) and np.array_equal(o_nl, nl)
    assert ms >= 0
    rg.free()
    g.free()


def _same(a, b, what):
    assert (a.exact, a.sg, a.aut, a.ref_compat, a.flags, a.motif_class) == \
        (b.exact, b.sg, b.aut, b.ref_compat, b.flags, b.motif_class), what


@pytest.mark.parametrize("spec", ["er:4096:32768:1", "rmat:12:16:0.57:0.19:0.19:7"])
def test_counts_invariant_under_relabel(gpu, spec):
    ni, nl = gpu.generate_csr(spec)
    g = gpu.Graph.from_csr(ni, nl)
    mos = [("C1", gpu.preprocess(*C1_MOTIF)[0]), ("C2", gpu.preprocess(*C2_MOTIF)[0]),
           ("C4", gpu.preprocess(*C4_MOTIF)[0])]
    heavy = spec.startswith("rmat")
    for p in PATTERNS:
        mo = gpu.preprocess_pattern(p)[0]
        if heavy and gpu.motif_class(mo) == "general":
            continue  # 4-vertex cores enumerate every core embedding; kept to the ER graph
        mos.append((f"pattern {p}", mo))
    base = {name: g.count(mo) for name, mo in mos}
    for order in ("desc", "asc"):
        rg, _, _ = g.relabel(order)
        for name, mo in mos:
            print("count", spec, order, name, flush=True)
            _same(rg.count(mo), base[name], f"{name} on {spec} relabelled {order}")
        batch, _ = rg.count_batch([mo for _, mo in mos])
        for (name, _), r in zip(mos, batch):
            _same(r, base[name], f"{name} batched on {spec} relabelled {order}")
        rg.free()
    # the oracle on the original graph agrees (C1 in full on the ER graph)
    if not heavy:
        o = oracle.count(mos[0][1], ni, nl)
        assert base["C1"].exact == o["exact"] and base["C1"].sg == o["sg"]
    g.free()
