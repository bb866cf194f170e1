"""The host model of the pivot schedule's algorithmic bytes (oracle.pivot_units):
its unit counts against closed forms of the schedule's traversal. CPU only; the
kernel's own reported bytes are compared with the model in test_gpu_roofline_model.py."""
import numpy as np

import fsgc
import oracle


def supports(ni, nl):
    n = len(ni) - 1
    adj = [set(nl[ni[u]:ni[u + 1]]) for u in range(n)]
    src = np.repeat(np.arange(n), np.diff(ni))
    return np.array([len(adj[u] & adj[x]) for u, x in zip(src, nl)], dtype=np.int64)


def test_pivot_units_closed_forms():
    for spec in ("rmat:9:8:0.57:0.19:0.19:21", "er:400:3000:5", "chunglu:2000:2.1:10:200:3"):
        ni, nl = fsgc.generate_csr(spec)
        d = np.diff(ni).astype(np.int64)
        t = supports(ni, nl)
        u = oracle.pivot_units(ni, nl, ecap=2048, hub_slots=0, c3=True)
        # small graphs: no pivot is split, each item repeats nothing
        assert u["paths_d"] == int((d * (d - 1) // 2).sum())        # 2-hop paths w-x-y, y > w
        assert u["paths_hub"] == u["marks_hub"] == u["walk_hub"] == 0
        assert u["marks_d"] == len(nl) // 2                         # neighbours x > w
        assert u["xvisits"] == len(nl)
        assert u["xvisit_bits"] == int(sum(int(x + 1).bit_length() * x for x in d))
        assert u["walk_d"] == int((d * (d - 1) // 2).sum())          # wedges (u; v > w)
        assert u["centres"] == int((d - 1)[d > 0].sum())
        assert u["list_elems"] == int((t * (t - 1) // 2).sum())
        # one window per centre: one search per c3 entry x of T(u,w), for every pivot w
        # that is not u's last neighbour (a centre without positions past w is skipped)
        last = np.zeros(len(nl), dtype=bool)
        last[ni[1:][d > 0] - 1] = True
        assert u["c3_x"] == int(t[~last].sum())
        # windows of 8 positions: every entry once per window of its centre
        u8 = oracle.pivot_units(ni, nl, ecap=8, hub_slots=0, c3=True)
        assert u8["c3_x"] > u["c3_x"] and u8["c3_search_bits"] > u["c3_search_bits"]
        assert u8["list_elems"] == u["list_elems"] and u8["walk_d"] == u["walk_d"]
        # hub launches: every vertex a hub on a small graph
        h = oracle.pivot_units(ni, nl, ecap=2048, hub_slots=1 << 20, c3=False)
        assert h["paths_d"] == h["marks_d"] == h["walk_d"] == h["c3_x"] == h["list_elems"] == 0
        assert h["paths_hub"] == u["paths_d"] and h["walk_hub"] == u["walk_d"]
