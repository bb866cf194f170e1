"""HIP counting path (through the C ABI) against the oracle: bit-exact class sums,
automorphism sums, exact counts, reference-compatible u64 values and core-match
counts. Marked gpu."""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mo_golden.json")))
C1_MOTIF = (5, [(0, 1), (1, 2), (0, 2), (0, 3), (0, 4)])
C2_MOTIF = (8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)])
C4_MOTIF = (9, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (0, 4), (0, 5), (0, 6), (1, 7), (1, 8)])

SMALL = ["er:12:30:1", "er:25:70:2", "rmat:6:4:0.57:0.19:0.19:5"]


def same(r, o, what=""):
    assert r.sg == o["sg"], what
    assert r.exact == o["exact"], what
    assert r.aut == o["aut"], what
    assert r.ref_compat == o["ref_compat"], what
    assert r.core_matches == o["core_matches"], what
    assert r.motif_class == o["motif_class"], what
    mask = 0x1f
    assert (r.flags & mask) == (o["flags"] & mask), what


@pytest.fixture(scope="module")
def small_graphs(gpu):
    out = []
    for spec in SMALL:
        ni, nl = gpu.generate_csr(spec)
        out.append((spec, ni, nl, gpu.Graph.from_csr(ni, nl)))
    return out


def all_motifs(fsgc):
    mos = []
    for r in GOLDEN["motifs"]:
        if "number" in r:
            mos.append((r["source"], fsgc.preprocess_pattern(r["number"])[0]))
        else:
            mos.append((r["source"], fsgc.preprocess(r["nodes"], [tuple(e) for e in r["edges"]])[0]))
    return mos


def test_every_golden_motif_on_small_graphs(gpu, small_graphs):
    for spec, ni, nl, g in small_graphs:
        print("graph", spec, flush=True)
        for name, mo in all_motifs(gpu):
            print("  motif", name, flush=True)
            o = oracle.count(mo, ni, nl)
            r = g.count(mo)
            same(r, o, f"{name} on {spec}")


def test_c1_config_full(gpu):
    """ER G(4096, 32768) seed 1, triangle core + 2 fringes on one core vertex."""
    ni, nl = gpu.generate_csr("er:4096:32768:1")
    g = gpu.Graph.from_csr(ni, nl)
    mo, _ = gpu.preprocess(*C1_MOTIF)
    o = oracle.count(mo, ni, nl)
    r = g.count(mo)
    same(r, o, "C1")
    assert r.exact > 0 and r.motif_class == 1


def test_batch_equals_single(gpu):
    ni, nl = gpu.generate_csr("er:2000:8000:4")
    g = gpu.Graph.from_csr(ni, nl)
    mos = gpu.random_motifs(3, 64)
    single = [g.count(mo) for mo in mos]
    batch, ms = g.count_batch(mos)
    assert ms > 0
    for i, (a, b) in enumerate(zip(single, batch)):
        assert (a.exact, a.sg, a.aut, a.ref_compat, a.core_matches, a.flags) == \
               (b.exact, b.sg, b.aut, b.ref_compat, b.core_matches, b.flags), i


def test_tuning_invariance(gpu):
    ni, nl = gpu.generate_csr("er:300:1200:13")
    g = gpu.Graph.from_csr(ni, nl)
    mos = gpu.random_motifs(3, 16) + [gpu.preprocess(*C2_MOTIF)[0], gpu.preprocess(*C4_MOTIF)[0]]
    base = [g.count(mo) for mo in mos]
    # block selects the pivot kernel's 4-, 8- or 16-wave workgroup shape
    for chunk, grid, hub, block in ((1, 7, 1, 256), (64, 0, 16, 512), (4096, 33, 100000, 1024), (256, 0, 4096, 0)):
        t = gpu.Tune(chunk=chunk, grid=grid, hub_degree=hub, block=block)
        for mo, b in zip(mos, base):
            r = g.count(mo, t)
            assert (r.exact, r.sg, r.core_matches) == (b.exact, b.sg, b.core_matches), (chunk, grid, hub, block)


def test_random_motifs_vs_oracle(gpu):
    for spec in ("er:300:1200:3", "er:2000:8000:4", "rmat:7:3:0.57:0.19:0.19:5"):
        ni, nl = gpu.generate_csr(spec)
        g = gpu.Graph.from_csr(ni, nl)
        for i, mo in enumerate(gpu.random_motifs(3, 64)):
            print("  ", spec, "random motif", i, flush=True)
            o = oracle.count(mo, ni, nl)
            same(g.count(mo), o, f"random motif {i} on {spec}")


def test_c2_c4_motifs_sliced_at_scale(gpu):
    """The C2 and C4 motifs on larger skewed graphs, compared on a deterministic slice of
    root vertices (the oracle's CPU time stays in seconds)."""
    for spec, motif, mod in (("rmat:12:16:0.57:0.19:0.19:7", C2_MOTIF, 31), ("chunglu:20000:2.1:20:2000:19", C4_MOTIF, 5)):
        ni, nl = gpu.generate_csr(spec)
        g = gpu.Graph.from_csr(ni, nl)
        mo, _ = gpu.preprocess(*motif)
        for rem in (1, 2):
            o = oracle.count(mo, ni, nl, slice_mod=mod, slice_rem=rem)
            r = g.count(mo, gpu.Tune(slice_mod=mod, slice_rem=rem))
            assert (r.sg, r.core_matches) == (o["sg"], o["core_matches"]), (spec, rem)


def test_wedge_pivot_schedule_matches_per_root(gpu):
    """The pivot schedule (default for wedge cores) and the per-root schedule agree on
    every wedge-class motif, and both match the oracle."""
    for spec in ("rmat:9:8:0.57:0.19:0.19:21", "er:500:3000:22"):
        ni, nl = gpu.generate_csr(spec)
        g = gpu.Graph.from_csr(ni, nl)
        mos = [m for m in gpu.random_motifs(3, 64) + [gpu.preprocess(*C2_MOTIF)[0]] if m[0] == 3]
        mos += [gpu.preprocess_pattern(p)[0] for p in (7, 13, 15, 17, 19, 22, 24, 26, 28)]
        for mo in mos:
            a = g.count(mo)
            b = g.count(mo, gpu.Tune(schedule=1))
            o = oracle.count(mo, ni, nl)
            if a.motif_class != 2:
                continue
            same(a, o, f"pivot {gpu.live_fields(mo)} on {spec}")
            assert (b.sg, b.core_matches) == (a.sg, a.core_matches)


def test_triangle_rank_schedule_matches_per_root(gpu):
    """The rank schedule (default for triangle cores and the triangle kinds of edge cores)
    agrees with the per-root schedule and with the oracle, cliques included."""
    for spec in ("rmat:9:8:0.57:0.19:0.19:23", "er:500:3000:24", "chunglu:3000:2.1:12:300:25"):
        ni, nl = gpu.generate_csr(spec)
        g = gpu.Graph.from_csr(ni, nl)
        mos = [m for m in gpu.random_motifs(3, 64) if m[0] in (2, 3)]
        mos += [gpu.preprocess_pattern(p)[0] for p in (3, 6, 9, 12, 20, 21, 25, 27, 29)]
        mos += [gpu.preprocess(*C4_MOTIF)[0], gpu.preprocess(*C1_MOTIF)[0]]
        for mo in mos:
            a = g.count(mo)
            b = g.count(mo, gpu.Tune(schedule=1))
            assert (a.sg, a.core_matches, a.exact) == (b.sg, b.core_matches, b.exact), (gpu.live_fields(mo), spec)
            o = oracle.count(mo, ni, nl)
            same(a, o, f"rank {gpu.live_fields(mo)} on {spec}")


def test_wedge_pivot_hub_and_global_counters(gpu):
    """Graphs with more vertices than the pivot kernel has LDS hub counters (39552).
    Launches without c3 (the closed-form wedge kinds and the motifs linear in the
    triple region) then keep the 2-hop counters of the highest-degree vertices in LDS
    and the others in global memory; launches with c3, and the 4-wave shape, keep all
    of them in global memory. Each shape (12-wave default, 16, 8 and 4 waves), and a
    batch that mixes c3 and closed-form motifs, matches the oracle and the per-root
    schedule."""
    for spec in ("er:120000:400000:41", "chunglu:120000:2.1:8:500:43"):
        ni, nl = gpu.generate_csr(spec)
        g = gpu.Graph.from_csr(ni, nl)
        mos = [gpu.preprocess(*C2_MOTIF)[0], gpu.preprocess_pattern(13)[0], gpu.preprocess_pattern(15)[0]]
        refs = []
        for mo in mos:
            print("  ", spec, gpu.live_fields(mo), flush=True)
            o = oracle.count(mo, ni, nl)
            a = g.count(mo)
            same(a, o, f"pivot {gpu.live_fields(mo)} on {spec}")
            b = g.count(mo, gpu.Tune(schedule=1))
            assert (b.sg, b.core_matches) == (a.sg, a.core_matches)
            for block in (256, 512, 1024):
                c = g.count(mo, gpu.Tune(block=block))
                assert (c.exact, c.sg, c.core_matches) == (a.exact, a.sg, a.core_matches), block
            refs.append(a)
        batch, _ = g.count_batch(mos)
        for a, b in zip(refs, batch):
            assert (a.exact, a.sg, a.core_matches, a.flags) == (b.exact, b.sg, b.core_matches, b.flags)


def test_wedge_pivot_large_pivot_degree(gpu):
    """A pivot of degree 33000 (a star on vertex 0 plus random edges, on the generator's
    ids): its items are split over several workgroups and its 2-hop counts reach the
    thousands. The default hub shape, the 4-wave shape (global counters only) and the
    per-root schedule agree on every field."""
    import numpy as np
    rng = np.random.default_rng(7)
    n, hub = 40000, 33000
    pairs = {(0, i) for i in range(1, hub + 1)}
    while len(pairs) < hub + 120000:
        a, b = (int(x) for x in rng.integers(1, n, 2))
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    e = np.array(sorted(pairs), dtype=np.int64)
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ni = np.zeros(n + 1, dtype=np.int32)
    np.add.at(ni, src + 1, 1)
    ni = np.cumsum(ni).astype(np.int32)
    g = gpu.Graph.from_csr(ni, dst.astype(np.int32))
    for mo in (gpu.preprocess(*C2_MOTIF)[0], gpu.preprocess_pattern(13)[0], gpu.preprocess_pattern(15)[0]):
        a = g.count(mo)
        for t in (gpu.Tune(block=256), gpu.Tune(schedule=1)):
            b = g.count(mo, t)
            assert (a.exact, a.sg, a.core_matches, a.flags) == (b.exact, b.sg, b.core_matches, b.flags), gpu.live_fields(mo)
    g.free()
