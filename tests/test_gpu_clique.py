"""The clique schedule (4-clique general cores, one visit per clique) through the C ABI
against the oracle (the reference's per-embedding DFS, FC:489-614) and against the
per-root GPU kernel: class sums, exact counts, the reference's u64, core matches and
flags, bit-exact. Marked gpu."""
import os

import pytest

import oracle

pytestmark = pytest.mark.gpu

SMALL = ["er:12:30:1", "er:25:70:2", "rmat:6:4:0.57:0.19:0.19:5", "er:48:500:2"]


def same(r, o, what=""):
    assert r.sg == o["sg"], what
    assert r.exact == o["exact"], what
    assert r.aut == o["aut"], what
    assert r.ref_compat == o["ref_compat"], what
    assert r.core_matches == o["core_matches"], what
    assert (r.flags & 0x1f) == (o["flags"] & 0x1f), what


def clique_motifs(fsgc, seeds, limit=None):
    out = []
    for seed in seeds:
        for i, mo in enumerate(fsgc.random_motifs(seed, 64)):
            if fsgc.motif_class(mo) == "general" and fsgc.live_fields(mo)["list"] == [1, 3, 7]:
                out.append((f"seed {seed} motif {i}", mo))
    return out[:limit] if limit else out


def test_c3_clique_motifs_vs_oracle(gpu):
    mos = clique_motifs(gpu, [3])
    assert len(mos) == 6
    for spec in SMALL:
        ni, nl = gpu.generate_csr(spec)
        maxdeg = int((ni[1:] - ni[:-1]).max())
        g = gpu.Graph.from_csr(ni, nl)
        print("graph", spec, flush=True)
        for name, mo in mos:
            assert gpu.plan_schedule(mo, maxdeg) == "clique"
            print("  ", name, flush=True)
            same(g.count(mo), oracle.count(mo, ni, nl), f"{name} on {spec}")
        g.free()


def test_random_clique_motifs_vs_oracle(gpu):
    mos = clique_motifs(gpu, range(1, 40), limit=40)
    for spec in SMALL[1:3]:
        ni, nl = gpu.generate_csr(spec)
        g = gpu.Graph.from_csr(ni, nl)
        print("graph", spec, flush=True)
        for name, mo in mos:
            same(g.count(mo), oracle.count(mo, ni, nl), f"{name} on {spec}")
        g.free()


def test_clique_schedule_vs_per_root_kernel(gpu):
    """A larger R-MAT graph with hub cliques: the clique schedule against the
    reference's per-embedding DFS on the GPU (schedule 1), which the tests above pin
    to the oracle."""
    mos = clique_motifs(gpu, [3]) + clique_motifs(gpu, range(1, 40), limit=16)[::2]
    ni, nl = gpu.generate_csr("rmat:8:8:0.57:0.19:0.19:3")
    g = gpu.Graph.from_csr(ni, nl)
    for name, mo in mos:
        a = g.count(mo)
        b = g.count(mo, gpu.Tune(schedule=1))
        print("  ", name, a.core_matches, a.kernel_ms, b.kernel_ms, flush=True)
        assert (a.sg, a.exact, a.aut, a.ref_compat, a.core_matches, a.flags) == \
            (b.sg, b.exact, b.aut, b.ref_compat, b.core_matches, b.flags), name
    g.free()


def test_clique_batch_equals_single(gpu):
    mos = [mo for _, mo in clique_motifs(gpu, [3]) + clique_motifs(gpu, range(1, 40), limit=6)]
    ni, nl = gpu.generate_csr("er:48:500:2")
    g = gpu.Graph.from_csr(ni, nl)
    batch, _ = g.count_batch(mos)
    for mo, b in zip(mos, batch):
        a = g.count(mo)
        assert (a.sg, a.exact, a.aut, a.ref_compat, a.core_matches, a.flags) == \
            (b.sg, b.exact, b.aut, b.ref_compat, b.core_matches, b.flags)
    g.free()


def test_clique_global_marks_and_search_path(gpu):
    """Marks in global memory (LDS cap of 16 positions) and the quadruple counts by
    search instead of bit rows (a one-word row budget), against the oracle."""
    mos = clique_motifs(gpu, [3])
    ni, nl = gpu.generate_csr("er:48:500:2")
    g = gpu.Graph.from_csr(ni, nl)
    os.environ["FSGC_CLIQUE_ROWCAP"] = "1"
    try:
        for name, mo in mos:
            o = oracle.count(mo, ni, nl)
            same(g.count(mo, gpu.Tune(lds_cap=16)), o, name + " lds_cap 16, searches")
            same(g.count(mo), o, name + " searches")
    finally:
        del os.environ["FSGC_CLIQUE_ROWCAP"]
    g.free()


def test_clique_root_slices(gpu):
    mos = clique_motifs(gpu, [3])
    ni, nl = gpu.generate_csr("rmat:6:4:0.57:0.19:0.19:5")
    g = gpu.Graph.from_csr(ni, nl)
    for name, mo in mos:
        for rem in (0, 3):
            o = oracle.count(mo, ni, nl, slice_mod=5, slice_rem=rem)
            r = g.count(mo, gpu.Tune(slice_mod=5, slice_rem=rem))
            assert (r.sg, r.core_matches) == (o["sg"], o["core_matches"]), (name, rem)
    g.free()
