"""General cores whose hub root edges are split over several work items (each item
takes one part of the root edge's first-level candidates, WorkItem::split): results
equal unsplit items and the oracle. Marked gpu."""
import pytest

import oracle
from test_gpu_parity import same

pytestmark = pytest.mark.gpu

# C3's random motifs (seed 3): 16, 24, 29 and 31 are path-shaped general cores, 6, 18
# and 32 other general cores
GENERAL_IDS = (16, 24, 29, 31, 6, 18, 32)


def key(r):
    return (r.exact, r.sg, r.aut, r.ref_compat, r.core_matches, r.flags)


def test_split_items_equal_oracle(gpu):
    """chunk 8: items of about 32 units (d_u^2 per root edge), so every root edge of
    degree above 5 is split, in up to 50 parts; chunk 4096 (16384 units): none is."""
    ni, nl = gpu.generate_csr("chunglu:250:2.1:6:60:5")
    g = gpu.Graph.from_csr(ni, nl)
    mos = gpu.random_motifs(3, 64)
    for i in GENERAL_IDS:
        print("motif", i, flush=True)
        fine = g.count(mos[i], gpu.Tune(chunk=8))
        coarse = g.count(mos[i], gpu.Tune(chunk=4096))
        assert key(fine) == key(coarse), i
        same(fine, oracle.count(mos[i], ni, nl), f"motif {i}")
    g.free()


def test_split_items_with_slices_and_batch(gpu):
    """Root slices and a batch of split general jobs beside other classes equal the
    counts with few splits (R-MAT s9, hubs of degree up to 279: about 1024 units per
    item by default, 64 parts per hub root edge; 16384 units and 5 parts at chunk 4096)."""
    g = gpu.Graph.generate("rmat:9:16:0.57:0.19:0.19:11")
    mos = gpu.random_motifs(3, 64)
    ids = (31, 6, 18, 32, 0, 1, 8)
    coarse = [g.count(mos[i], gpu.Tune(chunk=4096)) for i in ids]
    fine = [g.count(mos[i]) for i in ids]
    for i, a, b in zip(ids, fine, coarse):
        assert key(a) == key(b), i
    batch, _ = g.count_batch([mos[i] for i in ids])
    for i, a, b in zip(ids, batch, coarse):
        assert key(a) == key(b), i
    for rem in (0, 2):
        s = g.count(mos[31], gpu.Tune(slice_mod=5, slice_rem=rem, chunk=8))
        c = g.count(mos[31], gpu.Tune(slice_mod=5, slice_rem=rem, chunk=4096))
        assert key(s) == key(c), rem
    g.free()


def path_motif(gpu, f0, f2, f3):
    """Path core 0-1-2-3 with f0, f2 and f3 fringes on core vertices 0, 2 and 3."""
    edges = [(0, 1), (1, 2), (2, 3)]
    v = 4
    for a, c in ((0, f0), (2, f2), (3, f3)):
        for _ in range(c):
            edges.append((a, v))
            v += 1
    return gpu.preprocess(v, edges)[0]


@pytest.mark.parametrize("spec", ["rmat:6:6:0.57:0.19:0.19:5", "rmat:7:8:0.57:0.19:0.19:9", "er:30:120:3"])
def test_region_dp_word_size(gpu, spec):
    """The region DP runs in u64 for a match whose N regional vertices and F fringes give
    N^F < 2^64, in u128 otherwise. F = 14 switches at N = 24, F = 11 at N = 57, F = 21 at
    N = 9, F = 8 at N = 256: on these graphs (maximum degree 11 to about 60) one launch
    mixes both word sizes. Results and flags equal the oracle's."""
    ni, nl = gpu.generate_csr(spec)
    g = gpu.Graph.from_csr(ni, nl)
    for fr in ((12, 1, 1), (8, 2, 1), (20, 0, 1), (6, 1, 1)):
        mo = path_motif(gpu, *fr)
        print(spec, fr, flush=True)
        same(g.count(mo), oracle.count(mo, ni, nl), f"{fr} on {spec}")
    g.free()
