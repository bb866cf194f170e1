"""Wedge cores whose count is linear in the triple region (f fringes on the centre, one
on each endpoint; C2's motif is one): the pivot schedule evaluates them without the
per-wedge c3 and adds the per-centre terms (k_wedge_lin_centres). Compared through the
C ABI with the same schedule evaluating every wedge with its c3
(FSGC_PIVOT_C3_PER_WEDGE=1), with the oracle and with closed forms. Marked gpu."""
import os
from math import comb

import numpy as np
import pytest

import oracle
from test_gpu_parity import C2_MOTIF, same

pytestmark = pytest.mark.gpu


def wedge_motif(gpu, f):
    edges = [(0, 1), (1, 2), (0, 3), (2, 4)] + [(1, 5 + i) for i in range(f)]
    return gpu.preprocess(5 + f, edges)[0]


def per_wedge_c3(g, mo, tune=None):
    os.environ["FSGC_PIVOT_C3_PER_WEDGE"] = "1"
    try:
        return g.count(mo, tune)
    finally:
        os.environ.pop("FSGC_PIVOT_C3_PER_WEDGE", None)


def key(r):
    return (r.exact, r.sg, r.aut, r.ref_compat, r.core_matches, r.flags)


@pytest.mark.parametrize("spec", ["rmat:10:8:0.57:0.19:0.19:3", "er:2000:12000:5", "chunglu:3000:2.1:12:400:7"])
def test_linear_equals_per_wedge_and_oracle(gpu, spec):
    ni, nl = gpu.generate_csr(spec)
    g = gpu.Graph.from_csr(ni, nl)
    for f in (1, 2, 3, 4, 6):
        mo = wedge_motif(gpu, f)
        print(spec, "f", f, flush=True)
        a = g.count(mo)
        b = per_wedge_c3(g, mo)
        assert key(a) == key(b), (spec, f)
        same(a, oracle.count(mo, ni, nl), f"f={f} on {spec}")
    g.free()


def test_hub_graph_large_fringe_counts(gpu):
    """Hubs of degree about 3000 (centres with several walk chunks, shared 2-hop chunks)
    and seven centre fringes: the binomials leave the inline u64 range, so the u128
    path and the reference-wrap flag run; results and flags equal the per-wedge
    evaluation. Root slices (which take the per-wedge evaluation: the reference gives a
    triangle's three rotations to its smallest vertex, the per-centre terms give each
    to its own centre) equal the oracle's."""
    ni, nl = gpu.generate_csr("chunglu:20000:2.1:12:3000:9")
    g = gpu.Graph.from_csr(ni, nl)
    for f in (3, 7):
        mo = wedge_motif(gpu, f)
        a = g.count(mo)
        b = per_wedge_c3(g, mo)
        assert key(a) == key(b), f
        for rem in (0, 3):
            t = gpu.Tune(slice_mod=7, slice_rem=rem)
            s = g.count(mo, t)
            o = oracle.count(mo, ni, nl, slice_mod=7, slice_rem=rem)
            assert (s.sg, s.core_matches) == (o["sg"], o["core_matches"]), (f, rem)
    g.free()


def test_complete_graph_closed_form(gpu):
    """K_n: every pair of a centre's neighbours is a triangle (all three rotations);
    count = n C(n-1, 2) (n-3)(n-4) C(n-5, f)."""
    n = 60
    ni = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.int32)
    nl = np.array([v for u in range(n) for v in range(n) if v != u], dtype=np.int32)
    g = gpu.Graph.from_csr(ni, nl)
    for f in (1, 2, 3, 5):
        r = g.count(wedge_motif(gpu, f))
        assert r.exact == n * comb(n - 1, 2) * (n - 3) * (n - 4) * comb(n - 5, f), f
        assert key(r) == key(per_wedge_c3(g, wedge_motif(gpu, f))), f
    g.free()


def test_batch_and_capture_with_linear_jobs(gpu):
    """Linear jobs share c3-free launches with the specialised wedge kinds; a batch
    that also holds per-wedge-c3 motifs, and its captured replays, equal single counts."""
    ni, nl = gpu.generate_csr("rmat:11:8:0.57:0.19:0.19:13")
    g = gpu.Graph.from_csr(ni, nl)
    mos = [gpu.preprocess(*C2_MOTIF)[0], wedge_motif(gpu, 1), wedge_motif(gpu, 2)]
    mos += [gpu.preprocess_pattern(p)[0] for p in (13, 15, 19)]
    mos += [m for m in gpu.random_motifs(3, 64) if m[0] == 3][:4]
    single = [g.count(m) for m in mos]
    batch, _ = g.count_batch(mos)
    for a, b in zip(single, batch):
        assert key(a) == key(b)
    cb = g.capture_batch(mos)
    for _ in range(3):
        rs, _ = cb.run()
        for a, b in zip(single, rs):
            assert key(a) == key(b)
        g.count(mos[0])  # eager work between replays
    cb.free()
    g.free()
