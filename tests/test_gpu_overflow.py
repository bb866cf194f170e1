"""The HIP path at counts of 2^64 and more, and on the kernels' large-degree branches:
exact u128 counts, the class sums, the reference's u64 and the flags, against closed
forms and the oracle. Marked gpu.

Closed forms (Python big integers, pinned against brute force in test_overflow.py):
K_n and K_{a,b} counts of the C2 and C4 motifs. Tune.lds_cap shrinks the kernels' LDS
staging (the pivot kernel's c3 windows and shared chunks, the rank kernel's LDS marks)
so that small graphs take the branches the headline graphs take at degrees above 2048
and 8192; FSGC_PIVOT_LISTS32=1 makes the pivot kernel read the 32-bit common-neighbour
lists that graphs of maximum degree above 65536 use."""
import math
import os

import pytest

import oracle
from test_gpu_parity import C2_MOTIF, C4_MOTIF, same
from test_overflow import U64, c2_in_kab, c2_in_kn, c4_in_kn, complete, complete_bipartite, ref_compat, star

pytestmark = pytest.mark.gpu


def check_closed_form(r, want):
    assert r.exact == want
    assert r.sg == want * r.aut
    assert r.ref_compat == ref_compat(want, r.aut)
    assert bool(r.flags & 1) == (want >= U64)  # OVF64
    assert not (r.flags & 0b1010)  # OVF128, AUT_REMAINDER


def test_c4_motif_on_k260_rank_kernel(gpu):
    """Rank kernel (triangle core), count above 2^64, LDS and global marks."""
    n = 260
    ni, nl = complete(n)
    g = gpu.Graph.from_csr(ni, nl)
    mo, _ = gpu.preprocess(*C4_MOTIF)
    want = c4_in_kn(n)
    assert want >= U64
    o = oracle.count(mo, ni, nl)
    for cap in (0, 64):
        print("K260 C4 lds_cap", cap, flush=True)
        r = g.count(mo, gpu.Tune(lds_cap=cap))
        check_closed_form(r, want)
        same(r, o, f"K260 C4 lds_cap {cap}")
    g.free()


def test_c2_motif_on_k370_pivot_kernel(gpu):
    """Pivot kernel with c3 on every wedge, count above 2^64. lds_cap 64 cuts every
    centre's walk into c3 windows (phase B2) and every 2-hop list into shared chunks
    (phase A2); the 32-bit lists too."""
    n = 370
    ni, nl = complete(n)
    g = gpu.Graph.from_csr(ni, nl)
    mo, _ = gpu.preprocess(*C2_MOTIF)
    want = c2_in_kn(n)
    assert want >= U64
    base = None
    for cap, l32 in ((0, False), (64, False), (100, True)):
        print("K370 C2 lds_cap", cap, "lists32", l32, flush=True)
        if l32:
            os.environ["FSGC_PIVOT_LISTS32"] = "1"
        try:
            r = g.count(mo, gpu.Tune(lds_cap=cap))
        finally:
            os.environ.pop("FSGC_PIVOT_LISTS32", None)
        check_closed_form(r, want)
        if base is None:
            base = r
        assert (r.sg, r.core_matches, r.flags) == (base.sg, base.core_matches, base.flags)
    g.free()


def test_c2_motif_on_k5_5400_pivot_kernel(gpu):
    """Pivot kernel at degree 5400 with the default caps: phase A2 chunks (lists past
    the pivot longer than 2048) and phase B2 walk chunks; both terms of the closed
    form; the per-root schedule and the oracle agree."""
    a, b = 5, 5400
    ni, nl = complete_bipartite(a, b)
    g = gpu.Graph.from_csr(ni, nl)
    mo, _ = gpu.preprocess(*C2_MOTIF)
    want = c2_in_kab(a, b)
    assert want >= U64
    r = g.count(mo)
    check_closed_form(r, want)
    o = oracle.count(mo, ni, nl)
    same(r, o, "K5,5400 C2")
    p = g.count(mo, gpu.Tune(schedule=1))
    assert (p.sg, p.core_matches, p.exact) == (r.sg, r.core_matches, r.exact)
    g.free()


@pytest.mark.parametrize("d,k", [(16384, 6), (1000, 6), (200000, 4), (70000, 4)])
def test_star_ref_nck_inexact(gpu, d, k):
    """Edge-pass kernel (vertex stars): C(d, k) exactly, with REF_NCK_INEXACT where the
    reference's u64 nCk wraps, as the oracle sets it."""
    ni, nl = star(d)
    g = gpu.Graph.from_csr(ni, nl)
    mo, _ = gpu.preprocess(k + 1, [(0, i) for i in range(1, k + 1)])
    r = g.count(mo)
    o = oracle.count(mo, ni, nl)
    assert r.exact == math.comb(d, k)
    same(r, o, f"star {d} k {k}")
    g.free()


def test_c2_motif_slice_above_2_64_on_rmat14(gpu):
    """The default (pivot) schedule on R-MAT s14 ef16 seed 7 (maximum degree 3674:
    hub centres take several c3 windows) on the root slice u % 61 == 1, whose class
    sum exceeds 2^64, against the oracle on the same slice."""
    ni, nl = gpu.generate_csr("rmat:14:16:0.57:0.19:0.19:7")
    g = gpu.Graph.from_csr(ni, nl)
    mo, _ = gpu.preprocess(*C2_MOTIF)
    o = oracle.count(mo, ni, nl, slice_mod=61, slice_rem=1)
    r = g.count(mo, gpu.Tune(slice_mod=61, slice_rem=1))
    assert o["sg"] >= U64
    same(r, o, "rmat14 C2 slice 61/1")
    g.free()


LDS_CAP_GRAPHS = ["rmat:9:8:0.57:0.19:0.19:21", "chunglu:3000:2.1:12:300:25"]


@pytest.mark.parametrize("spec", LDS_CAP_GRAPHS)
def test_lds_cap_branches_vs_oracle(gpu, spec):
    """Wedge and triangle motifs (closed-form and full fc3, cliques) with the LDS
    staging capped at 16 and 48 positions, and with the 32-bit lists, against the
    oracle: every multi-window, shared-chunk and global-mark branch runs."""
    ni, nl = gpu.generate_csr(spec)
    g = gpu.Graph.from_csr(ni, nl)
    mos = [gpu.preprocess(*C2_MOTIF)[0], gpu.preprocess(*C4_MOTIF)[0]]
    mos += [gpu.preprocess_pattern(p)[0] for p in (9, 13, 15, 19, 20, 22, 25, 29)]
    mos += [m for m in gpu.random_motifs(3, 64) if gpu.motif_class(m) in ("wedge", "triangle")][:8]
    for mo in mos:
        o = oracle.count(mo, ni, nl)
        for cap, l32 in ((16, False), (48, True)):
            print(spec, gpu.live_fields(mo), "lds_cap", cap, "lists32", l32, flush=True)
            if l32:
                os.environ["FSGC_PIVOT_LISTS32"] = "1"
            try:
                r = g.count(mo, gpu.Tune(lds_cap=cap))
            finally:
                os.environ.pop("FSGC_PIVOT_LISTS32", None)
            same(r, o, f"{spec} {gpu.live_fields(mo)} lds_cap {cap}")
    batch, _ = g.count_batch(mos, gpu.Tune(lds_cap=16))
    for mo, b in zip(mos, batch):
        same(b, oracle.count(mo, ni, nl), f"batch lds_cap 16 {gpu.live_fields(mo)}")
    g.free()
