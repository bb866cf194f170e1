"""Counts of 2^64 and more (the regime of the large configurations, where the
reference's u64 wraps): the oracle against closed forms on complete and complete
bipartite graphs, the reference's nCk wrap flag on a star, and the host-side guards
(schedule choice by maximum degree, CSR validation before upload). CPU only."""
import math

import numpy as np
import pytest

import fsgc
import oracle

C2_MOTIF = (8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)])
C4_MOTIF = (9, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (0, 4), (0, 5), (0, 6), (1, 7), (1, 8)])
U64 = 1 << 64


def complete(n):
    """K_n as CSR."""
    nl = np.array([v for u in range(n) for v in range(n) if v != u], dtype=np.int32)
    ni = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.int32)
    return ni, nl


def complete_bipartite(a, b):
    """K_{a,b}: vertices 0..a-1 on one side, a..a+b-1 on the other."""
    lists = [list(range(a, a + b)) for _ in range(a)] + [list(range(a)) for _ in range(b)]
    ni = np.zeros(a + b + 1, dtype=np.int32)
    ni[1:] = np.cumsum([len(x) for x in lists])
    return ni, np.array([v for x in lists for v in x], dtype=np.int32)


def star(d):
    """K_{1,d} with the hub at vertex 0."""
    ni = np.concatenate([[0, d], d + np.arange(1, d + 1)]).astype(np.int32)
    nl = np.concatenate([np.arange(1, d + 1), np.zeros(d)]).astype(np.int32)
    return ni, nl


# Closed forms (Python big integers). A count of non-induced occurrences of H is the
# number of injective homomorphisms H -> G divided by |Aut(H)|.
#  * C4 motif (4-clique a b c d, 3 fringes on a, 2 on b; 9 vertices) in K_n: every
#    injection is a homomorphism, |Aut| = 3! * 2! * 2 (fringes of a, fringes of b,
#    c <-> d) = 24.
#  * C2 motif (wedge 0-1-2, fringes 3 4 5 on the centre 1, 6 on 0, 7 on 2) is a tree
#    with sides {1, 6, 7} and {0, 2, 3, 4, 5}; in K_{a,b} the side of 3 goes to either
#    side of the graph: P(a,3) P(b,5) + P(b,3) P(a,5) injections; |Aut| = 3! * 2 = 12.
def c4_in_kn(n):
    return math.perm(n, 9) // 24


def c2_in_kn(n):
    return math.perm(n, 8) // 12


def c2_in_kab(a, b):
    return (math.perm(a, 3) * math.perm(b, 5) + math.perm(b, 3) * math.perm(a, 5)) // 12


def ref_compat(count, aut):
    """The u64 the reference prints when nothing but the final sum wrapped:
    floor(((aut * count) mod 2^64) / aut) (fringeCount.cu:2252-2261)."""
    return ((count * aut) % U64) // aut


def ref_nck(n, k):
    """fringeCount.cu:189-197 in u64 arithmetic, with the wrap of r * (n + 1 - i); and
    whether any step wrapped."""
    if k > n - k:
        k = n - k
    if k < 0:
        return 0, False
    if k == 0:
        return 1, False
    r, wrapped = n, False
    for i in range(2, k + 1):
        p = r * (n + 1 - i)
        wrapped = wrapped or p >= U64
        r = (p % U64) // i
    return r, wrapped


def test_closed_forms_small_vs_bruteforce():
    """The closed forms themselves, against brute-force enumeration on small graphs."""
    ni, nl = complete(10)
    assert c4_in_kn(10) == oracle.bruteforce(*C4_MOTIF, ni, nl)
    assert c2_in_kn(10) == oracle.bruteforce(*C2_MOTIF, ni, nl)
    ni, nl = complete_bipartite(3, 6)
    assert c2_in_kab(3, 6) == oracle.bruteforce(*C2_MOTIF, ni, nl)
    ni, nl = complete_bipartite(5, 6)
    assert c2_in_kab(5, 6) == oracle.bruteforce(*C2_MOTIF, ni, nl)


def test_oracle_c4_motif_on_k260_above_2_64():
    """Triangle core (relabel + general_triangle_core path) with a count above 2^64."""
    n = 260
    ni, nl = complete(n)
    mo, _ = fsgc.preprocess(*C4_MOTIF)
    o = oracle.count(mo, ni, nl)
    want = c4_in_kn(n)
    assert want >= U64
    assert o["exact"] == want
    assert o["sg"] == want * o["aut"]
    assert o["flags"] & fsgc.FLAG_OVF64
    assert not (o["flags"] & (fsgc.FLAG_OVF128 | fsgc.FLAG_AUT_REMAINDER))
    assert o["ref_compat"] == ref_compat(want, o["aut"])


def test_oracle_c2_motif_on_k5_5400_above_2_64():
    """Wedge core (general_wedge_core path) with a count above 2^64: both terms of the
    closed form are non-zero."""
    a, b = 5, 5400
    ni, nl = complete_bipartite(a, b)
    mo, _ = fsgc.preprocess(*C2_MOTIF)
    o = oracle.count(mo, ni, nl)
    want = c2_in_kab(a, b)
    assert want >= U64
    assert o["exact"] == want
    assert o["sg"] == want * o["aut"]
    assert o["flags"] & fsgc.FLAG_OVF64
    assert o["ref_compat"] == ref_compat(want, o["aut"])


@pytest.mark.parametrize("d,k,wraps", [(16384, 6, True), (1000, 6, False), (200000, 4, True), (70000, 4, False)])
def test_oracle_star_ref_nck_inexact(d, k, wraps):
    """Vertex core: sum over vertices of C(d_u, k) (fringeCount.cu:2152-2191). On the
    star K_{1,d} that is C(d, k) (the leaves have degree 1). REF_NCK_INEXACT is set
    exactly when the reference's u64 loop wraps on the hub."""
    ni, nl = star(d)
    mo, _ = fsgc.preprocess(k + 1, [(0, i) for i in range(1, k + 1)])
    o = oracle.count(mo, ni, nl)
    assert o["motif_class"] == 0
    assert o["exact"] == math.comb(d, k)
    ref, wrapped = ref_nck(d, k)
    assert wrapped == wraps
    assert bool(o["flags"] & fsgc.FLAG_REF_NCK_INEXACT) == wraps
    assert bool(o["flags"] & fsgc.FLAG_OVF64) == (math.comb(d, k) >= U64)
    if not wraps:
        assert o["ref_compat"] == ref % U64 == math.comb(d, k) % U64
    else:
        # what the reference prints differs from the exact count mod 2^64: the flag is
        # the only way to tell
        assert ref != math.comb(d, k) % U64


def test_schedule_guard_by_maximum_degree():
    """Wedge cores leave the pivot schedule when a codegree could reach 2^23 (the pivot
    kernel's 24-bit 2-hop counters); the other classes do not depend on the degree."""
    c2, _ = fsgc.preprocess(*C2_MOTIF)
    c4, _ = fsgc.preprocess(*C4_MOTIF)
    assert fsgc.plan_schedule(c2, (1 << 23) - 1) == "pivot"
    assert fsgc.plan_schedule(c2, 1 << 23) == "per-root"
    assert fsgc.plan_schedule(c2, 2_000_000_000) == "per-root"
    assert fsgc.plan_schedule(c4, 1 << 23) == "rank"
    assert fsgc.plan_schedule(fsgc.preprocess_pattern(1)[0], 1 << 23) == "host"
    assert fsgc.plan_schedule(fsgc.preprocess(4, [(0, 1), (0, 2), (0, 3)])[0], 10) == "edge-pass"


@pytest.mark.parametrize("case", ["nonmonotone", "range", "selfloop", "unsorted", "asymmetric", "span"])
def test_upload_rejects_malformed_csr(case):
    """Malformed graphs are refused on the host with -22 before any device work (the
    check runs before the GPU is initialised, so this holds without one)."""
    ni = np.array([0, 2, 4, 6], dtype=np.int32)
    nl = np.array([1, 2, 0, 2, 0, 1], dtype=np.int32)  # triangle
    if case == "nonmonotone":
        ni = np.array([0, 4, 2, 6], dtype=np.int32)
    elif case == "range":
        nl = nl.copy(); nl[1] = 7
    elif case == "selfloop":
        nl = nl.copy(); nl[0] = 0
    elif case == "unsorted":
        nl = np.array([2, 1, 0, 2, 0, 1], dtype=np.int32)
    elif case == "asymmetric":
        ni = np.array([0, 2, 3, 5], dtype=np.int32)
        nl = np.array([1, 2, 2, 0, 1], dtype=np.int32)  # 0 -> 1 without 1 -> 0
    elif case == "span":
        ni = np.array([0, 2, 4, 5], dtype=np.int32)
    with pytest.raises(fsgc.FsgcError) as e:
        fsgc.Graph.from_csr(ni, nl)
    assert e.value.code == -22, str(e.value)
