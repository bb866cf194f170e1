"""The triple-region decomposition of wedge cores with one fringe on each endpoint
(the pivot schedule's c3-free evaluation), restated with Python integers and pinned
against brute force and the oracle's per-wedge fc3 (FC:233-327, FC:1513-1650).

A wedge core with f fringes on its centre X and one fringe on each endpoint Y, Z
(single-vertex anchor sets) has the fringe count
    F = V - b0 (A + B) + (b0 - b2) s,
    V = b0 [(A - P)(B - Q) - R + A + B] + b1 [P (B - Q) + Q (A - P)] + b2 P Q,
with b_c = C(n_X - c, f), n_X = |N(X) minus the core|, A = |N(Y) minus the core|,
B = |N(Z) minus the core|, P = |A & N(X)|, Q = |B & N(X)|, R = |A & B| and
s = |N(X) & N(Y) & N(Z)|. Summed over all wedges, the last two terms only need, per
centre u, S_u = sum_{x in N(u)} C(t_ux, 2) (the pairs of N(u) inside one N(x)) and
K_u = (d_u - 1) sum_x d_x - d_u (d_u - 1) - sum_x t_ux (the sum of A + B over the pairs
of N(u)), so the kernel never needs the per-wedge triple count.
"""
import itertools
import random
from math import comb

import numpy as np
import pytest

import fsgc
import oracle


def binom(n, k):
    return comb(n, k) if 0 <= k <= n else 0


def v_term(f, nx, a, b, p, q, r):
    b0, b1, b2 = binom(nx, f), binom(nx - 1, f), binom(nx - 2, f)
    return b0 * ((a - p) * (b - q) - r + a + b) + b1 * (p * (b - q) + q * (a - p)) + b2 * p * q


def test_identity_against_brute_force_on_sets():
    """F = V - b0 (A + B) + (b0 - b2) s equals the number of (a, b, S) with a in A,
    b in B, a != b, S an f-subset of U minus {a, b}, on random set systems."""
    rng = random.Random(5)
    for _ in range(400):
        universe = range(rng.randint(0, 12))
        U = {x for x in universe if rng.random() < 0.6}
        A = {x for x in universe if rng.random() < 0.5}
        B = {x for x in universe if rng.random() < 0.5}
        f = rng.randint(0, 4)
        brute = sum(binom(len(U - {a, b}), f) for a in A for b in B if a != b)
        nx, a, b = len(U), len(A), len(B)
        p, q, r, s = len(A & U), len(B & U), len(A & B), len(A & B & U)
        b0, b2 = binom(nx, f), binom(nx - 2, f)
        assert v_term(f, nx, a, b, p, q, r) - b0 * (a + b) + (b0 - b2) * s == brute
        # V itself is never negative (the kernel accumulates it unsigned)
        assert v_term(f, nx, a, b, p, q, r) >= 0


def decomposed_sg(f, ni, nl):
    """Sum over every centre u and every pair {w, v} of N(u) of V (the triangle pairs
    with their three rotations are the same sum), plus the per-centre terms."""
    n = len(ni) - 1
    adj = [set(map(int, nl[ni[u]:ni[u + 1]])) for u in range(n)]
    deg = [len(s) for s in adj]
    total = 0
    for u in range(n):
        nbrs = sorted(adj[u])
        nx = deg[u] - 2
        for w, v in itertools.combinations(nbrs, 2):
            e = 1 if v in adj[w] else 0
            a, b = deg[w] - 1 - e, deg[v] - 1 - e
            p, q = len(adj[u] & adj[w]) - e, len(adj[u] & adj[v]) - e
            r = len(adj[w] & adj[v]) - 1
            total += v_term(f, nx, a, b, p, q, r)
        b0, b2 = binom(deg[u] - 2, f), binom(deg[u] - 4, f)
        t = [len(adj[u] & adj[x]) for x in nbrs]
        s_u = sum(comb(tx, 2) for tx in t)
        k_u = (deg[u] - 1) * sum(deg[x] for x in nbrs) - deg[u] * (deg[u] - 1) - sum(t)
        total += (b0 - b2) * s_u - b0 * k_u
    return total


def wedge_motif(f):
    # centre 1 with f fringes, one fringe on each endpoint 0 and 2
    edges = [(0, 1), (1, 2), (0, 3), (2, 4)] + [(1, 5 + i) for i in range(f)]
    return fsgc.preprocess(5 + f, edges)[0]


@pytest.mark.parametrize("spec", ["er:60:400:3", "rmat:7:8:0.57:0.19:0.19:5", "chunglu:150:2.1:8:40:2"])
@pytest.mark.parametrize("f", [1, 2, 3, 5])
def test_decomposition_equals_oracle(spec, f):
    """The decomposed sum equals the oracle's per-wedge fc3 sum (the reference's
    per-root schedule with its filters and rotations), triangles included."""
    ni, nl = fsgc.generate_csr(spec)
    mo = wedge_motif(f)
    assert fsgc.motif_class(mo) == "wedge"
    o = oracle.count(mo, ni, nl)
    assert decomposed_sg(f, ni, nl) == o["sg"]


def test_decomposition_on_complete_graph():
    """K_n: every pair of a centre's neighbours is a triangle; the count is
    n (n-1)(n-2)/2 * (n-3)(n-4) * C(n-5, f) / |Aut| embeddings of the labelled motif."""
    n = 12
    ni = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.int32)
    nl = np.array([v for u in range(n) for v in range(n) if v != u], dtype=np.int32)
    for f in (1, 2, 3):
        o = oracle.count(wedge_motif(f), ni, nl)
        assert decomposed_sg(f, ni, nl) == o["sg"]
        # centre, unordered endpoint pair, an ordered pair of endpoint fringes, f centre fringes
        assert o["exact"] == n * comb(n - 1, 2) * (n - 3) * (n - 4) * comb(n - 5, f)
