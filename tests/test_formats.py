"""ECLgraph I/O, CSR validation, synthetic generators, and the C ABI surface. CPU only."""
import ctypes
import hashlib
import os
import re
import subprocess
import struct

import numpy as np
import pytest

import fsgc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "fsgc.h")).read()
    names = set(re.findall(r"\b(fsgc_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 18
    lib = ctypes.CDLL(fsgc.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_generators_are_deterministic_and_valid():
    for spec in ("er:4096:32768:1", "rmat:12:16:0.57:0.19:0.19:7", "chunglu:20000:2.1:20:1000:19"):
        a = fsgc.generate_csr(spec)
        b = fsgc.generate_csr(spec)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        ni, nl = a
        n = len(ni) - 1
        assert ni[0] == 0 and ni[-1] == len(nl)
        deg = np.diff(ni)
        src = np.repeat(np.arange(n), deg)
        assert np.all(nl != src)                                   # no self loops
        same = src[1:] == src[:-1]
        assert np.all(nl[1:][same] > nl[:-1][same])                # sorted, no duplicates
        key = np.sort(src.astype(np.int64) * n + nl)
        rkey = np.sort(nl.astype(np.int64) * n + src)
        assert np.array_equal(key, rkey)                           # symmetric


def test_er_config_shape():
    """G(4096, 32768) from seed 1: 32,768 distinct undirected pairs, no self loops, and
    a different seed gives a different graph of the same size."""
    ni, nl = fsgc.generate_csr("er:4096:32768:1")
    assert len(ni) == 4097 and len(nl) == 65536
    src = np.repeat(np.arange(4096), np.diff(ni))
    pairs = {(int(a), int(b)) for a, b in zip(src, nl) if a < b}
    assert len(pairs) == 32768
    ni2, nl2 = fsgc.generate_csr("er:4096:32768:2")
    assert len(nl2) == 65536
    assert hashlib.sha256(ni.tobytes() + nl.tobytes()).digest() != hashlib.sha256(ni2.tobytes() + nl2.tobytes()).digest()


def test_chung_lu_degree_profile():
    ni, nl = fsgc.generate_csr("chunglu:200000:2.1:20:5000:19")
    deg = np.diff(ni)
    assert 15 < deg.mean() < 22
    assert deg.max() > 2000


def test_egr_round_trip_and_errors(tmp_path):
    exe = os.path.join(os.path.dirname(fsgc.LIB_PATH), "bin", "fsgc_gen")
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    p = tmp_path / "g.egr"
    r = subprocess.run([exe, "graph", "er:100:300:3", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    raw = p.read_bytes()
    n, m = struct.unpack("<ii", raw[:8])
    assert n == 100 and m == 600 and len(raw) == 8 + 4 * (n + 1) + 4 * m
    ni, nl = fsgc.generate_csr(f"egr:{p}")
    a = fsgc.generate_csr("er:100:300:3")
    assert np.array_equal(ni, a[0]) and np.array_equal(nl, a[1])
    info = subprocess.run([exe, "info", str(p)], capture_output=True, text=True, timeout=60).stdout
    assert "nodes: 100" in info and "edges: 600" in info and "valid: yes" in info
    # truncated file and bad header are rejected with the reference's messages
    (tmp_path / "t.egr").write_bytes(raw[:20])
    with pytest.raises(fsgc.FsgcError, match="neighbor index"):
        fsgc.generate_csr(f"egr:{tmp_path / 't.egr'}")
    (tmp_path / "z.egr").write_bytes(struct.pack("<ii", 0, 0))
    with pytest.raises(fsgc.FsgcError, match="too low"):
        fsgc.generate_csr(f"egr:{tmp_path / 'z.egr'}")
    with pytest.raises(fsgc.FsgcError, match="could not open"):
        fsgc.generate_csr(f"egr:{tmp_path / 'missing.egr'}")


def test_random_motif_classes():
    mos = fsgc.random_motifs(3, 64)
    classes = [min(4, {1: 0, 2: 1}.get(m[0], 2 if m[0] == 3 and bin(m[9]).count("1") == 1 else (3 if m[0] == 3 else 4))) for m in mos]
    assert set(classes) == {0, 1, 2, 3, 4}


def test_motif_class_helper_matches_oracle():
    """fsgc.motif_class (host-side, from the matching order) names the class the
    oracle dispatches on, for the C3 generator's motifs."""
    import oracle
    ni, nl = fsgc.generate_csr("er:12:30:1")
    for mo in fsgc.random_motifs(3, 64):
        o = oracle.count(mo, ni, nl)
        assert fsgc.CLASS_NAMES[o["motif_class"]] == fsgc.motif_class(mo)


def test_reproduce_library_matches_reference_families():
    """The edges/s harness finds every motif file family of the reference's library."""
    from fsgc import reproduce
    sizes = {f: len(reproduce.library(f)) for f in reproduce.FAMILIES}
    assert sizes == {"vertex": 5, "edge": 6, "wedge": 4, "triangle": 4, "tails": 11, "tri_fringe": 11,
                     "fig3wedge": 11}


@pytest.mark.parametrize("spec", ["er:4096:32768:1", "rmat:20:16:0.57:0.19:0.19:7", "rmat:22:16:0.57:0.19:0.19:11",
                                  "chunglu:4000000:2.1:20:100000:19", "rmat:23:16:0.57:0.19:0.19:23",
                                  "rmat:21:32:0.57:0.19:0.19:29"])
def test_config_graphs_match_committed_sha256(spec):
    """The generators still produce the committed benchmark graphs: the ER, R-MAT and
    Chung-Lu graphs of all five configurations (graph_sha256.json, written by
    make_graph_sha.py)."""
    import hashlib
    import json
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "graph_sha256.json")))
    ni, nl = fsgc.generate_csr(spec)
    h = hashlib.sha256()
    h.update(ni.astype("<i4").tobytes())
    h.update(nl.astype("<i4").tobytes())
    assert h.hexdigest() == ref[spec]["sha256"], spec
    assert len(nl) == ref[spec]["directed_edges"]


def test_relabel_order_is_the_reference_key_order():
    """The host renumbering the relabel GPU tests compare against sorts vertices as the
    reference's relabelByDegree does: keys (degree << idbits) + id sorted descending
    (fringeCount.cu:1064-1075, 1241-1251), or ascending for the opposite order; and it
    keeps every count (oracle on both numberings)."""
    import oracle
    from test_gpu_relabel import host_relabel
    ni, nl = fsgc.generate_csr("rmat:8:8:0.57:0.19:0.19:3")
    n = len(ni) - 1
    deg = np.diff(ni).astype(np.uint64)
    idbits = max(1, int(n - 1).bit_length())
    keys = (deg << np.uint64(idbits)) | np.arange(n, dtype=np.uint64)
    for order in ("desc", "asc"):
        mp, a, b = host_relabel(ni, nl, order)
        srt = np.sort(keys)[::-1] if order == "desc" else np.sort(keys)
        perm = (srt & np.uint64((1 << idbits) - 1)).astype(np.int64)
        assert np.array_equal(mp[perm], np.arange(n))
        for p in (2, 7, 9, 23):
            mo = fsgc.preprocess_pattern(p)[0]
            x, y = oracle.count(mo, ni, nl), oracle.count(mo, a, b)
            assert (x["exact"], x["sg"], x["aut"]) == (y["exact"], y["sg"], y["aut"]), (order, p)
