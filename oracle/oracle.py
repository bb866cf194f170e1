"""ctypes wrapper of liboracle.so — TEST INFRASTRUCTURE ONLY.

The oracle is the exact CPU restatement of the reference's core+fringe counting
method plus a brute-force counter (oracle/oracle.cpp). Only tests/, the smoke check
and bench.py's cpu_baseline leg use it, as the checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PREPROCESS = os.path.join(HERE, "_ref", "fringePreprocess")


class ORes(C.Structure):
    _fields_ = [
        ("exact_lo", C.c_uint64), ("exact_hi", C.c_uint64), ("ref_compat", C.c_uint64),
        ("sg_lo", C.c_uint64), ("sg_hi", C.c_uint64), ("aut", C.c_uint64), ("core_matches", C.c_uint64),
        ("flags", C.c_uint32), ("motif_class", C.c_int32), ("seconds", C.c_double), ("threads", C.c_int32),
    ]


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def _lib():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.oracle_count.argtypes = [C.c_char_p, C.c_int, C.c_longlong, P(C.c_int), P(C.c_int), C.c_int, C.c_int,
                                 C.c_int, P(ORes)]
    lib.oracle_bruteforce.argtypes = [C.c_int, C.c_int, P(C.c_int), C.c_int, C.c_longlong, P(C.c_int), P(C.c_int),
                                      P(C.c_uint64), P(C.c_uint64)]
    lib.oracle_fc3.argtypes = [P(C.c_int), P(C.c_int), P(C.c_uint64), P(C.c_uint64), P(C.c_uint32)]
    lib.oracle_functionx.argtypes = [C.c_int, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_uint64),
                                     P(C.c_uint64), P(C.c_uint32)]
    lib.oracle_ncr.argtypes = [C.c_int, C.c_int, P(C.c_uint64), P(C.c_uint64), P(C.c_uint32)]
    lib.oracle_pivot_units.argtypes = [C.c_int, C.c_longlong, P(C.c_int), P(C.c_int), C.c_int, C.c_int, C.c_longlong,
                                       C.c_int, C.c_int, P(C.c_uint64)]
    return lib


_L = None


def lib():
    global _L
    if _L is None:
        _L = _lib()
    return _L


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def count(mo: bytes, nindex, nlist, threads: int = 0, slice_mod: int = 1, slice_rem: int = 0) -> dict:
    ni = np.ascontiguousarray(nindex, dtype=np.int32)
    nl = np.ascontiguousarray(nlist, dtype=np.int32)
    r = ORes()
    rc = lib().oracle_count(mo, len(ni) - 1, len(nl), _ip(ni), _ip(nl), threads, slice_mod, slice_rem, C.byref(r))
    if rc != 0:
        raise RuntimeError(f"oracle_count failed: {rc}")
    return {
        "exact": r.exact_lo | (r.exact_hi << 64), "ref_compat": r.ref_compat, "sg": r.sg_lo | (r.sg_hi << 64),
        "aut": r.aut, "core_matches": r.core_matches, "flags": r.flags, "motif_class": r.motif_class,
        "seconds": r.seconds, "threads": r.threads,
    }


def bruteforce(nodes: int, edges, nindex, nlist) -> int:
    e = np.array([x for p in edges for x in p] or [0], dtype=np.int32)
    ni = np.ascontiguousarray(nindex, dtype=np.int32)
    nl = np.ascontiguousarray(nlist, dtype=np.int32)
    lo, hi = C.c_uint64(), C.c_uint64()
    rc = lib().oracle_bruteforce(nodes, len(edges), _ip(e), len(ni) - 1, len(nl), _ip(ni), _ip(nl), C.byref(lo),
                                 C.byref(hi))
    if rc < 0:
        raise RuntimeError("bruteforce failed")
    return lo.value | (hi.value << 64)


def fc3(fr8, venn8) -> int:
    a = (C.c_int * 8)(*fr8)
    b = (C.c_int * 8)(*venn8)
    lo, hi, fl = C.c_uint64(), C.c_uint64(), C.c_uint32()
    lib().oracle_fc3(a, b, C.byref(lo), C.byref(hi), C.byref(fl))
    return lo.value | (hi.value << 64)


def functionx(nodes: int, anch, fcnt, venn64) -> int:
    a = (C.c_int * len(anch))(*anch)
    f = (C.c_int * len(fcnt))(*fcnt)
    v = (C.c_int * 64)(*venn64)
    lo, hi, fl = C.c_uint64(), C.c_uint64(), C.c_uint32()
    lib().oracle_functionx(nodes, len(anch), a, f, v, C.byref(lo), C.byref(hi), C.byref(fl))
    return lo.value | (hi.value << 64)


def ncr(n: int, k: int) -> tuple[int, int]:
    lo, hi, fl = C.c_uint64(), C.c_uint64(), C.c_uint32()
    lib().oracle_ncr(n, k, C.byref(lo), C.byref(hi), C.byref(fl))
    return lo.value | (hi.value << 64), fl.value


def motif_graph_edges(mo: bytes):
    """(n, edges) of the motif a matching order describes: core in matching order,
    then the fringes of each anchor set (the reference's matchingOrder2ECLgraph)."""
    nodes, nas = mo[0], mo[1]
    edges = []
    for i in range(nodes - 1):
        for j in range(i + 1):
            if (mo[8 + i] >> j) & 1:
                edges.append((j, i + 1))
    v = nodes
    for a, c in zip(mo[13:13 + nas], mo[76:76 + nas]):
        for _ in range(c):
            for j in range(6):
                if (a >> j) & 1:
                    edges.append((j, v))
            v += 1
    return v, edges


PIVOT_UNITS = ("paths_d", "paths_hub", "marks_d", "marks_hub", "xvisits", "xvisit_bits", "centres", "centre_bits",
               "c3_x", "c3_search_bits", "list_elems", "walk_d", "walk_hub")


def pivot_split_budget(total_cost: int, chunk: int = 256) -> int:
    """The library's pivot work-item split threshold (pivot.hip pivot_split_budget)."""
    return max(1 << 16, chunk << 12, total_cost // 4096)


def pivot_units(nindex, nlist, ecap: int, hub_slots: int, c3: bool, chunk: int = 256, threads: int = 0) -> dict:
    """Unit counts of one pivot-schedule launch on this CSR (oracle.cpp oracle_pivot_units).
    hub_slots: the LDS hub counters of the launch (0 for launches with c3)."""
    ni = np.ascontiguousarray(nindex, dtype=np.int32)
    nl = np.ascontiguousarray(nlist, dtype=np.int32)
    deg = np.diff(ni).astype(np.int64)
    budget = pivot_split_budget(int((deg * deg).sum()), chunk)
    out = (C.c_uint64 * 13)()
    rc = lib().oracle_pivot_units(len(ni) - 1, len(nl), _ip(ni), _ip(nl), ecap, hub_slots, budget, int(c3), threads,
                                  out)
    if rc != 0:
        raise RuntimeError("oracle_pivot_units failed")
    return dict(zip(PIVOT_UNITS, (int(x) for x in out)))


def pivot_bytes(units: dict, list_elem_bytes: int, hub_launch: bool) -> int:
    """The pivot kernel's algorithmic bytes from its units (the per-unit sizes of
    pivot.hip k_wedge_pivot)."""
    s = list_elem_bytes
    # a pivot neighbour x reads adj, idx[x + 1] and the reverse-edge index (12 B); a centre
    # adj, idx twice, the support and the reverse-edge index (20 B): no searches (the
    # *_bits units are the probes the searches they replace would take)
    return (20 * units["paths_d"] + 4 * units["paths_hub"] + (20 if hub_launch else 16) * units["marks_d"]
            + 4 * units["marks_hub"] + 12 * units["xvisits"] + 20 * units["centres"]
            + (16 + s) * units["c3_x"] + (8 + s) * units["c3_search_bits"]
            + s * units["list_elems"] + 16 * units["walk_d"] + 12 * units["walk_hub"])
