"""ctypes binding of libfsgc.so (the C ABI in include/fsgc.h).

This is the host-side mirror used by the tests and the benchmark: graphs are made
resident with `Graph.*`, motifs are 139-byte matching orders (`bytes`), and counts
come back as `Result` with the exact unsigned 128-bit value and the u64 value the
reference's fringeCount prints. There is no fallback: if libfsgc.so is missing the
import raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libfsgc.so")
MO_BYTES = 139

FLAG_OVF64 = 1 << 0
FLAG_OVF128 = 1 << 1
FLAG_REF_NCK_INEXACT = 1 << 2
FLAG_AUT_REMAINDER = 1 << 3
FLAG_REF_AUT_TRUNC = 1 << 4

CLASS_NAMES = ["vertex", "edge", "wedge", "triangle", "general"]


class _Result(C.Structure):
    _fields_ = [
        ("exact_lo", C.c_uint64), ("exact_hi", C.c_uint64), ("ref_compat", C.c_uint64),
        ("sg_lo", C.c_uint64), ("sg_hi", C.c_uint64), ("aut", C.c_uint64),
        ("core_matches", C.c_uint64), ("flags", C.c_uint32), ("motif_class", C.c_int32),
        ("kernel_ms", C.c_double), ("wall_ms", C.c_double), ("alg_bytes", C.c_double), ("index_ms", C.c_double),
    ]


class Tune(C.Structure):
    _fields_ = [
        ("block", C.c_int32), ("grid", C.c_int32), ("chunk", C.c_int32), ("hub_degree", C.c_int32),
        ("lds_cap", C.c_int32), ("slice_mod", C.c_int32), ("slice_rem", C.c_int32), ("schedule", C.c_int32),
    ]


class TuneReport(C.Structure):
    _fields_ = [("schedule", C.c_int32), ("candidates", C.c_int32), ("best", Tune),
                ("best_ms", C.c_double), ("default_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_batch", C.c_int32), ("scratch_cap", C.c_uint64),
                ("grid", C.c_int32), ("reserved", C.c_int32)]


class _Query(C.Structure):
    _fields_ = [("result", _Result), ("status", C.c_int32), ("graph", C.c_int32), ("batch_size", C.c_int32),
                ("slot", C.c_int32), ("submit_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("scratch_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
                ("rejected", C.c_uint64), ("scratch_cap", C.c_uint64), ("scratch_peak", C.c_uint64),
                ("slots", C.c_int32), ("peak_concurrent", C.c_int32)]


SCHEDULE_NAMES = ["per-root", "pivot", "rank", "edge-pass", "clique"]
RELABEL_ORDERS = {"desc": 0, "asc": 1}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libfsgc.so not built at {LIB_PATH}; run `make -C fringe-sgc_amd`")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "fsgc_init": (C.c_int, [C.c_int]),
        "fsgc_graph_load_egr": (C.c_int, [C.c_char_p, P(C.c_void_p)]),
        "fsgc_graph_from_csr": (C.c_int, [C.c_int32, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_void_p)]),
        "fsgc_graph_validate": (C.c_int, [C.c_void_p]),
        "fsgc_graph_info": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
        "fsgc_graph_free": (None, [C.c_void_p]),
        "fsgc_graph_relabel": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_void_p), P(C.c_int32), P(C.c_double)]),
        "fsgc_graph_copy_csr": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_int32)]),
        "fsgc_count": (C.c_int, [C.c_void_p, C.c_char_p, P(Tune), P(_Result)]),
        "fsgc_count_batch": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, P(Tune), P(_Result), P(C.c_double)]),
        "fsgc_plan_schedule": (C.c_int, [C.c_char_p, C.c_int32, P(C.c_int32), P(C.c_int32)]),
        "fsgc_count_noalloc": (C.c_int, [C.c_void_p, C.c_char_p, P(_Result)]),
        "fsgc_preprocess": (C.c_int, [C.c_int, C.c_int, P(C.c_int), C.c_char_p, C.c_char_p, C.c_size_t,
                                      C.c_char_p, C.c_size_t]),
        "fsgc_preprocess_pattern": (C.c_int, [C.c_int, C.c_char_p, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]),
        "fsgc_print_matching_order": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
        "fsgc_generate": (C.c_int, [C.c_char_p, P(C.c_int32), P(C.c_int32), P(P(C.c_int32)), P(P(C.c_int32))]),
        "fsgc_graph_generate": (C.c_int, [C.c_char_p, P(C.c_void_p)]),
        "fsgc_free_host": (None, [C.c_void_p]),
        "fsgc_random_motifs": (C.c_int, [C.c_uint64, C.c_int, C.c_char_p]),
        "fsgc_random_patterns": (C.c_int, [C.c_uint64, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int)]),
        "fsgc_batch_create": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, P(Tune), P(C.c_void_p)]),
        "fsgc_batch_run": (C.c_int, [C.c_void_p, P(_Result), P(C.c_double)]),
        "fsgc_batch_free": (None, [C.c_void_p]),
        "fsgc_autotune": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_int32, P(TuneReport), P(C.c_int32)]),
        "fsgc_graph_untune": (C.c_int, [C.c_void_p]),
        "fsgc_server_create": (C.c_int, [P(ServerConfig), P(C.c_void_p)]),
        "fsgc_server_add_graph": (C.c_int, [C.c_void_p, C.c_void_p, P(C.c_int32)]),
        "fsgc_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, P(C.c_uint64)]),
        "fsgc_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, P(_Query)]),
        "fsgc_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, P(_Query)]),
        "fsgc_server_stats_get": (C.c_int, [C.c_void_p, P(ServerStats)]),
        "fsgc_server_clock_ms": (C.c_double, [C.c_void_p]),
        "fsgc_server_destroy": (None, [C.c_void_p]),
        "fsgc_last_error": (C.c_char_p, []),
        "fsgc_strerror": (C.c_char_p, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class FsgcError(RuntimeError):
    """A negative return code of the C ABI (`code`) with its message."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def _check(rc: int):
    if rc != 0:
        raise FsgcError(rc, f"{lib.fsgc_strerror(rc).decode()} ({rc}): {lib.fsgc_last_error().decode()}")


@dataclass
class Result:
    exact: int
    ref_compat: int
    sg: int
    aut: int
    core_matches: int
    flags: int
    motif_class: int
    kernel_ms: float
    wall_ms: float
    alg_bytes: float = 0.0
    index_ms: float = 0.0

    @property
    def class_name(self) -> str:
        return CLASS_NAMES[self.motif_class]


def _result(r: _Result) -> Result:
    return Result(exact=r.exact_lo | (r.exact_hi << 64), ref_compat=r.ref_compat, sg=r.sg_lo | (r.sg_hi << 64),
                  aut=r.aut, core_matches=r.core_matches, flags=r.flags, motif_class=r.motif_class,
                  kernel_ms=r.kernel_ms, wall_ms=r.wall_ms, alg_bytes=r.alg_bytes, index_ms=r.index_ms)


def motif_class(mo: bytes) -> str:
    """Class of a matching order by its core (fringeCount.cu:2244-2252): 1 vertex,
    2 edge, 3 wedge or triangle (list[1] names one or two earlier core vertices),
    4-6 general."""
    nodes = mo[0]
    if nodes <= 1:
        return "vertex"
    if nodes == 2:
        return "edge"
    if nodes == 3:
        return "wedge" if bin(mo[9]).count("1") == 1 else "triangle"
    return "general"


def plan_schedule(mo: bytes, maxdeg: int) -> str:
    """Name of the kernel schedule fsgc_count uses for `mo` on a graph of maximum
    degree `maxdeg` ("host" for the vertex and edge counts). Host only."""
    s, c = C.c_int32(), C.c_int32()
    _check(lib.fsgc_plan_schedule(mo, maxdeg, C.byref(s), C.byref(c)))
    return "host" if s.value < 0 else SCHEDULE_NAMES[s.value]


def init(device: int = 0):
    _check(lib.fsgc_init(device))


def _csr_ptrs(nindex, nlist):
    import numpy as np
    ni = np.ascontiguousarray(nindex, dtype=np.int32)
    nl = np.ascontiguousarray(nlist, dtype=np.int32)
    return ni, nl, ni.ctypes.data_as(C.POINTER(C.c_int32)), nl.ctypes.data_as(C.POINTER(C.c_int32))


class Graph:
    """A CSR graph resident in GPU memory."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_csr(cls, nindex, nlist) -> "Graph":
        ni, nl, pi, pl = _csr_ptrs(nindex, nlist)
        h = C.c_void_p()
        _check(lib.fsgc_graph_from_csr(len(ni) - 1, len(nl), pi, pl, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_egr(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(lib.fsgc_graph_load_egr(path.encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, spec: str) -> "Graph":
        h = C.c_void_p()
        _check(lib.fsgc_graph_generate(spec.encode(), C.byref(h)))
        return cls(h.value)

    def info(self):
        n, m, d = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib.fsgc_graph_info(self._h, C.byref(n), C.byref(m), C.byref(d)))
        return n.value, m.value, d.value

    def validate(self):
        _check(lib.fsgc_graph_validate(self._h))

    def relabel(self, order: str = "desc"):
        """A new resident graph renumbered by degree (relabelByDegree, fringeCount.cu:
        1221-1288): "desc" highest degree first, ties by the larger id (the reference's
        order); "asc" lowest first, ties by the smaller id. Returns (graph, map, ms) with
        map[old] = new as an int32 array."""
        import numpy as np
        n, _, _ = self.info()
        mp = np.empty(n, dtype=np.int32)
        h = C.c_void_p()
        ms = C.c_double()
        _check(lib.fsgc_graph_relabel(self._h, RELABEL_ORDERS[order], C.byref(h),
                                      mp.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(ms)))
        return Graph(h.value), mp, ms.value

    def csr(self):
        """The resident CSR copied to the host: (nindex, nlist) int32 arrays."""
        import numpy as np
        n, m, _ = self.info()
        ni = np.empty(n + 1, dtype=np.int32)
        nl = np.empty(max(1, m), dtype=np.int32)
        _check(lib.fsgc_graph_copy_csr(self._h, ni.ctypes.data_as(C.POINTER(C.c_int32)),
                                       nl.ctypes.data_as(C.POINTER(C.c_int32))))
        return ni, nl[:m]

    def count(self, mo: bytes, tune: Tune | None = None) -> Result:
        assert len(mo) == MO_BYTES
        r = _Result()
        _check(lib.fsgc_count(self._h, mo, C.byref(tune) if tune else None, C.byref(r)))
        return _result(r)

    def count_noalloc(self, mo: bytes) -> Result:
        """fsgc_count_noalloc: the same result as count() for vertex and closed-form
        edge motifs, with no allocation (the graph's supports must be built)."""
        r = _Result()
        _check(lib.fsgc_count_noalloc(self._h, mo, C.byref(r)))
        return _result(r)

    def count_batch(self, mos: list[bytes], tune: Tune | None = None):
        k = len(mos)
        buf = b"".join(mos)
        out = (_Result * max(1, k))()
        ms = C.c_double()
        _check(lib.fsgc_count_batch(self._h, buf, k, C.byref(tune) if tune else None, out, C.byref(ms)))
        return [_result(out[i]) for i in range(k)], ms.value

    def capture_batch(self, mos: list[bytes], tune: Tune | None = None) -> "CapturedBatch":
        """The motifs' launches captured once as a HIP graph (replay with .run())."""
        h = C.c_void_p()
        _check(lib.fsgc_batch_create(self._h, b"".join(mos), len(mos), C.byref(tune) if tune else None, C.byref(h)))
        return CapturedBatch(h, len(mos), self)

    def autotune(self, mos: list[bytes], sample: int = 16) -> list[dict]:
        """Times launch shapes per schedule on every sample-th work item and keeps the
        fastest for this graph (used by later counts for fields left at zero)."""
        buf = b"".join(mos)
        reps = (TuneReport * 5)()
        nr = C.c_int32()
        _check(lib.fsgc_autotune(self._h, buf, len(mos), sample, reps, C.byref(nr)))
        out = []
        for r in reps[: nr.value]:
            out.append({"schedule": SCHEDULE_NAMES[r.schedule], "candidates": r.candidates,
                        "block": r.best.block, "grid": r.best.grid, "chunk": r.best.chunk, "hub_degree": r.best.hub_degree,
                        "lds_cap": r.best.lds_cap, "best_ms": r.best_ms, "default_ms": r.default_ms})
        return out

    def untune(self):
        _check(lib.fsgc_graph_untune(self._h))

    def free(self):
        if self._h:
            lib.fsgc_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CapturedBatch:
    """A batch replayed from one captured HIP graph."""

    def __init__(self, handle, k: int, graph: Graph):
        self._h = handle
        self._k = k
        self._graph = graph  # the graph must outlive the batch

    def run(self):
        out = (_Result * max(1, self._k))()
        ms = C.c_double()
        _check(lib.fsgc_batch_run(self._h, out, C.byref(ms)))
        return [_result(out[i]) for i in range(self._k)], ms.value

    def free(self):
        if self._h:
            lib.fsgc_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_csr(spec: str):
    """Host CSR arrays (numpy int32) for a synthetic graph spec."""
    import numpy as np
    n, m = C.c_int32(), C.c_int32()
    pi, pl = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    _check(lib.fsgc_generate(spec.encode(), C.byref(n), C.byref(m), C.byref(pi), C.byref(pl)))
    try:
        ni = np.ctypeslib.as_array(pi, shape=(n.value + 1,)).copy()
        nl = np.ctypeslib.as_array(pl, shape=(max(1, m.value),)).copy()[: m.value]
    finally:
        lib.fsgc_free_host(pi)
        lib.fsgc_free_host(pl)
    return ni, nl


def preprocess(nodes: int, edges) -> tuple[bytes, str]:
    flat = [int(x) for e in edges for x in e]
    arr = (C.c_int * max(1, len(flat)))(*flat)
    mo = C.create_string_buffer(MO_BYTES)
    log = C.create_string_buffer(1 << 16)
    _check(lib.fsgc_preprocess(nodes, len(flat) // 2, arr, mo, log, len(log), None, 0))
    return mo.raw, log.value.decode()


def preprocess_pattern(number: int) -> tuple[bytes, str]:
    mo = C.create_string_buffer(MO_BYTES)
    log = C.create_string_buffer(1 << 16)
    _check(lib.fsgc_preprocess_pattern(number, mo, log, len(log), None, 0))
    return mo.raw, log.value.decode()


def random_motifs(seed: int, count: int) -> list[bytes]:
    buf = C.create_string_buffer(MO_BYTES * max(1, count))
    _check(lib.fsgc_random_motifs(seed, count, buf))
    return [buf.raw[i * MO_BYTES:(i + 1) * MO_BYTES] for i in range(count)]


def random_patterns(seed: int, count: int):
    """The random motifs as (nodes, edge list) before preprocessing."""
    nodes = (C.c_int * max(1, count))()
    ned = (C.c_int * max(1, count))()
    edges = (C.c_int * (128 * max(1, count)))()
    _check(lib.fsgc_random_patterns(seed, count, nodes, ned, edges))
    return [(nodes[i], [(edges[i * 128 + 2 * k], edges[i * 128 + 2 * k + 1]) for k in range(ned[i])])
            for i in range(count)]


def print_matching_order(mo: bytes) -> str:
    out = C.create_string_buffer(8192)
    _check(lib.fsgc_print_matching_order(mo, out, len(out)))
    return out.value.decode()


def live_fields(mo: bytes) -> dict:
    """The fields of a matching order that carry meaning (the reference leaves the
    rest uninitialised)."""
    nodes, nas = mo[0], mo[1]
    return {
        "nodes": nodes, "anchorsets": nas, "startdeg": mo[2],
        "destdeg": list(mo[3:3 + max(0, nodes - 1)]), "list": list(mo[8:8 + max(0, nodes - 1)]),
        "anch": list(mo[13:13 + nas]), "fcnt": list(mo[76:76 + nas]),
    }


@dataclass
class QueryResult:
    result: Result
    status: int
    graph: int
    batch_size: int
    slot: int
    submit_ms: float
    start_ms: float
    done_ms: float
    scratch_bytes: int

    @property
    def latency_ms(self) -> float:
        return self.done_ms - self.submit_ms


def _query(q: _Query) -> QueryResult:
    return QueryResult(result=_result(q.result), status=q.status, graph=q.graph, batch_size=q.batch_size,
                       slot=q.slot, submit_ms=q.submit_ms, start_ms=q.start_ms, done_ms=q.done_ms,
                       scratch_bytes=q.scratch_bytes)


class Server:
    """The library's serving loop: resident graphs, queries folded into shared
    launches on `slots` concurrent streams, admitted by device scratch."""

    def __init__(self, slots: int = 2, max_batch: int = 8, scratch_cap: int = 0, grid: int = 0):
        cfg = ServerConfig(slots=slots, max_batch=max_batch, scratch_cap=scratch_cap, grid=grid)
        h = C.c_void_p()
        _check(lib.fsgc_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._graphs = []

    def add_graph(self, g: Graph) -> int:
        gid = C.c_int32()
        _check(lib.fsgc_server_add_graph(self._h, g._h, C.byref(gid)))
        self._graphs.append(g)  # keep it alive as long as the server
        return gid.value

    def submit(self, graph: int, mo: bytes) -> int:
        assert len(mo) == MO_BYTES
        t = C.c_uint64()
        _check(lib.fsgc_server_submit(self._h, graph, mo, C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> QueryResult:
        q = _Query()
        _check(lib.fsgc_server_wait(self._h, ticket, C.byref(q)))
        return _query(q)

    def poll(self, ticket: int):
        q = _Query()
        rc = lib.fsgc_server_poll(self._h, ticket, C.byref(q))
        if rc == 0:
            return None
        if rc < 0:
            _check(rc)
        return _query(q)

    def clock_ms(self) -> float:
        return lib.fsgc_server_clock_ms(self._h)

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib.fsgc_server_stats_get(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib.fsgc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
