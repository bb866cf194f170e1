"""Autotuner and serving loop through the C ABI (marked gpu).

The autotuner may only change time: counts after tuning equal counts before. The
server's results must equal fsgc_count of the same motif alone, whatever the mix,
the number of slots or the scratch cap; admission never holds more device scratch
than the cap."""
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("exact", "sg", "aut", "ref_compat", "core_matches", "flags", "motif_class")


def key(r):
    return tuple(getattr(r, k) for k in KEYS)


def mixed_motifs(fsgc, n=24):
    # random motifs of every class; cores of 4-5 vertices only on the small graphs
    return fsgc.random_motifs(3, n)


@pytest.fixture(scope="module")
def graphs(gpu):
    specs = ["er:3000:12000:5", "rmat:11:8:0.57:0.19:0.19:29"]
    return [(s, gpu.Graph.generate(s)) for s in specs]


def test_autotune_changes_time_not_counts(gpu):
    g = gpu.Graph.generate("rmat:12:8:0.57:0.19:0.19:3")
    mos = [m for m in gpu.random_motifs(3, 32) if m[0] <= 3]
    before = [key(g.count(m)) for m in mos]
    print("tuning", len(mos), "motifs", flush=True)
    reps = g.autotune(mos, sample=4)
    print(reps, flush=True)
    assert {r["schedule"] for r in reps} >= {"pivot", "rank", "edge-pass"}
    for r in reps:
        assert r["candidates"] >= 2 and r["best_ms"] <= r["default_ms"] + 1e-9
    after = [key(g.count(m)) for m in mos]
    assert before == after
    batch, _ = g.count_batch(mos)
    assert [key(r) for r in batch] == before
    g.untune()
    assert [key(g.count(m)) for m in mos] == before


def test_server_alone_equals_mix(gpu, graphs):
    mos = mixed_motifs(gpu)
    alone = {}
    for gi, (spec, g) in enumerate(graphs):
        for i, mo in enumerate(mos):
            if spec.startswith("rmat") and mo[0] >= 4:
                continue
            alone[(gi, i)] = key(g.count(mo))
    print("alone:", len(alone), "queries", flush=True)
    for slots in (1, 3):
        srv = gpu.Server(slots=slots)
        ids = [srv.add_graph(g) for _, g in graphs]
        tickets = {}
        # interleave the graphs so launches of both run side by side
        for i in range(len(mos)):
            for gi in ids:
                if (gi, i) in alone:
                    tickets[(gi, i)] = srv.submit(gi, mos[i])
        got = {k: srv.wait(t) for k, t in tickets.items()}
        st = srv.stats()
        srv.close()
        print("slots", slots, st, flush=True)
        for k, q in got.items():
            assert q.status == 0, k
            assert key(q.result) == alone[k], k
            assert q.done_ms >= q.start_ms >= q.submit_ms
        assert st["completed"] == len(tickets) and st["rejected"] == 0
        assert st["scratch_peak"] <= st["scratch_cap"]
        assert st["peak_concurrent"] <= slots
        assert max(q.batch_size for q in got.values()) > 1  # queries were folded


def test_server_scratch_cap_admission(gpu, graphs):
    """A cap that fits one launch at a time: results unchanged, peak within the cap."""
    mos = [m for m in mixed_motifs(gpu) if m[0] <= 3]
    spec, g = graphs[1]
    alone = [key(g.count(mo)) for mo in mos]
    cap = 64 << 20
    srv = gpu.Server(slots=4, scratch_cap=cap, max_batch=2)
    gid = srv.add_graph(g)
    ts = [srv.submit(gid, mo) for mo in mos]
    got = [srv.wait(t) for t in ts]
    st = srv.stats()
    srv.close()
    print(st, flush=True)
    assert [key(q.result) for q in got] == alone
    assert st["scratch_peak"] <= cap
    assert max(q.batch_size for q in got) <= 2


def test_server_rejects_query_over_cap(gpu, graphs):
    """No launch fits in 1 KiB of scratch: non-trivial queries fail with -12 (no kernel runs)."""
    spec, g = graphs[0]
    mo = gpu.preprocess(3, [(0, 1), (1, 2)])[0]
    srv = gpu.Server(slots=1, scratch_cap=1024)
    gid = srv.add_graph(g)
    t = srv.submit(gid, mo)
    with pytest.raises(gpu.FsgcError):
        srv.wait(t)
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_served_c5_stream_vs_oracle(gpu):
    """The served stream of the two-graph configuration, scaled down to sizes the oracle
    covers: the same generators, seeds and edge factors (R-MAT ef16 seed 23 and ef32
    seed 29), motifs of every class from the same generator (seed 5), Poisson arrivals
    from seed 5 submitted at their arrival times while earlier queries run. Every
    served result is compared with the oracle. Cores of 4-5 vertices go to the scale-4
    pair (the oracle's general-core fringe enumeration is exponential), the others to
    the scale-10/9 pair."""
    import threading
    import time

    import numpy as np

    import oracle

    pairs = {"small": ["rmat:4:16:0.57:0.19:0.19:23", "rmat:4:32:0.57:0.19:0.19:29"],
             "large": ["rmat:10:16:0.57:0.19:0.19:23", "rmat:9:32:0.57:0.19:0.19:29"]}
    csr = {s: gpu.generate_csr(s) for p in pairs.values() for s in p}
    rng = np.random.default_rng(5)
    mos = gpu.random_motifs(5, 64)
    classes = {gpu.motif_class(m) for m in mos}
    assert classes == {"vertex", "edge", "wedge", "triangle", "general"}
    queries = []
    for mo in mos:
        side = "small" if gpu.motif_class(mo) == "general" else "large"
        queries.append((pairs[side][int(rng.integers(2))], mo))
    gaps = rng.exponential(1.0 / 200.0, size=len(queries))  # 200 queries/s offered
    srv = gpu.Server(slots=2)
    graphs = {s: gpu.Graph.from_csr(*csr[s]) for s in csr}
    ids = {s: srv.add_graph(g) for s, g in graphs.items()}
    tickets = [None] * len(queries)

    def submitter():
        start = time.perf_counter()
        at = 0.0
        for i, (s, mo) in enumerate(queries):
            at += gaps[i]
            d = start + at - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(ids[s], mo)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    got = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    print("served", st, flush=True)
    assert st["completed"] == len(queries) and st["rejected"] == 0
    assert max(q.batch_size for q in got) > 1
    from test_gpu_parity import same
    for i, ((s, mo), q) in enumerate(zip(queries, got)):
        print("query", i, gpu.motif_class(mo), s, flush=True)
        assert q.status == 0
        same(q.result, oracle.count(mo, *csr[s]), f"query {i} ({gpu.motif_class(mo)}) on {s}")
    for g in graphs.values():
        g.free()
