#!/usr/bin/env python3
"""Benchmark of the Fringe-SGC counting path on one MI355X.

Workload (default): R-MAT scale 20, edge factor 16, (a, b, c) = (0.57, 0.19, 0.19),
seed 7, symmetrised and deduplicated (1,048,576 vertices, 31,400,418 directed
entries), resident in HBM; motif: wedge core with 3 fringe vertices on the centre and
one on each endpoint (`8 0 1 1 2 1 3 1 4 1 5 0 6 2 7`). One step = one exact count
of that motif over the whole graph (fsgc_count). The graph is renumbered by ascending
degree once at load (fsgc_graph_relabel, reported as relabel_ms; --relabel none
counts on the generator's ids, with the same result), and the per-graph edge index
is built once, before the timed region, and reported as index_build_ms.

Metric: core matches per second (core embeddings whose fringe formula is evaluated).
Also reported: the roofline of the dominant kernel (k_wedge_pivot for the default
workload, k_triangle_rank for --workload c1 / c4) from its
algorithmic bytes and its HIP-event time, and a CPU baseline: the oracle (the
reference's core+fringe method restated in C++/OpenMP) on a deterministic slice of
root vertices, checked bit-exact against the GPU on the same slice. Beside it, the
two other parts of the metric on small graphs: batched vs one-motif-per-launch motifs/s
over the random-motif generator (seed 3) and the served query stream's p50/p99.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-extras] [--relabel asc|desc|none]
    python bench.py --batch <graph spec> [--classes vertex,edge,...]   (plain vs batched)
    python bench.py --serve <spec>,<spec> [--rate Q] [--queries N]    (served stream)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fringe-sgc_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "c2": {
        "spec": "rmat:20:16:0.57:0.19:0.19:7",
        "motif": (8, [(0, 1), (1, 2), (1, 3), (1, 4), (1, 5), (0, 6), (2, 7)]),
        "name": "R-MAT s20 ef16 (0.57,0.19,0.19) seed 7; wedge core + 3 centre / 1+1 endpoint fringes",
        "cpu_slice": (4096,),
        "kernel": "k_wedge_pivot",
        # counted on the graph renumbered by ascending degree (2.24 s against 3.47 s on
        # the generator's ids and 3.31 s on the reference's descending order)
        "relabel": "asc",
        # the motif's count is linear in the triple region: the pivot kernel runs
        # without c3 and k_wedge_lin_centres adds the per-centre terms
        "linear": True,
    },
    "c1": {
        "spec": "er:4096:32768:1",
        "motif": (5, [(0, 1), (1, 2), (0, 2), (0, 3), (0, 4)]),
        "name": "ER G(4096, 32768) seed 1; triangle + 2 fringes on one core vertex",
        "cpu_slice": (1,),
        "kernel": "k_triangle_rank",
    },
    "c4": {
        "spec": "chunglu:4000000:2.1:20:100000:19",
        "motif": (9, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (0, 4), (0, 5), (0, 6), (1, 7), (1, 8)]),
        "name": "Chung-Lu n=4e6 gamma 2.1 mean 20 max 1e5 seed 19; 4-clique + 3/2 fringes",
        "cpu_slice": (4096,),
        "kernel": "k_triangle_rank",
    },
}


# the small batched comparison and served stream reported beside the main workload
EXTRA_BATCH = "rmat:16:16:0.57:0.19:0.19:11"
EXTRA_SERVE = ["rmat:16:16:0.57:0.19:0.19:23", "rmat:14:32:0.57:0.19:0.19:29"]


def physical_cores() -> int:
    try:
        import psutil
        c = psutil.cpu_count(logical=False)
        if c:
            return int(c)
    except Exception:
        pass
    return os.cpu_count() or 1


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    p = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return float(d["traffic_bytes_per_launch"])
    except Exception:
        return None


def _result_fields(r):
    return {"class": r.class_name, "exact": str(r.exact), "sg": str(r.sg), "aut": r.aut, "ref_compat": r.ref_compat,
            "core_matches": r.core_matches, "flags": r.flags}


def batch_compare(spec, seed=3, count=64, classes=None, out_path=None, emit=True, quick=False, part="both",
                  relabel="none"):
    """Plain path (one fsgc_count per motif) against the batched path (one
    fsgc_count_batch over all motifs) on the same resident graph: motifs/s of each,
    per-class times, and per-motif equality of every result. `classes` restricts the
    motif set to the named classes (the others are listed as excluded). `part` runs
    only the plain or only the batched side ("plain" / "batched"), for graphs where
    both do not fit one run; batch_merge() then compares two such records. `relabel`
    counts on the graph renumbered by degree (counts are label-invariant)."""
    import fsgc
    fsgc.init(0)
    t0 = time.perf_counter()
    g = fsgc.Graph.generate(spec)
    relabel_ms = None
    if relabel != "none":
        g0 = g
        g, _, relabel_ms = g0.relabel(relabel)
        g0.free()
    n, m, maxdeg = g.info()
    print(f"# graph {spec}: n={n} m={m} maxdeg={maxdeg} relabel={relabel} ({time.perf_counter() - t0:.1f} s)",
          flush=True)
    mos_all = fsgc.random_motifs(seed, count)
    # general cores split by schedule: "clique" (4-clique cores) and "general" (the rest)
    cls_of = {i: ("clique" if fsgc.plan_schedule(mo, maxdeg) == "clique" else fsgc.motif_class(mo))
              for i, mo in enumerate(mos_all)}
    keep = [i for i in range(count) if classes is None or cls_of[i] in classes]
    mos = [mos_all[i] for i in keep]
    # the automorphism sums (per motif, always whole) and the per-graph index are built
    # once, by a batch over a thin root slice, so neither side below pays for them
    t0 = time.perf_counter()
    g.count_batch(mos, fsgc.Tune(slice_mod=1 << 20, slice_rem=1))
    print(f"# warm batch (index + automorphism sums, root slice u % 2^20 == 1) {time.perf_counter() - t0:.1f} s",
          flush=True)
    plain, t_plain, per_class, plain_secs = [], 0.0, {}, []
    if part in ("both", "plain"):
        for i, mo in zip(keep, mos):
            t0 = time.perf_counter()
            r = g.count(mo)
            dt = time.perf_counter() - t0
            t_plain += dt
            plain.append(r)
            plain_secs.append(round(dt, 4))
            pc = per_class.setdefault(r.class_name, {"motifs": 0, "plain_s": 0.0})
            pc["motifs"] += 1
            pc["plain_s"] += dt
            print(f"# plain motif {i} ({r.class_name}): {dt:.3f} s {json.dumps(_result_fields(r))}", flush=True)
    batch, t_batch = [], None
    if part in ("both", "batched"):
        t0 = time.perf_counter()
        batch, batch_ms = g.count_batch(mos)
        t_batch = time.perf_counter() - t0
        print(f"# batched: {t_batch:.3f} s", flush=True)
    same = None
    if part == "both":
        same = all(_result_fields(a) == _result_fields(b) for a, b in zip(plain, batch))
    # the same batch captured once as a HIP graph, then replayed
    t_cap, same_cap = None, None
    if not quick and part == "both":
        cb = g.capture_batch(mos)
        t0 = time.perf_counter()
        cap, cap_ms = cb.run()
        t_cap = time.perf_counter() - t0
        cb.free()
        same_cap = all((a.exact, a.sg, a.core_matches) == (b.exact, b.sg, b.core_matches) for a, b in zip(plain, cap))
        print(f"# captured batch replay: {t_cap:.3f} s, identical {same_cap}", flush=True)
    # the same per class (one count_batch per class)
    for c, pc in (per_class.items() if not quick and part == "both" else []):
        sub = [mo for mo, r in zip(mos, plain) if r.class_name == c]
        t0 = time.perf_counter()
        g.count_batch(sub)
        pc["batched_s"] = time.perf_counter() - t0
        pc["speedup"] = round(pc["plain_s"] / pc["batched_s"], 3)
        print(f"# class {c}: {pc}", flush=True)
    for pc in per_class.values():
        for k in ("plain_s", "batched_s"):
            if k in pc:
                pc[k] = round(pc[k], 4)
    excluded = {}
    for i in range(count):
        if i not in keep:
            excluded[cls_of[i]] = excluded.get(cls_of[i], 0) + 1
    res = {
        "graph": spec, "nodes": n, "directed_edges": m, "maxdeg": maxdeg, "relabel": relabel,
        "relabel_ms": round(relabel_ms, 1) if relabel_ms is not None else None, "part": part,
        "motifs": len(mos), "motifs_generated": count, "seed": seed, "excluded": excluded,
        "motif_ids": keep,
        "plain_s": round(t_plain, 3) if plain else None, "batched_s": round(t_batch, 3) if batch else None,
        "plain_motifs_per_s": round(len(mos) / t_plain, 4) if plain else None,
        "batched_motifs_per_s": round(len(mos) / t_batch, 4) if batch else None,
        "speedup": round(t_plain / t_batch, 3) if plain and batch else None, "identical": same,
        "captured_s": round(t_cap, 3) if t_cap else None,
        "captured_speedup": round(t_plain / t_cap, 3) if t_cap else None, "captured_identical": same_cap,
        "per_class": per_class,
    }
    print(json.dumps(res) if emit else "# " + json.dumps(res), flush=True)
    if out_path:
        rows = batch if batch else plain
        rec = {"summary": res, "counts": [str(r.exact) for r in rows]}
        if plain:
            rec["plain_results"] = [_result_fields(r) for r in plain]
            rec["plain_seconds_per_motif"] = plain_secs
        if batch:
            rec["batched_results"] = [_result_fields(r) for r in batch]
        json.dump(rec, open(out_path, "w"), indent=1)
    return res


def batch_merge(paths, out_path=None):
    """Combines a plain-only and a batched-only record of batch_compare (same graph,
    seed and motif set) into one comparison: times, speedup and per-motif equality of
    every result field."""
    recs = [json.load(open(p)) for p in paths]
    pl = next(r for r in recs if "plain_results" in r)
    bt = next(r for r in recs if "batched_results" in r)
    a, b = pl["summary"], bt["summary"]
    for k in ("graph", "seed", "motifs_generated", "motif_ids"):
        if a.get(k) != b.get(k):
            raise SystemExit(f"batch_merge: records differ in {k}: {a.get(k)} vs {b.get(k)}")
    same = pl["plain_results"] == bt["batched_results"]
    res = dict(a)
    res.update({"part": "merged", "batched_s": b["batched_s"], "batched_motifs_per_s": b["batched_motifs_per_s"],
                "speedup": round(a["plain_s"] / b["batched_s"], 3), "identical": same,
                "relabel_batched": b.get("relabel"), "runs": [os.path.basename(p) for p in paths]})
    print(json.dumps(res), flush=True)
    if out_path:
        json.dump({"summary": res, "counts": bt["counts"], "plain_results": pl["plain_results"],
                   "batched_results": bt["batched_results"]}, open(out_path, "w"), indent=1)
    return res


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = max(0, min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def serve_bench(specs, rate, nq, slots, classes=None, seed=5, out_path=None, check=True, emit=True):
    """Served query stream: graphs resident, Poisson arrivals at `rate` queries/s from
    `seed`, each query a uniform graph and the next motif of the random-motif generator
    (seed `seed`), submitted to the library's server at its arrival time. Latency is
    completion minus arrival on the server's clock. Every result is then recomputed
    alone with fsgc_count and compared."""
    import threading

    import numpy as np

    import fsgc
    fsgc.init(0)
    graphs = []
    for spec in specs:
        t0 = time.perf_counter()
        g = fsgc.Graph.generate(spec)
        n, m, maxdeg = g.info()
        graphs.append(g)
        print(f"# graph {spec}: n={n} m={m} maxdeg={maxdeg} ({time.perf_counter() - t0:.1f} s)", flush=True)
    rng = np.random.default_rng(seed)
    pool = fsgc.random_motifs(seed, 4 * nq)
    queries = []
    for mo in pool:
        if len(queries) == nq:
            break
        c = fsgc.motif_class(mo)
        if classes is not None and c not in classes:
            continue
        queries.append((int(rng.integers(len(graphs))), mo, c))
    gaps = rng.exponential(1.0 / rate, size=len(queries))
    srv = fsgc.Server(slots=slots)
    t0 = time.perf_counter()
    ids = [srv.add_graph(g) for g in graphs]
    print(f"# server: {slots} slots, graphs registered (index build {time.perf_counter() - t0:.1f} s)", flush=True)
    tickets = [None] * len(queries)

    def submitter():
        start = time.perf_counter()
        at = 0.0
        for i, (gi, mo, c) in enumerate(queries):
            at += gaps[i]
            delay = start + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(ids[gi], mo)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    results = []
    for i, t in enumerate(tickets):
        q = srv.wait(t)
        results.append(q)
        print(f"# query {i} ({queries[i][2]} on graph {queries[i][0]}): latency {q.latency_ms:.1f} ms, "
              f"batch {q.batch_size}, slot {q.slot}", flush=True)
    st = srv.stats()
    srv.close()
    lat = [q.latency_ms for q in results]
    first = min(q.submit_ms for q in results)
    last = max(q.done_ms for q in results)
    identical = None
    if check:
        identical = True
        for (gi, mo, c), q in zip(queries, results):
            a = graphs[gi].count(mo)
            b = q.result
            identical = identical and (a.exact, a.sg, a.aut, a.ref_compat, a.core_matches, a.flags) == \
                (b.exact, b.sg, b.aut, b.ref_compat, b.core_matches, b.flags)
        print(f"# alone == served for every query: {identical}", flush=True)
    by_class = {}
    for (gi, mo, c), q in zip(queries, results):
        by_class.setdefault(c, []).append(q.latency_ms)
    res = {
        "graphs": specs, "offered_rate_qps": rate, "queries": len(queries), "slots": slots, "seed": seed,
        "classes": sorted(by_class), "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
        "mean_ms": round(sum(lat) / len(lat), 3), "max_ms": round(max(lat), 3),
        "motifs_per_s": round(len(results) / ((last - first) / 1000.0), 3),
        "launches": st["launches"], "peak_concurrent": st["peak_concurrent"],
        "scratch_peak_bytes": st["scratch_peak"], "scratch_cap_bytes": st["scratch_cap"],
        "mean_batch": round(sum(q.batch_size for q in results) / len(results), 3),
        "per_class_p50_p99_ms": {c: [round(percentile(v, 50), 3), round(percentile(v, 99), 3)] for c, v in by_class.items()},
        "identical_alone_vs_served": identical,
    }
    print(json.dumps(res) if emit else "# " + json.dumps(res), flush=True)
    if out_path:
        json.dump(res, open(out_path, "w"), indent=1)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the small batched and served runs")
    ap.add_argument("--tune", action="store_true", help="autotune the launch shape on the workload first")
    ap.add_argument("--tune-sample", type=int, default=16)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle work for the CPU baseline")
    ap.add_argument("--alg-model", action="store_true",
                    help="recompute the pivot kernel's algorithmic bytes from unit counts on the host (oracle model) "
                         "and check them against the kernel's (c2 workload)")
    ap.add_argument("--gpu-check-budget", type=float, default=60.0,
                    help="seconds of GPU slice counts checked against the CPU baseline's slices")
    ap.add_argument("--batch", default=None, help="graph spec: compare plain vs batched over the random motifs")
    ap.add_argument("--batch-out", default=None)
    ap.add_argument("--batch-quick", action="store_true", help="--batch: skip the captured replay and per-class runs")
    ap.add_argument("--batch-part", default="both", choices=["both", "plain", "batched"],
                    help="--batch: run only one side (merge two such records with --batch-merge)")
    ap.add_argument("--batch-relabel", default="none", choices=["none", "asc", "desc"],
                    help="--batch: count on the graph renumbered by degree")
    ap.add_argument("--batch-merge", default=None, help="comma-separated plain-only and batched-only --batch-out files")
    ap.add_argument("--classes", default=None, help="comma-separated motif classes for --batch / --serve "
                    "(--batch also takes clique: the general cores with a 4-clique core)")
    ap.add_argument("--serve", default=None, help="comma-separated graph specs: served Poisson query stream")
    ap.add_argument("--rate", type=float, default=20.0, help="--serve: offered queries per second")
    ap.add_argument("--queries", type=int, default=64, help="--serve: number of queries")
    ap.add_argument("--slots", type=int, default=2, help="--serve: concurrent launches")
    ap.add_argument("--serve-out", default=None)
    ap.add_argument("--relabel", default="auto", choices=["auto", "none", "desc", "asc"],
                    help="count on the graph renumbered by degree (desc: the reference's relabelByDegree order; "
                         "auto: the workload's choice); the relabelling runs once at load, outside the timed "
                         "region, and is reported as relabel_ms")
    args = ap.parse_args()
    classes = set(args.classes.split(",")) if args.classes else None
    if args.batch_merge:
        batch_merge(args.batch_merge.split(","), out_path=args.batch_out)
        return
    if args.batch:
        batch_compare(args.batch, classes=classes, out_path=args.batch_out, quick=args.batch_quick,
                      part=args.batch_part, relabel=args.batch_relabel)
        return
    if args.serve:
        serve_bench(args.serve.split(","), args.rate, args.queries, args.slots, classes=classes,
                    out_path=args.serve_out)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the count shards by root vertices with one 16-byte reduction; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import fsgc
    W = WORKLOADS[args.workload]
    fsgc.init(0)
    t0 = time.perf_counter()
    ni, nl = fsgc.generate_csr(W["spec"])
    g = fsgc.Graph.from_csr(ni, nl)
    gen_s = time.perf_counter() - t0
    n, m, maxdeg = g.info()
    # the graph the steps count on: the loaded one, or its degree relabelling (counts
    # are label-invariant; the CPU slices below stay on the loaded ids)
    gc, relabel_ms = g, None
    if args.relabel == "auto":
        args.relabel = W.get("relabel", "none")
    mo, _ = fsgc.preprocess(*W["motif"])
    print(f"# graph {W['spec']}: n={n} m={m} maxdeg={maxdeg} (generated+uploaded in {gen_s:.1f}s)", flush=True)
    # The cold count: the reference's timed scope (fringeCount.cu:2317-2321: the
    # automorphism run, the degree relabel, the worklist / per-graph index build, the
    # kernels and the result copy; not the graph read and upload), from the resident
    # graph to the first result. The steps below reuse the relabelled graph, its index
    # and the automorphism sum.
    t_cold = time.perf_counter()
    if args.relabel != "none":
        gc, _, relabel_ms = g.relabel(args.relabel)
    first = gc.count(mo)
    cold_ms = (time.perf_counter() - t_cold) * 1000.0
    if relabel_ms is not None:
        print(f"# relabelled by degree ({args.relabel}) in {relabel_ms:.1f} ms", flush=True)
    print(f"# first call: index {first.index_ms:.1f} ms, kernel {first.kernel_ms:.1f} ms, class {first.class_name}, "
          f"exact {first.exact}; cold count {cold_ms:.1f} ms", flush=True)
    tuned = None
    if args.tune:
        # the full launch with the default shape, then the tuner on every
        # tune_sample-th work item; the timed steps below use the shape it keeps
        default_ms = gc.count(mo).kernel_ms
        t0 = time.perf_counter()
        reps = gc.autotune([mo], sample=args.tune_sample)
        tuned = {"reports": reps, "tune_s": round(time.perf_counter() - t0, 2), "default_kernel_ms": round(default_ms, 3)}
        print(f"# autotune ({tuned['tune_s']} s): {reps}; default shape {default_ms:.1f} ms", flush=True)
    for i in range(args.warmup):
        r = gc.count(mo)
        print(f"# warmup {i}: kernel {r.kernel_ms:.1f} ms", flush=True)
    kms, results = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = gc.count(mo)
        kms.append(r.kernel_ms)
        results.append(r)
        print(f"# step {i}: kernel {r.kernel_ms:.1f} ms wall {r.wall_ms:.1f} ms", flush=True)
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1000.0
    r = results[-1]
    assert all(x.exact == r.exact and x.sg == r.sg for x in results), "counts differ between steps"
    value = r.core_matches / (ms_per_step / 1000.0)
    kernel_ms = sum(kms) / len(kms)
    if tuned is not None:
        tuned["tuned_kernel_ms"] = round(kernel_ms, 3)
    achieved = r.alg_bytes / (kernel_ms / 1000.0) / 1e9 if r.alg_bytes else None
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
        "traffic": load_traffic(W["kernel"], args.workload),
        "kernel": W["kernel"],
        "kernel_ms": round(kernel_ms, 3),
        "alg_bytes_per_launch": r.alg_bytes,
    }

    if W.get("linear"):
        # the same count with every wedge evaluated with its own triple count (the
        # kernel's c3 instantiation): an independent evaluation of the headline number
        os.environ["FSGC_PIVOT_C3_PER_WEDGE"] = "1"
        try:
            pw = gc.count(mo)
        finally:
            os.environ.pop("FSGC_PIVOT_C3_PER_WEDGE", None)
        pw_ach = pw.alg_bytes / (pw.kernel_ms / 1000.0) / 1e9 if pw.alg_bytes else None
        roofline["launch"] = "k_wedge_pivot without c3 + k_wedge_lin_centres (per-centre triple-region terms)"
        roofline["per_wedge_c3"] = {
            "kernel_ms": round(pw.kernel_ms, 3), "alg_bytes_per_launch": pw.alg_bytes,
            "achieved": round(pw_ach, 2) if pw_ach else None,
            "frac": round(pw_ach / HBM_PEAK_GBS, 5) if pw_ach else None,
            "identical": (pw.exact, pw.sg, pw.ref_compat, pw.core_matches, pw.flags) ==
                         (r.exact, r.sg, r.ref_compat, r.core_matches, r.flags)}
        print(f"# per-wedge c3: kernel {pw.kernel_ms:.1f} ms, identical {roofline['per_wedge_c3']['identical']}",
              flush=True)

    cpu = None
    if not args.no_cpu:
        # The oracle (the reference's per-root schedule on the host cores) on a
        # deterministic sample: root vertices u with u % mod == rem for rem = mod-1,
        # mod-2, ... until about cpu_budget seconds of CPU work (on R-MAT these are
        # mostly low-degree roots: one root of a hub centre alone takes the oracle
        # minutes); every sub-slice is also
        # counted on the GPU with the kernel the steps time (default schedule, on the
        # loaded ids the slices are defined on) and compared bit-exact: class sum,
        # exact count, the reference's u64, core matches and flags.
        import oracle
        mod, budget = W["cpu_slice"][0], args.cpu_budget
        cores = physical_cores()
        done, matches, secs, parity, checked, gpu_check_s = [], 0, 0.0, True, 0, 0.0
        rem = mod - 1
        while rem >= 0:
            o = oracle.count(mo, ni, nl, threads=cores, slice_mod=mod, slice_rem=rem)
            ok = None
            # the GPU check of a slice runs the timed kernel on the whole graph with the
            # other roots filtered out; it stops after gpu_check_budget seconds
            if gpu_check_s < args.gpu_check_budget:
                t1 = time.perf_counter()
                gs = g.count(mo, fsgc.Tune(slice_mod=mod, slice_rem=rem)) if mod > 1 else r
                gpu_check_s += time.perf_counter() - t1
                ok = (gs.sg, gs.exact, gs.ref_compat, gs.core_matches, gs.flags & 0x1f) == \
                    (o["sg"], o["exact"], o["ref_compat"], o["core_matches"], o["flags"] & 0x1f)
                parity = parity and ok
                checked += 1
            done.append(rem)
            matches += o["core_matches"]
            secs += o["seconds"]
            print(f"# cpu slice u % {mod} == {rem}: {o['core_matches']} matches in {o['seconds']:.2f} s "
                  f"({o['threads']} threads), GPU identical: {ok}", flush=True)
            if secs >= budget or mod == 1:
                break
            rem -= 1
        cpu = {
            "value": round(matches / secs, 1) if secs > 0 else None,
            "unit": "core_matches/s",
            "cores": cores,
            "kind": "port",
            "sample": (f"roots u % {mod} in [{done[-1]}, {done[0]}] ({len(done)} slices, {matches} core matches, "
                       f"{secs:.2f} s; reference per-root schedule, OpenMP); GPU ({W['kernel']}"
                       f"{', per-wedge c3 on root slices' if W.get('linear') else ''}) results on "
                       f"{checked} of these slices identical: {parity}"),
        }
        if not parity:
            print("# CPU/GPU slice mismatch!", file=sys.stderr)
        if args.alg_model and W["kernel"] == "k_wedge_pivot":
            # the unit counts of the timed launch, from the CSR it ran on, and the
            # per-unit sizes of the kernel: they must give the kernel's own figure
            t1 = time.perf_counter()
            cni, cnl = gc.csr()
            lin = bool(W.get("linear"))
            # a launch without c3 keeps the hub vertices' 2-hop counters in LDS; the
            # per-centre pass reads 8 bytes per vertex and per position
            units = oracle.pivot_units(cni, cnl, ecap=2048, hub_slots=39552 if lin else 0, c3=not lin, threads=cores)
            model = oracle.pivot_bytes(units, 2 if maxdeg <= 65536 else 4, hub_launch=lin)
            if lin:
                units["centre_pass_vertices"], units["centre_pass_positions"] = n, m
                model += 8 * (n + m)
            roofline["alg_model"] = {"units": units, "bytes_per_unit": "20 paths_d + 4 paths_hub + 16 marks_d (20 in a "
                                     "launch without c3) + 4 marks_hub + 12 xvisits + 20 centres + (16+s) c3_x + (8+s) c3_search_bits + s list_elems + 16 walk_d + "
                                     "12 walk_hub + 8 centre_pass_vertices + 8 centre_pass_positions, s = list element "
                                     "bytes", "list_elem_bytes": 2 if maxdeg <= 65536 else 4, "bytes": model,
                                     "equals_kernel": int(r.alg_bytes) == model,
                                     "seconds": round(time.perf_counter() - t1, 2)}
            print(f"# alg-bytes model: {model} (kernel {int(r.alg_bytes)}), units {units}", flush=True)

    # the other two parts of the metric, on small graphs so that the run stays within
    # minutes (the full-size runs: bench.py --batch / --serve, results under profiles/)
    batched = served = None
    if not args.no_extras:
        b = batch_compare(EXTRA_BATCH, classes={"vertex", "edge", "wedge", "triangle"}, emit=False)
        batched = {k: b[k] for k in ("graph", "motifs", "excluded", "plain_motifs_per_s", "batched_motifs_per_s",
                                     "speedup", "captured_speedup", "identical", "captured_identical")}
        sv = serve_bench(EXTRA_SERVE, 20.0, 32, 2, classes={"vertex", "edge", "wedge", "triangle"}, emit=False)
        served = {k: sv[k] for k in ("graphs", "offered_rate_qps", "queries", "slots", "p50_ms", "p99_ms",
                                     "motifs_per_s", "identical_alone_vs_served")}

    # the full-size batched (C3: R-MAT s22) and served (C5: R-MAT s23 + s21) measurements,
    # as committed under profiles/ by their own runs (minutes each; not re-run here)
    records = {}
    for key, name, fields in (("c3_batched", "batch_c3_rmat22_54.json",
                               ("graph", "motifs", "excluded", "plain_s", "batched_s", "speedup", "identical")),
                              ("c5_served", "serve_c5_graphs_all_fast_classes_0.03qps.json",
                               ("graphs", "classes", "offered_rate_qps", "queries", "slots", "p50_ms", "p99_ms",
                                "motifs_per_s", "identical_alone_vs_served"))):
        path = os.path.join(ROOT, "profiles", name)
        if os.path.exists(path):
            rec = json.load(open(path))
            rec = rec.get("summary", rec)
            records[key] = {f: rec.get(f) for f in fields}
            records[key]["source"] = f"profiles/{name} (committed record, not re-run by this invocation)"

    out = {
        "metric": "core_matches_per_s",
        "value": round(value, 1),
        "unit": "core_matches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u128",
        "data": "synthetic",
        "config": {
            "workload": W["name"],
            "graph": W["spec"],
            "nodes": n,
            "directed_edges": m,
            "core_matches": r.core_matches,
            "exact_count": str(r.exact),
            "ref_compat_u64": r.ref_compat,
            "flags": r.flags,
            "index_build_ms": round(first.index_ms, 3),
            "cold_ms": round(cold_ms, 3),
            "relabel": args.relabel,
            "relabel_ms": round(relabel_ms, 3) if relabel_ms is not None else None,
            "autotuned": tuned,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
        "full_size_records": records or None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
