/*
 * fsgc.h — C ABI of the MI355X Fringe-SGC counting library (libfsgc.so).
 *
 * The library counts non-induced occurrences of a motif (given as the reference's
 * 139-byte matching order, "motif.mo") in an undirected graph resident in GPU memory
 * (given as the reference's ECLgraph CSR). Plain pointers and sizes only; every
 * function returns 0 on success or a negative errno-style code (see fsgc_strerror);
 * nothing here calls exit().
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   fsgc_graph_load_egr  readECLgraph + host-to-device copy   src/ECLgraph.h:58-85, src/fringeCount.cu:2285-2297
 *   fsgc_graph_from_csr  the same from caller memory          src/fringeCount.cu:2292-2297
 *   fsgc_graph_free      freeECLgraph + cudaFree              src/ECLgraph.h:88-96, src/fringeCount.cu:2325-2329
 *   fsgc_graph_relabel   relabelByDegree (into a new graph)    src/fringeCount.cu:1221-1288, 1065-1164
 *   fsgc_count           occurrences(mo, d_g, wl, ...)         src/fringeCount.cu:2199-2262
 *   fsgc_count_batch     (no counterpart: the reference counts one motif per process)
 *   fsgc_batch_*         (no counterpart: the same, captured once as a HIP graph and replayed)
 *   fsgc_preprocess      fringePreprocess n s d s d ...        src/fringePreprocess.cpp:287-365
 *   fsgc_preprocess_pattern  fringePreprocess <pattern#>       src/fringePreprocess.cpp:296-298, src/patterns.h:57-354
 *   fsgc_print_matching_order  printMatchingOrder              src/fringes.h:61-98
 *   fsgc_autotune        (no counterpart: the reference's launch shape is fixed,
 *                         512 threads per block, fringeCount.cu:82-84, 2276-2282)
 *   fsgc_server_*        (no counterpart: the reference is one motif per process,
 *                         driven by reproduce.py:69-73 through subprocess)
 */
#ifndef FSGC_H
#define FSGC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSGC_MO_BYTES 139

/* result flags */
#define FSGC_OVF64 (1u << 0)           /* exact count >= 2^64: the u64 the reference prints wrapped */
#define FSGC_OVF128 (1u << 1)          /* an exact intermediate reached 2^128: count invalid */
#define FSGC_REF_NCK_INEXACT (1u << 2) /* the reference's u64 nCk (fringeCount.cu:195) would have wrapped */
#define FSGC_AUT_REMAINDER (1u << 3)   /* class sum not divisible by the automorphism sum */
#define FSGC_REF_AUT_TRUNC (1u << 4)   /* the reference's int automorphism count (fringeCount.cu:2213) truncates */

/* motif classes (core size 1, 2, 3 open, 3 closed, 4-6) */
#define FSGC_CLASS_VERTEX 0
#define FSGC_CLASS_EDGE 1
#define FSGC_CLASS_WEDGE 2
#define FSGC_CLASS_TRIANGLE 3
#define FSGC_CLASS_GENERAL 4

typedef struct fsgc_graph fsgc_graph;

typedef struct fsgc_result {
  uint64_t exact_lo, exact_hi; /* exact count = floor(S_G / aut), unsigned 128-bit */
  uint64_t ref_compat;         /* the u64 the reference computes: (S_G mod 2^64) / aut */
  uint64_t sg_lo, sg_hi;       /* S_G: the class sum over the data graph, before / aut */
  uint64_t aut;                /* automorphism sum (the reference prints it as "automorphisms") */
  uint64_t core_matches;       /* core embeddings for which a fringe formula was evaluated */
  uint32_t flags;              /* FSGC_* flags */
  int32_t motif_class;         /* FSGC_CLASS_* */
  double kernel_ms;            /* device time of the counting launch on the data graph */
  double wall_ms;              /* host wall time of the whole call */
  double alg_bytes;            /* algorithmic global-memory bytes of the dominant launch's schedule (0 if not counted) */
  double index_ms;             /* device time spent building the graph's per-edge index in this call (0 if cached) */
} fsgc_result;

/* Launch shape. A zero field takes the value fsgc_autotune picked for the graph and
   the motif's schedule, or the default when the graph was not tuned. */
typedef struct fsgc_tune {
  int32_t block;       /* threads per workgroup where a kernel has several shapes (pivot: 256, 512, 1024); 0 = default */
  int32_t grid;        /* persistent workgroups; 0 = tuned or default (4-8 per CU) */
  int32_t chunk;       /* roots per work item; 0 = tuned or default */
  int32_t hub_degree;  /* roots above this degree get items of their own; 0 = tuned or default */
  int32_t lds_cap;     /* > 0: caps the list positions staged in LDS: the pivot kernel's c3 windows (at most
                          4096 positions per wave for 256/512 threads, 2048 for 1024) and its shared
                          phase-A and walk chunks (2048), and the degree up to which the rank kernel keeps
                          a vertex's marks in LDS (8192; larger ones use global memory); 0 = those maxima */
  int32_t slice_mod;   /* > 1: count only roots u with u % slice_mod == slice_rem */
  int32_t slice_rem;
  int32_t schedule;    /* 0: fastest schedule per class; 1: per-root schedule (the reference's) for every class */
} fsgc_tune;

/* Selects the GPU (the reference hard-codes device 0, fringeCount.cu:2276). */
int fsgc_init(int device);

/* Graph residency. The graph must be undirected with sorted, duplicate-free lists and
   no self loops (fringeCount.cu:50-56, which states this but never checks it).
   fsgc_graph_from_csr, fsgc_graph_load_egr and fsgc_graph_generate check it on the
   host before anything reaches the GPU and return -22 (message in fsgc_last_error)
   for a malformed graph; fsgc_graph_validate re-checks a resident graph. */
int fsgc_graph_load_egr(const char* path, fsgc_graph** out);
int fsgc_graph_from_csr(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist, fsgc_graph** out);
int fsgc_graph_validate(const fsgc_graph* g);
int fsgc_graph_info(const fsgc_graph* g, int32_t* nodes, int32_t* edges, int32_t* maxdeg);
void fsgc_graph_free(fsgc_graph* g);

/* Degree relabelling (relabelByDegree, fringeCount.cu:1221-1288, with its kernels
   init_sort/create_map/create_edge_list/create_nindex/create_nlist, 1065-1164). The
   reference renumbers the caller's device graph in place before a triangle-core count
   (fringeCount.cu:1348), and skips that when the maximum degree is at most 16. Here the
   caller's graph is left unchanged and *out receives a new resident graph in which
   vertex u of g is vertex map[u]:
     FSGC_RELABEL_DEGREE_DESC  degree highest first, ties by the larger id first (the
                               reference's descending sort of (degree << idbits) + id)
     FSGC_RELABEL_DEGREE_ASC   degree lowest first, ties by the smaller id first
   Lists stay sorted. Counts on *out equal counts on g. map (n entries) and relabel_ms
   (device time) may be NULL. */
#define FSGC_RELABEL_DEGREE_DESC 0
#define FSGC_RELABEL_DEGREE_ASC 1
int fsgc_graph_relabel(const fsgc_graph* g, int32_t order, fsgc_graph** out, int32_t* map, double* relabel_ms);
/* Copies a resident graph's CSR to the host (nindex: nodes + 1 entries, nlist: edges).
   No reference counterpart: the reference never reads its device graph back. */
int fsgc_graph_copy_csr(const fsgc_graph* g, int32_t* nindex, int32_t* nlist);

/* Counts one motif (one persistent launch on the data graph, plus one on the motif
   graph for the automorphism sum, as the reference does). tune may be NULL. */
int fsgc_count(fsgc_graph* g, const uint8_t mo[FSGC_MO_BYTES], const fsgc_tune* tune, fsgc_result* out);

/* fsgc_count for callers that cannot allocate (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc). No reference counterpart. It calls
   no hipMalloc, hipFree or other allocating routine: its device scratch (work-queue
   head, accumulator, job descriptor, work items) is a fixed 4 MiB array of the library,
   one copy per GPU, held until the process ends. Supported: vertex motifs and edge
   motifs of the closed-form kind (the ones fsgc_plan_schedule maps to the edge pass,
   or to the host); others return -95. The graph's per-edge supports must already be
   built (any fsgc_count of an edge, wedge or triangle motif on g builds them), else
   -11. Graphs of more than 130,960 * 8192 = 1,072,824,320 directed edges (the work items
   of 8192 edges each then fill the array) return -27.
   Calls on one GPU share the array; the library serialises them, so concurrent calls
   wait for each other. out equals fsgc_count's result; kernel_ms is host-timed. */
int fsgc_count_noalloc(fsgc_graph* g, const uint8_t mo[FSGC_MO_BYTES], fsgc_result* out);

/* The kernel schedule fsgc_count uses for a motif on a graph of maximum degree maxdeg:
   0 per-root, 1 pivot, 2 rank, 3 edge pass, 4 clique, -1 none (vertex or edge count,
   computed on the host). Wedge cores take the per-root schedule when maxdeg >= 2^23,
   where the pivot kernel's 24-bit 2-hop counters could overflow. 4-clique general cores
   take the clique schedule when their fringe count stays below 2^128 for every
   clique of the graph (maxdeg^fringes < 2^128, maxdeg < 2^21, at most 7 fringes), else
   the per-root schedule. Host only: no GPU is touched.
   No reference counterpart (the reference's dispatch sits inside occurrences,
   fringeCount.cu:2199-2262, and has one kernel per class). motif_class may be NULL. */
int fsgc_plan_schedule(const uint8_t mo[FSGC_MO_BYTES], int32_t maxdeg, int32_t* schedule, int32_t* motif_class);

/* Counts k motifs in one persistent launch over one graph; out[i] is identical to
   fsgc_count of mos + i*139 alone. batch_ms receives the device time of the launch. */
int fsgc_count_batch(fsgc_graph* g, const uint8_t* mos, int k, const fsgc_tune* tune, fsgc_result* out,
                     double* batch_ms);

/* A batch captured as one HIP graph: fsgc_batch_create plans every launch the k motifs
   need on g (work items, descriptors and scratch resident; the graph's edge index and
   the automorphism sums computed by one eager run) and captures the zeroing, the
   kernels and the result copies into a HIP graph; fsgc_batch_run replays it with one
   launch and one synchronisation. out[i] equals fsgc_count of motif i. g must outlive
   the batch; tune may be NULL (slices and the per-root schedule are refused). */
typedef struct fsgc_batch fsgc_batch;
int fsgc_batch_create(fsgc_graph* g, const uint8_t* mos, int k, const fsgc_tune* tune, fsgc_batch** out);
int fsgc_batch_run(fsgc_batch* b, fsgc_result* out, double* graph_ms);
void fsgc_batch_free(fsgc_batch* b);

/* Autotuner. For every schedule the given motifs use (per-root kernel, pivot, rank,
   edge pass, clique), times candidate launch shapes (workgroups, roots per work item, hub
   threshold) on every sample-th work item of the graph (sample <= 1: 16) and keeps the
   fastest for the graph; later counts on the graph use it for fields left at zero.
   Counts never depend on the shape. reports[0..*nreports) describe each schedule
   tuned (up to FSGC_NUM_SCHEDULES). */
#define FSGC_NUM_SCHEDULES 5
typedef struct fsgc_tune_report {
  int32_t schedule;     /* 0 per-root, 1 pivot, 2 rank, 3 edge pass, 4 clique */
  int32_t candidates;   /* shapes timed */
  fsgc_tune best;       /* the shape kept */
  double best_ms;       /* device time of the sample with it */
  double default_ms;    /* device time of the sample with the default shape */
} fsgc_tune_report;
int fsgc_autotune(fsgc_graph* g, const uint8_t* mos, int k, int32_t sample, fsgc_tune_report* reports,
                  int32_t* nreports);
/* forget the tuned shapes of a graph */
int fsgc_graph_untune(fsgc_graph* g);

/* Serving loop. A server keeps registered graphs resident (with their per-edge
   index), takes count queries at any time from any thread, and runs them on `slots`
   concurrent HIP streams. Whenever a slot is free it folds the waiting queries that
   share a graph and a schedule (up to max_batch, oldest first) into one launch,
   admitted only while the device scratch of the launches in flight stays within
   scratch_cap; queries arriving while others run join the next launch on a free
   slot. Each result is identical to fsgc_count of the same motif alone. */
typedef struct fsgc_server fsgc_server;
typedef struct fsgc_server_config {
  int32_t slots;         /* concurrent launches (streams); 0 = 2 */
  int32_t max_batch;     /* motifs folded into one launch, 1..8; 0 = 8 */
  uint64_t scratch_cap;  /* device bytes the launches in flight may hold; 0 = half the free memory */
  int32_t grid;          /* workgroups per launch; 0 = the machine's share of one slot */
  int32_t reserved;
} fsgc_server_config;
typedef struct fsgc_query {
  fsgc_result result;
  int32_t status;        /* 0, or the negative error code of the query */
  int32_t graph;         /* graph id the query ran on */
  int32_t batch_size;    /* motifs in the launch that counted it */
  int32_t slot;          /* stream slot */
  double submit_ms;      /* server clock (fsgc_server_clock_ms) at submission */
  double start_ms;       /* ... when its launch was admitted */
  double done_ms;        /* ... when its result was ready */
  uint64_t scratch_bytes;/* device scratch of its launch */
} fsgc_query;
typedef struct fsgc_server_stats {
  uint64_t submitted, completed, launches, rejected;
  uint64_t scratch_cap, scratch_peak; /* bytes */
  int32_t slots, peak_concurrent;     /* launches in flight at once (max seen) */
} fsgc_server_stats;
int fsgc_server_create(const fsgc_server_config* cfg, fsgc_server** out);
/* registers a graph (builds its per-edge supports now, the common-neighbour lists
   on the first query that needs them); g must outlive the server */
int fsgc_server_add_graph(fsgc_server* s, fsgc_graph* g, int32_t* id);
int fsgc_server_submit(fsgc_server* s, int32_t graph, const uint8_t mo[FSGC_MO_BYTES], uint64_t* ticket);
/* blocks until the query is done, fills out and forgets the ticket */
int fsgc_server_wait(fsgc_server* s, uint64_t ticket, fsgc_query* out);
/* 1 and out filled (ticket forgotten) if done, 0 if still pending */
int fsgc_server_poll(fsgc_server* s, uint64_t ticket, fsgc_query* out);
int fsgc_server_stats_get(fsgc_server* s, fsgc_server_stats* out);
double fsgc_server_clock_ms(const fsgc_server* s);
/* finishes the queries already submitted, then stops */
void fsgc_server_destroy(fsgc_server* s);

/* Motif compiler: pattern with `nodes` vertices and `nedges` edges (pairs in `edges`)
   -> 139-byte matching order (unused slots zero). `log` receives the reference's
   report text (may be NULL), `dot` the motif.dot text (may be NULL). */
int fsgc_preprocess(int nodes, int nedges, const int* edges, uint8_t mo[FSGC_MO_BYTES], char* log, size_t logcap,
                    char* dot, size_t dotcap);
int fsgc_preprocess_pattern(int number, uint8_t mo[FSGC_MO_BYTES], char* log, size_t logcap, char* dot,
                            size_t dotcap);
int fsgc_print_matching_order(const uint8_t mo[FSGC_MO_BYTES], char* out, size_t cap);

/* Synthetic inputs (the reference downloads its graphs, setup.sh:3-7; these replace
   that offline). spec is one of
     "er:<n>:<m_undirected>:<seed>"
     "rmat:<scale>:<edge_factor>:<a>:<b>:<c>:<seed>"
     "chunglu:<n>:<gamma>:<mean_degree>:<max_degree>:<seed>"
     "egr:<path>"
   fsgc_generate returns host arrays to be released with fsgc_free_host. */
int fsgc_generate(const char* spec, int32_t* nodes, int32_t* edges, int32_t** nindex, int32_t** nlist);
int fsgc_graph_generate(const char* spec, fsgc_graph** out);
void fsgc_free_host(void* p);
/* count random motifs (core 2..5 vertices, density U[0.4,1), 0..6 degree-1 fringes)
   from `seed`, preprocessed into count * 139 bytes of matching orders. */
int fsgc_random_motifs(uint64_t seed, int count, uint8_t* mos);
/* the same motifs before preprocessing: nodes[i], nedges[i] and up to 64 edge pairs
   at edges + i * 128 */
int fsgc_random_patterns(uint64_t seed, int count, int* nodes, int* nedges, int* edges);

/* Text of the last error on this thread, and of an error code. */
const char* fsgc_last_error(void);
const char* fsgc_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* FSGC_H */
