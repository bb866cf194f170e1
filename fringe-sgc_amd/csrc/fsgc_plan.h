// Motif planning (host) and the device driver interface.
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fsgc_host.h"
#include "fsgc_job.h"

namespace fsgc {

int classify(const MatchingOrder& mo);
int validate_mo(const MatchingOrder& mo, std::string* err);
int make_job(const MatchingOrder& mo, JobDesc* j, std::string* err);
int general_quirk_factor(const MatchingOrder& mo);
int triangle_rotations(const int f[8], int fr[6][8]);
int wedge_rotations(const int f[8], int fr[6][8]);
HostGraph motif_graph(const MatchingOrder& mo);
// 4-clique cores (general cores whose four vertices are pairwise adjacent)
bool clique_core(const JobDesc& j);
int job_fringes(const JobDesc& j);
// whether the clique schedule counts this job exactly on a graph of maximum degree maxdeg
bool clique_fits(const JobDesc& j, int32_t maxdeg);
int build_clique_tab(const JobDesc& j, CliqueTab* tab, std::vector<CliqueTerm>* terms, std::string* err);

// Launch parameters of the persistent counting kernel.
struct Tune {
  int block = 0;          // threads per workgroup for kernels with several shapes (0: the kernel's default)
  int grid = 0;           // workgroups (0: 8 per CU)
  int chunk = 256;        // roots (CSR positions or vertices) per work item
  bool chunk_set = false; // chunk given by the caller or the tuner (else the schedule's default: 256,
                          // 64 for rank launches that count triple intersections)
  int hub_degree = 4096;  // roots of degree above this are split into single-root items
  int lds_cap = 0;        // > 0: cap on the positions staged in LDS per window / chunk / mark array (0: the kernel's maximum)
  int sample = 1;         // > 1: run only every sample-th work item (timing runs; partial counts)
};

// keeps every t.sample-th element of a work-item list (all of them when sample <= 1)
template <class T>
void sample_items(std::vector<T>* items, int sample)
{
  if (sample <= 1) return;
  size_t k = 0;
  for (size_t i = 0; i < items->size(); i += (size_t)sample) (*items)[k++] = (*items)[i];
  items->resize(k);
}

// A graph resident in device memory (owned by the library).
struct DeviceGraph {
  int32_t n = 0;
  int64_t m = 0;
  int32_t maxdeg = 0;
  int32_t* idx = nullptr;   // n + 1
  int32_t* adj = nullptr;   // m
  std::vector<int32_t> h_idx;  // host copy of nindex (used to cut work items)
  // derived once per graph: supports and pivot costs, then (pivot and rank
  // schedules) the common-neighbour lists
  bool support_ready = false;
  bool derived_ready = false;   // lists built too
  int32_t* tsup = nullptr;      // m: |N(u) & N(x)| per directed edge
  int32_t* adjdeg = nullptr;    // m: degree of adj[e] (read in CSR order instead of through idx)
  int32_t* adjhub = nullptr;    // m: LDS slot of adj[e] in the pivot kernel when it is a hub, else ~adj[e]
  int32_t* rev = nullptr;       // m: CSR index of the reverse edge (x, u) of (u, x) (built with the lists)
  long long* tl_off = nullptr;  // m + 1: offsets of the common-neighbour lists
  int32_t* tl = nullptr;        // positions in N(u) of the common neighbours of (u, x)
  uint16_t* tl16 = nullptr;     // the same lists in 16 bits (built when maxdeg <= 65536; pivot kernel)
  int32_t* c3tl = nullptr;      // parallel to tl: entry of c in T(a, b), a < b < c by rank: |N(a) & N(b) & N(c)|
                                // (built for the clique schedule)
  long long tl_total = 0;
  long long* piv_cost = nullptr;         // n: sum of neighbour degrees
  std::vector<long long> h_piv_cost;
  // pivot-schedule work items on the device, per (chunk << 32 | sample): pointer, count
  std::map<long long, std::pair<void*, int>> piv_items;
  // rank-schedule work items, the same way
  std::map<long long, std::pair<void*, int>> tri_items;
  size_t derived_bytes = 0;              // device bytes of the derived arrays
};

struct DevScratch;



struct JobResult {
  u128 sum = 0;
  uint64_t matches = 0;
  uint32_t flags = 0;
  uint64_t bytes = 0;     // algorithmic bytes of the launch (pivot schedule)
  double index_ms = 0;    // per-graph index build time spent in this call
};

// One kernel launch prepared on the host and in device memory (work items, job
// descriptors and scratch resident). Running it only zeroes its accumulators (a fill
// kernel), launches the kernel and copies the accumulators to pinned host memory, so
// a sequence of runs can be captured into a HIP graph and replayed. Scratch the
// kernel dirties is left clean by the kernel itself (the pivot D arrays, the rank
// marks), so replays need no other reset.
struct LaunchPlan {
  int nj = 0;
  Accum* d_acc = nullptr;    // nj accumulators, then the work-queue head
  size_t acc_bytes = 0;      // bytes zeroed per run (multiple of 8)
  Accum* h_acc = nullptr;    // pinned host copy of the accumulators
  double index_ms = 0;       // per-graph index build spent while planning
  int nitems = 0;            // work items of the launch
  // launches the kernel over items [first, first + count) with work-queue head `head`
  std::function<void(void* stream, int first, int count, int* head)> kernel;
  // optional: launched once after all items (per-centre terms of the pivot schedule)
  std::function<void(void* stream)> post;
  // the accumulators nj .. 2 nj - 1 hold negative parts, subtracted from the sums
  bool neg = false;
  int* head = nullptr;       // work-queue heads (32 slots), zeroed with the accumulators
  std::shared_ptr<DevScratch> scratch;
  LaunchPlan() = default;
  LaunchPlan(const LaunchPlan&) = delete;
  LaunchPlan& operator=(const LaunchPlan&) = delete;
  LaunchPlan(LaunchPlan&& o) noexcept { *this = std::move(o); }
  LaunchPlan& operator=(LaunchPlan&& o) noexcept;
  ~LaunchPlan();
};
// allocates the accumulators (device and pinned host) of a plan for nj jobs, and with
// neg a negative part per job
int plan_accumulators(LaunchPlan* lp, int nj, void* stream, std::string* err, bool neg = false);
// zero + kernel (+ timing events when given) + copy of the results; no allocation, no sync
int launch_enqueue(const LaunchPlan& lp, void* stream, void* ev_begin, void* ev_end, std::string* err);
// results of a finished run
void launch_collect(const LaunchPlan& lp, std::vector<JobResult>* out);
// enqueue, wait, collect. With FSGC_SEGMENT_ITEMS=k in the environment the items run
// in launches of at most k items each (at most 32), one progress line per launch
// with FSGC_LOG; the result is the same.
int launch_run(const LaunchPlan& lp, void* stream, std::vector<JobResult>* out, double* kernel_ms, std::string* err);

// pivot.hip
int device_derived_build(DeviceGraph* g, void* stream, double* ms, std::string* err, bool lists);
void device_derived_free(DeviceGraph* g);
constexpr int kMaxPivotJobs = 8;
size_t pivot_scratch_bytes(const DeviceGraph& g, const Tune& t);
int plan_wedge_pivot(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream, LaunchPlan* lp,
                     std::string* err);
int device_run_wedge_pivot(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream,
                           std::vector<JobResult>* out, double* kernel_ms, std::string* err);

// tri.hip
constexpr int kMaxTriJobs = 8;
bool triangle_schedule_ok(const JobDesc& j);
bool edgepass_schedule_ok(const JobDesc& j);
size_t triangle_scratch_bytes(const DeviceGraph& g, const Tune& t);
size_t edgepass_scratch_bytes(const DeviceGraph& g, const Tune& t);
int plan_edgepass(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream, LaunchPlan* lp,
                  std::string* err);
int plan_triangles(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream, LaunchPlan* lp,
                   std::string* err);
int device_run_edgepass(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream,
                        std::vector<JobResult>* out, double* kernel_ms, std::string* err);
int device_run_triangles(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream,
                         std::vector<JobResult>* out, double* kernel_ms, std::string* err);
// the triple counts c3tl of every triangle, once per graph (after the lists)
int device_c3tl_build(DeviceGraph* g, void* stream, double* ms, std::string* err);

// clique.hip
constexpr int kMaxCliqueJobs = 8;
size_t clique_scratch_bytes(const DeviceGraph& g, const Tune& t);
int plan_cliques(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream, LaunchPlan* lp,
                 std::string* err);
// One edge-pass job (closed-form edge core or stars) with all its device scratch in a
// fixed array (no allocation); the graph's supports must be built. kernel_ms: host
// time of the device part.
int edgepass_count_noalloc(const DeviceGraph* g, JobDesc job, JobResult* out, double* kernel_ms, std::string* err);

// kernels.hip
int device_init(int device, std::string* err);
int device_graph_upload(const HostGraph& g, DeviceGraph* dg, std::string* err);
int device_graph_upload_async(const HostGraph& g, DeviceGraph* dg, DevScratch* sc, std::string* err);
void device_graph_free(DeviceGraph* dg);
// relabel.hip: a new resident graph renumbered by degree (order 0: highest first, ties
// by the larger id first, as the reference's relabelByDegree; 1: lowest first, ties by
// the smaller id first); h_map (optional) receives old -> new ids, ms the device time
int device_graph_relabel(const DeviceGraph& src, int order, DeviceGraph* dst, std::vector<int32_t>* h_map,
                         double* ms, std::string* err);
// Runs the given jobs (each bound to a graph) in one persistent launch on `stream`
// and returns one result per job. kernel_ms receives the device time of the launch.
size_t perroot_scratch_bytes(const DeviceGraph& g, const Tune& t);
int plan_jobs(std::vector<JobDesc>& jobs, const std::vector<const DeviceGraph*>& graphs, const Tune& t, void* stream,
              LaunchPlan* lp, std::string* err);
int device_run_jobs(std::vector<JobDesc>& jobs, const std::vector<const DeviceGraph*>& graphs, const Tune& t,
                    void* stream, std::vector<JobResult>* out, double* kernel_ms, std::string* err);

}  // namespace fsgc
