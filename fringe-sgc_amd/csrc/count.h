// Host side of one counting call: motif preparation, kernel scheduling per class,
// automorphism sums and the final division (shared by the C ABI, the autotuner and
// the serving loop).
#pragma once
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fsgc.h"
#include "fsgc_plan.h"

namespace fsgc {

// Which kernel counts a motif on the data graph.
enum Schedule { kPersistent = 0, kPivot = 1, kRank = 2, kEdgePass = 3, kClique = 4, kNumSchedules = 5 };
const char* schedule_name(int s);

// Automorphism sum of a motif (run on the motif graph, cached by the live fields).
struct AutEntry {
  u128 sum;
  uint32_t flags;
  int32_t hn;
  int64_t hm;
};

struct Prepared {
  MatchingOrder mo;
  JobDesc job;
  int qf = 1;
  bool trivial = false;  // vertex count / edge count: no kernel
  bool need_aut = false;
  bool aut_cached = false;
  AutEntry aut{};
  std::string key;
  HostGraph mgraph;
};

}  // namespace fsgc

// A graph resident on the GPU, with the launch shapes the autotuner picked for it.
struct fsgc_graph {
  fsgc::DeviceGraph dg;
  std::mutex tune_mu;
  bool tuned[fsgc::kNumSchedules] = {false, false, false, false, false};
  fsgc::Tune best[fsgc::kNumSchedules];
};

namespace fsgc {

int ensure_init(std::string* err);  // selects the library's GPU on first use
int current_device();
void set_error(const std::string& msg);  // text returned by fsgc_last_error on this thread
bool log_on();
int prepare(const uint8_t* mob, const fsgc_tune* tune, Prepared* p, std::string* err);
// The pivot kernel keeps 2 * c(w, y) + adjacency of a pivot w and a vertex y in the low
// 24 bits of a counter (the top 8 hold an epoch); c(w, y) <= maxdeg, so the pivot
// schedule takes graphs of maximum degree up to 2^23 - 1 only.
constexpr int32_t kPivotMaxDegree = (1 << 23) - 1;
// The kernel that counts a motif on a graph of maximum degree `maxdeg`.
Schedule schedule_of(const Prepared& p, bool fast, int32_t maxdeg);
// Launch shape for one schedule: the caller's non-zero fields, else the graph's tuned
// values, else the defaults.
Tune tune_for(fsgc_graph* g, const fsgc_tune* t, int sched);
// Device scratch (bytes) one launch of `sched` takes on g with the given shape.
size_t scratch_for(const fsgc_graph* g, int sched, const Tune& t);

// The launches a set of prepared motifs needs: the per-root launch, then groups of up
// to 8 motifs per fast schedule that share one enumeration.
struct Group {
  int sched;
  std::vector<size_t> idx;  // indices into the prepared motifs
};
std::vector<Group> make_groups(const std::vector<Prepared>& ps, bool fast, int32_t maxdeg);
int plan_group(fsgc_graph* g, const std::vector<Prepared>& ps, const Group& gr, const fsgc_tune* tune, int sample,
               void* stream, LaunchPlan* lp, std::string* err);

// Runs the data-graph jobs of all prepared motifs (and the motif-graph jobs of
// motifs whose automorphism sum is not cached) on `stream`.
// `sample` > 1 runs only every sample-th work item (timing runs of the autotuner).
int run_prepared(fsgc_graph* g, std::vector<Prepared>& ps, const fsgc_tune* tune, std::vector<JobResult>* res,
                 double* kernel_ms, void* stream, std::string* err, int sample = 1);
void finish(const Prepared& p, const JobResult& r, int32_t n, int64_t m, fsgc_result* out);

}  // namespace fsgc
