// A batch of motifs on one graph as one captured HIP graph: every launch the batch
// needs is planned once (work items, job descriptors and scratch resident in device
// memory), the zeroing of the accumulators, the kernels and the copies of the results
// are captured from a stream into a graph, and each run replays that graph with one
// hipGraphLaunch and one synchronisation.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "count.h"
#include "devscratch.h"

struct fsgc_batch {
  fsgc_graph* g = nullptr;
  std::vector<fsgc::Prepared> ps;
  std::vector<fsgc::Group> groups;
  std::vector<fsgc::LaunchPlan> plans;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~fsgc_batch()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    plans.clear();  // frees the plans' scratch on the stream
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

namespace {

int fail(int rc, const std::string& msg)
{
  fsgc::set_error(msg);
  return rc;
}

#define HIP_OR_FAIL(expr)                                                           \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return fail(-5, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" {

int fsgc_batch_create(fsgc_graph* g, const uint8_t* mos, int k, const fsgc_tune* tune, fsgc_batch** out)
{
  if (!g || !out || k < 0 || (k > 0 && !mos)) return fail(-22, "null argument");
  std::string err;
  int rc = fsgc::ensure_init(&err);
  if (rc) return fail(rc, err);
  if (tune && (tune->slice_mod > 1 || tune->schedule == 1))
    return fail(-22, "a captured batch runs the full graph on the fast schedules");
  std::unique_ptr<fsgc_batch> b(new fsgc_batch());
  b->g = g;
  b->ps.resize(k);
  for (int i = 0; i < k; i++) {
    rc = fsgc::prepare(mos + (size_t)i * FSGC_MO_BYTES, tune, &b->ps[i], &err);
    if (rc) return fail(rc, err);
  }
  HIP_OR_FAIL(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  // one eager run: builds the graph's edge index and caches every automorphism sum,
  // so the captured work is the data-graph launches alone
  std::vector<fsgc::JobResult> res;
  double ms = 0;
  rc = fsgc::run_prepared(g, b->ps, tune, &res, &ms, b->stream, &err);
  if (rc) return fail(rc, err);
  b->groups = fsgc::make_groups(b->ps, true, g->dg.maxdeg);
  for (const fsgc::Group& gr : b->groups) {
    fsgc::LaunchPlan lp;
    rc = fsgc::plan_group(g, b->ps, gr, tune, 1, b->stream, &lp, &err);
    if (rc) return fail(rc, err);
    b->plans.push_back(std::move(lp));
  }
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  HIP_OR_FAIL(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (const fsgc::LaunchPlan& lp : b->plans) {
    rc = fsgc::launch_enqueue(lp, b->stream, nullptr, nullptr, &err);
    if (rc) {
      hipGraph_t dropped = nullptr;
      (void)hipStreamEndCapture(b->stream, &dropped);
      if (dropped) (void)hipGraphDestroy(dropped);
      return fail(rc, err);
    }
  }
  HIP_OR_FAIL(hipStreamEndCapture(b->stream, &b->graph));
  HIP_OR_FAIL(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  *out = b.release();
  return 0;
}

int fsgc_batch_run(fsgc_batch* b, fsgc_result* out, double* graph_ms)
{
  if (!b || (!out && !b->ps.empty())) return fail(-22, "null argument");
  const int k = (int)b->ps.size();
  fsgc::EventPair ev;
  HIP_OR_FAIL(ev.create());
  HIP_OR_FAIL(hipEventRecord(ev.a, b->stream));
  HIP_OR_FAIL(hipGraphLaunch(b->exec, b->stream));
  HIP_OR_FAIL(hipEventRecord(ev.b, b->stream));
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  float ms = 0;
  HIP_OR_FAIL(hipEventElapsedTime(&ms, ev.a, ev.b));
  if (graph_ms) *graph_ms = ms;
  std::vector<fsgc::JobResult> res(k);
  for (size_t p = 0; p < b->plans.size(); p++) {
    std::vector<fsgc::JobResult> r;
    fsgc::launch_collect(b->plans[p], &r);
    for (size_t q = 0; q < r.size(); q++) {
      r[q].index_ms = 0;
      res[b->groups[p].idx[q]] = r[q];
    }
  }
  for (int i = 0; i < k; i++) {
    std::memset(&out[i], 0, sizeof(fsgc_result));
    fsgc::finish(b->ps[i], res[i], b->g->dg.n, b->g->dg.m, &out[i]);
    out[i].kernel_ms = ms;
    out[i].wall_ms = ms;
    out[i].alg_bytes = (double)res[i].bytes;
  }
  return 0;
}

void fsgc_batch_free(fsgc_batch* b) { delete b; }

}  // extern "C"
