// Serving loop: resident graphs, count queries admitted by device scratch, folded
// into shared launches and run on several streams at once (continuous batching).
//
// Each slot is a host thread with its own non-blocking HIP stream. A slot that is
// free takes, in arrival order, the first waiting query whose launch fits in the
// scratch budget next to the launches in flight, folds into that launch every other
// waiting query on the same graph and schedule (up to max_batch), and runs it. The
// launch's workgroups are the slot's share of the machine, so launches on different
// slots run side by side on the device. A query that arrives while others run waits
// only for the next free slot, not for a whole round of queries.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "count.h"

namespace {

struct Query {
  uint64_t ticket = 0;
  int graph = 0;
  int sched = 0;
  fsgc::Prepared p;
  fsgc_query out{};
  bool done = false;
  std::string err;  // message of a failed query
};

}  // namespace

struct fsgc_server {
  fsgc_server_config cfg{};
  int device = 0;
  int cus = 256;
  std::vector<fsgc_graph*> graphs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Query*> pending;
  std::map<uint64_t, Query*> live;
  uint64_t next_ticket = 1;
  bool stopping = false;
  size_t inflight_scratch = 0;
  int inflight = 0;
  fsgc_server_stats st{};
  std::vector<std::thread> workers;
  std::vector<hipStream_t> streams;  // one non-blocking stream per slot
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double clock_ms() const
  {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  void worker(int slot);
  bool pick(std::vector<Query*>* batch, int* grid, size_t* scratch);
};

namespace {

// default workgroups of one launch of a schedule when it has the machine to itself
int full_grid(int sched, int cus)
{
  if (sched == fsgc::kClique) return cus * 3;
  return (sched == fsgc::kPivot || sched == fsgc::kRank) ? cus * 4 : cus * 8;
}

}  // namespace

// Called with mu held. Chooses the next launch, or returns false when nothing waiting
// fits now. Queries whose launch cannot fit even on an idle device are failed here.
bool fsgc_server::pick(std::vector<Query*>* batch, int* grid, size_t* scratch)
{
  std::vector<std::pair<int, int>> blocked;  // (graph, schedule) kept in arrival order
  for (size_t i = 0; i < pending.size(); i++) {
    Query* q = pending[i];
    const std::pair<int, int> key(q->graph, q->sched);
    bool skip = false;
    for (auto& b : blocked) skip = skip || b == key;
    if (skip) continue;
    fsgc_graph* g = graphs[q->graph];
    int gr = cfg.grid > 0 ? cfg.grid : std::max(cus, full_grid(q->sched, cus) / std::max(1, cfg.slots));
    fsgc::Tune t = fsgc::tune_for(g, nullptr, q->sched);
    t.grid = gr;
    size_t need = q->p.trivial ? 0 : fsgc::scratch_for(g, q->sched, t) + (1 << 20);
    if (inflight == 0 && inflight_scratch + need > cfg.scratch_cap) {
      // alone on the device: narrow the launch until it fits
      while (gr > 8 && need > cfg.scratch_cap) {
        gr /= 2;
        t.grid = gr;
        need = fsgc::scratch_for(g, q->sched, t) + (1 << 20);
      }
      if (need > cfg.scratch_cap) {
        q->out.status = -12;
        q->out.done_ms = clock_ms();
        q->done = true;
        st.rejected++;
        st.completed++;
        pending.erase(pending.begin() + (long)i);
        cv_done.notify_all();
        i--;
        continue;
      }
    }
    if (inflight_scratch + need > cfg.scratch_cap) {
      blocked.push_back(key);
      continue;
    }
    batch->clear();
    for (size_t j = i; j < pending.size() && (int)batch->size() < cfg.max_batch;) {
      Query* r = pending[j];
      if (r->graph == q->graph && r->sched == q->sched) {
        batch->push_back(r);
        pending.erase(pending.begin() + (long)j);
      } else {
        j++;
      }
    }
    *grid = gr;
    *scratch = need;
    return true;
  }
  return false;
}

void fsgc_server::worker(int slot)
{
  (void)hipSetDevice(device);
  hipStream_t stream = streams[slot];
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    std::vector<Query*> batch;
    int grid = 0;
    size_t scratch = 0;
    while (!pick(&batch, &grid, &scratch)) {
      if (stopping && pending.empty()) return;
      cv_work.wait(lk);
    }
    inflight_scratch += scratch;
    inflight++;
    st.launches++;
    st.scratch_peak = std::max<uint64_t>(st.scratch_peak, inflight_scratch);
    st.peak_concurrent = std::max(st.peak_concurrent, inflight);
    const double start = clock_ms();
    fsgc_graph* g = graphs[batch[0]->graph];
    std::vector<fsgc::Prepared> ps;
    for (Query* q : batch) ps.push_back(q->p);
    lk.unlock();

    fsgc_tune ft;
    std::memset(&ft, 0, sizeof(ft));
    ft.grid = grid;
    std::vector<fsgc::JobResult> res;
    double ms = 0;
    std::string err;
    const int rc = fsgc::run_prepared(g, ps, &ft, &res, &ms, stream, &err);

    lk.lock();
    const double done = clock_ms();
    for (size_t i = 0; i < batch.size(); i++) {
      Query* q = batch[i];
      q->out.status = rc;
      if (rc) q->err = err;
      if (!rc) {
        fsgc::finish(ps[i], res[i], g->dg.n, g->dg.m, &q->out.result);
        q->out.result.kernel_ms = ms;
        q->out.result.wall_ms = done - start;
      }
      q->out.batch_size = (int)batch.size();
      q->out.slot = slot;
      q->out.start_ms = start;
      q->out.done_ms = done;
      q->out.scratch_bytes = scratch;
      q->done = true;
      st.completed++;
    }
    inflight_scratch -= scratch;
    inflight--;
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

int fsgc_server_create(const fsgc_server_config* cfg, fsgc_server** out)
{
  if (!out) { fsgc::set_error("null argument"); return -22; }
  std::string err;
  int rc = fsgc::ensure_init(&err);
  if (rc) { fsgc::set_error(err); return rc; }
  fsgc_server* s = new fsgc_server();
  if (cfg) s->cfg = *cfg;
  if (s->cfg.slots <= 0) s->cfg.slots = 2;
  if (s->cfg.slots > 16) s->cfg.slots = 16;
  if (s->cfg.max_batch <= 0 || s->cfg.max_batch > fsgc::kMaxPivotJobs) s->cfg.max_batch = fsgc::kMaxPivotJobs;
  s->device = fsgc::current_device();
  (void)hipGetDevice(&s->device);
  (void)hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device);
  if (s->cfg.scratch_cap == 0) {
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    s->cfg.scratch_cap = free_b / 2;
  }
  s->st.slots = s->cfg.slots;
  s->st.scratch_cap = s->cfg.scratch_cap;
  for (int i = 0; i < s->cfg.slots; i++) {
    hipStream_t st = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
      delete s;
      fsgc::set_error(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
      return -5;
    }
    s->streams.push_back(st);
  }
  for (int i = 0; i < s->cfg.slots; i++) s->workers.emplace_back([s, i] { s->worker(i); });
  *out = s;
  return 0;
}

int fsgc_server_add_graph(fsgc_server* s, fsgc_graph* g, int32_t* id)
{
  if (!s || !g) { fsgc::set_error("null argument"); return -22; }
  std::string err;
  double ms = 0;
  // the supports now; the common-neighbour lists on the first query that needs them
  const int rc = fsgc::device_derived_build(&g->dg, nullptr, &ms, &err, false);
  if (rc) { fsgc::set_error(err); return rc; }
  std::lock_guard<std::mutex> lk(s->mu);
  s->graphs.push_back(g);
  if (id) *id = (int32_t)s->graphs.size() - 1;
  return 0;
}

int fsgc_server_submit(fsgc_server* s, int32_t graph, const uint8_t mo[FSGC_MO_BYTES], uint64_t* ticket)
{
  if (!s || !mo || !ticket) { fsgc::set_error("null argument"); return -22; }
  Query* q = new Query();
  std::string err;
  const int rc = fsgc::prepare(mo, nullptr, &q->p, &err);
  if (rc) { delete q; fsgc::set_error(err); return rc; }
  std::lock_guard<std::mutex> lk(s->mu);
  if (graph < 0 || graph >= (int32_t)s->graphs.size()) { delete q; fsgc::set_error("no such graph"); return -22; }
  q->sched = fsgc::schedule_of(q->p, true, s->graphs[graph]->dg.maxdeg);
  if (s->stopping) { delete q; fsgc::set_error("server is stopping"); return -22; }
  q->graph = graph;
  q->ticket = s->next_ticket++;
  q->out.graph = graph;
  q->out.submit_ms = s->clock_ms();
  s->live[q->ticket] = q;
  s->pending.push_back(q);
  s->st.submitted++;
  *ticket = q->ticket;
  s->cv_work.notify_one();
  return 0;
}

static int take(fsgc_server* s, std::map<uint64_t, Query*>::iterator it, fsgc_query* out)
{
  Query* q = it->second;
  if (out) *out = q->out;
  s->live.erase(it);
  const int status = q->out.status;
  if (status) fsgc::set_error(status == -12 ? "query needs more device scratch than the server's cap" : q->err);
  delete q;
  return status;
}

int fsgc_server_wait(fsgc_server* s, uint64_t ticket, fsgc_query* out)
{
  if (!s) { fsgc::set_error("null argument"); return -22; }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) { fsgc::set_error("unknown ticket"); return -2; }
  Query* q = it->second;
  s->cv_done.wait(lk, [q] { return q->done; });
  return take(s, s->live.find(ticket), out);
}

int fsgc_server_poll(fsgc_server* s, uint64_t ticket, fsgc_query* out)
{
  if (!s) { fsgc::set_error("null argument"); return -22; }
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) { fsgc::set_error("unknown ticket"); return -2; }
  if (!it->second->done) return 0;
  const int rc = take(s, it, out);
  return rc ? rc : 1;
}

int fsgc_server_stats_get(fsgc_server* s, fsgc_server_stats* out)
{
  if (!s || !out) { fsgc::set_error("null argument"); return -22; }
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  return 0;
}

double fsgc_server_clock_ms(const fsgc_server* s) { return s ? s->clock_ms() : 0.0; }

void fsgc_server_destroy(fsgc_server* s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
  for (auto& kv : s->live) delete kv.second;
  delete s;
}

}  // extern "C"
