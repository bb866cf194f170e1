// Autotuner: per graph and schedule, the launch shape (workgroup size where the kernel
// has several, workgroups, roots per work item, hub threshold, LDS staging cap) with
// the shortest device time on a sample of the work items.
// The counts never depend on the shape (every shape enumerates the same roots), so
// tuning only changes time.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "count.h"

namespace {

struct Cand {
  int block, grid, chunk, hub, lds;
};

std::vector<Cand> candidates(int sched, int cus)
{
  std::vector<Cand> c;
  std::vector<int> grids, chunks = {64, 256, 1024}, hubs = {4096}, blocks = {256}, lds = {0};
  switch (sched) {
    // the LDS cap shortens the pivot kernel's c3 windows (more windows, shorter LDS
    // clears) and sends more of the rank kernel's vertices to global marks
    case fsgc::kPivot: grids = {cus, cus * 2, cus * 4}; blocks = {1024, 768, 256}; lds = {0, 1024}; break;
    case fsgc::kRank: grids = {cus * 2, cus * 4, cus * 8}; lds = {0, 2048}; break;
    case fsgc::kEdgePass: grids = {cus * 4, cus * 8, cus * 16}; break;
    case fsgc::kClique: grids = {cus * 2, cus * 3}; lds = {0, 1024}; break;
    default: grids = {cus * 4, cus * 8, cus * 16}; hubs = {1024, 4096}; break;
  }
  for (int b : blocks)
    for (int g : grids) {
      if (b > 256 && g > cus * 2) continue;  // one 1024-thread workgroup per CU fits; more only queue
      for (int ch : chunks)
        for (int h : hubs)
          for (int l : lds) c.push_back(Cand{b, g, ch, h, l});
    }
  return c;
}

}  // namespace

extern "C" {

int fsgc_autotune(fsgc_graph* g, const uint8_t* mos, int k, int32_t sample, fsgc_tune_report* reports,
                  int32_t* nreports)
{
  if (nreports) *nreports = 0;
  if (!g || k < 0 || (k > 0 && !mos)) { fsgc::set_error("null argument"); return -22; }
  std::string err;
  int rc = fsgc::ensure_init(&err);
  if (rc) { fsgc::set_error(err); return rc; }
  if (sample <= 1) sample = 16;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    fsgc::set_error("cannot query the GPU");
    return -5;
  }
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  std::vector<std::vector<fsgc::Prepared>> by(fsgc::kNumSchedules);
  for (int i = 0; i < k; i++) {
    fsgc::Prepared p;
    rc = fsgc::prepare(mos + (size_t)i * FSGC_MO_BYTES, nullptr, &p, &err);
    if (rc) { fsgc::set_error(err); return rc; }
    if (p.trivial) continue;
    by[fsgc::schedule_of(p, true, g->dg.maxdeg)].push_back(p);
  }
  int nr = 0;
  for (int s = 0; s < fsgc::kNumSchedules; s++) {
    if (by[s].empty()) continue;
    // at most one launch per schedule: the first 8 motifs of the schedule
    if (by[s].size() > (size_t)fsgc::kMaxPivotJobs) by[s].resize(fsgc::kMaxPivotJobs);
    auto time_shape = [&](const fsgc_tune& ft, double* ms) {
      std::vector<fsgc::JobResult> res;
      std::vector<fsgc::Prepared> ps = by[s];
      return fsgc::run_prepared(g, ps, &ft, &res, ms, nullptr, &err, sample);
    };
    const fsgc::Tune d;
    fsgc_tune def;
    std::memset(&def, 0, sizeof(def));
    def.block = s == fsgc::kPivot ? 0 : 256;  // the pivot kernel's default depends on the launch (c3 or not)
    def.grid = s == fsgc::kPivot ? cus : (s == fsgc::kRank ? cus * 4 : (s == fsgc::kClique ? cus * 3 : cus * 8));
    def.chunk = 0;  // each schedule's default item size
    def.hub_degree = d.hub_degree;
    def.lds_cap = d.lds_cap;
    double ms = 0, def_ms = 0;
    rc = time_shape(def, &ms);  // first run also builds the graph's per-edge index
    if (rc) { fsgc::set_error(err); return rc; }
    rc = time_shape(def, &def_ms);
    if (rc) { fsgc::set_error(err); return rc; }
    fsgc_tune best = def;
    double best_ms = def_ms;
    int ncand = 1;
    for (const Cand& c : candidates(s, cus)) {
      fsgc_tune ft = def;
      ft.block = c.block;
      ft.grid = c.grid;
      ft.chunk = c.chunk;
      ft.hub_degree = c.hub;
      ft.lds_cap = c.lds;
      fsgc::Tune t;
      t.block = c.block; t.grid = c.grid; t.chunk = c.chunk; t.hub_degree = c.hub; t.lds_cap = c.lds;
      if (fsgc::scratch_for(g, s, t) > free_b / 4) continue;  // leave room for everything else
      rc = time_shape(ft, &ms);
      if (rc) { fsgc::set_error(err); return rc; }
      ncand++;
      if (fsgc::log_on())
        fprintf(stderr, "[fsgc] tune %s block %d grid %d chunk %d hub %d lds %d: %.2f ms\n", fsgc::schedule_name(s),
                c.block, c.grid, c.chunk, c.hub, c.lds, ms);
      if (ms < best_ms) { best_ms = ms; best = ft; }
    }
    {
      std::lock_guard<std::mutex> lk(g->tune_mu);
      fsgc::Tune& b = g->best[s];
      b.block = best.block;
      b.grid = best.grid;
      b.chunk = best.chunk > 0 ? best.chunk : d.chunk;
      b.chunk_set = best.chunk > 0;
      b.hub_degree = best.hub_degree;
      b.lds_cap = best.lds_cap;
      b.sample = 1;
      g->tuned[s] = true;
    }
    if (reports && nr < FSGC_NUM_SCHEDULES) {
      fsgc_tune_report& r = reports[nr];
      r.schedule = s;
      r.candidates = ncand;
      r.best = best;
      r.best_ms = best_ms;
      r.default_ms = def_ms;
    }
    nr++;
  }
  if (nreports) *nreports = nr;
  return 0;
}

int fsgc_graph_untune(fsgc_graph* g)
{
  if (!g) { fsgc::set_error("null graph"); return -22; }
  std::lock_guard<std::mutex> lk(g->tune_mu);
  for (int s = 0; s < fsgc::kNumSchedules; s++) { g->tuned[s] = false; g->best[s] = fsgc::Tune(); }
  return 0;
}

}  // extern "C"
