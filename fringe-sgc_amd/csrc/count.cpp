// Host side of a counting call (see count.h). The class dispatch, the automorphism
// run on the motif graph and the final division follow occurrences() of the
// reference (fringeCount.cu:2199-2262); which kernel runs each class is this
// library's choice (pivot / rank / edge-pass schedules, or the per-root kernel).
#include "count.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

#include "devscratch.h"

namespace fsgc {

namespace {
std::mutex g_aut_mu;
std::map<std::string, AutEntry> g_aut;
}  // namespace

const char* schedule_name(int s)
{
  switch (s) {
    case kPersistent: return "per-root";
    case kPivot: return "pivot";
    case kRank: return "rank";
    case kEdgePass: return "edge-pass";
    case kClique: return "clique";
    default: return "?";
  }
}

// FSGC_LOG=1 in the environment: one line per kernel launch on stderr
bool log_on()
{
  static const bool on = [] { const char* e = getenv("FSGC_LOG"); return e && *e && *e != '0'; }();
  return on;
}

int prepare(const uint8_t* mob, const fsgc_tune* tune, Prepared* p, std::string* err)
{
  std::memcpy(&p->mo, mob, sizeof(p->mo));
  int rc = make_job(p->mo, &p->job, err);
  if (rc) return rc;
  if (tune && tune->slice_mod > 1) { p->job.smod = tune->slice_mod; p->job.srem = tune->slice_rem; }
  p->trivial = (p->job.cls == kVertex && p->job.kind != kVertexStars);
  p->need_aut = p->job.cls >= kWedge;
  p->qf = general_quirk_factor(p->mo);
  if (p->need_aut) {
    // only the live fields determine the motif
    MatchingOrder z;
    std::memset(&z, 0, sizeof(z));
    z.nodes = p->mo.nodes; z.anchorsets = p->mo.anchorsets; z.startdeg = p->mo.startdeg;
    for (int i = 0; i + 1 < p->mo.nodes; i++) { z.destdeg[i] = p->mo.destdeg[i]; z.list[i] = p->mo.list[i]; }
    for (int i = 0; i < p->mo.anchorsets; i++) { z.anch[i] = p->mo.anch[i]; z.fcnt[i] = p->mo.fcnt[i]; }
    p->key.assign((const char*)&z, sizeof(z));
    std::lock_guard<std::mutex> lk(g_aut_mu);
    auto it = g_aut.find(p->key);
    if (it != g_aut.end()) { p->aut = it->second; p->aut_cached = true; }
    else { p->mgraph = motif_graph(p->mo); p->aut.hn = p->mgraph.n; p->aut.hm = p->mgraph.m; }
  }
  return 0;
}

void finish(const Prepared& p, const JobResult& r, int32_t n, int64_t m, fsgc_result* out)
{
  out->motif_class = p.job.cls;
  out->core_matches = r.matches;
  uint32_t flags = r.flags;
  u128 sg = r.sum, aut = 1;
  if (p.trivial) sg = (p.job.kind == kVertexCount) ? (u128)n : (u128)(m / 2);
  if (p.need_aut) {
    if (p.aut.hn > n || p.aut.hm > m) {  // FC:2216-2219: the motif cannot fit
      std::memset(out, 0, offsetof(fsgc_result, kernel_ms));
      out->motif_class = p.job.cls;
      return;
    }
    aut = p.aut.sum * (u128)p.qf;
    sg = sg * (u128)p.qf;
    flags |= p.aut.flags;
  }
  out->sg_lo = (uint64_t)r.sum;
  out->sg_hi = (uint64_t)(r.sum >> 64);
  if (p.trivial) { out->sg_lo = (uint64_t)sg; out->sg_hi = (uint64_t)(sg >> 64); }
  out->aut = (uint64_t)aut;
  if (aut == 0) { out->exact_lo = out->exact_hi = out->ref_compat = 0; out->flags = flags | FSGC_AUT_REMAINDER; return; }
  const u128 q = sg / aut;
  if (p.job.smod <= 1 && (sg % aut) != 0) flags |= FSGC_AUT_REMAINDER;
  if ((q >> 64) != 0) flags |= FSGC_OVF64;
  out->exact_lo = (uint64_t)q;
  out->exact_hi = (uint64_t)(q >> 64);
  // fringeCount.cu:2213, 2230: the automorphism sum passes through an int
  const uint64_t ref_aut = (uint64_t)(int64_t)(int32_t)(uint32_t)(uint64_t)aut;
  if ((u128)ref_aut != aut) flags |= FSGC_REF_AUT_TRUNC;
  out->ref_compat = ref_aut ? (uint64_t)sg / ref_aut : 0;
  out->flags = flags;
}

// The pivot schedule for wedge cores, the rank schedule for triangle cores and the
// triangle kinds of the edge core, one edge pass for the other edge kinds and the
// stars, the clique schedule for 4-clique cores whose fringe counts it evaluates
// exactly on this graph, the per-root persistent kernel otherwise (or for everything
// when !fast).
// Wedge cores on a graph whose maximum degree exceeds kPivotMaxDegree take the
// per-root kernel: the pivot kernel's 24-bit 2-hop counters could not hold a codegree.
Schedule schedule_of(const Prepared& p, bool fast, int32_t maxdeg)
{
  if (!fast) return kPersistent;
  if (p.job.cls == kWedge) return maxdeg <= kPivotMaxDegree ? kPivot : kPersistent;
  if (triangle_schedule_ok(p.job)) return kRank;
  if (edgepass_schedule_ok(p.job)) return kEdgePass;
  if (clique_fits(p.job, maxdeg)) return kClique;
  return kPersistent;
}

Tune tune_for(fsgc_graph* g, const fsgc_tune* t, int sched)
{
  Tune r;
  if (g && sched >= 0 && sched < kNumSchedules) {
    std::lock_guard<std::mutex> lk(g->tune_mu);
    if (g->tuned[sched]) r = g->best[sched];
  }
  if (!t) return r;
  if (t->block > 0) r.block = t->block;
  if (t->grid > 0) r.grid = t->grid;
  if (t->chunk > 0) { r.chunk = t->chunk; r.chunk_set = true; }
  if (t->hub_degree > 0) r.hub_degree = t->hub_degree;
  if (t->lds_cap > 0) r.lds_cap = t->lds_cap;
  return r;
}

size_t scratch_for(const fsgc_graph* g, int sched, const Tune& t)
{
  switch (sched) {
    case kPivot: return pivot_scratch_bytes(g->dg, t);
    case kRank: return triangle_scratch_bytes(g->dg, t);
    case kEdgePass: return edgepass_scratch_bytes(g->dg, t);
    case kClique: return clique_scratch_bytes(g->dg, t);
    default: return perroot_scratch_bytes(g->dg, t);
  }
}

namespace {

// Motifs share a launch only within a tier: a launch runs the kernel variant its most
// demanding motif needs (the c3 pass, the full fc3 loop), so mixing tiers would make
// the cheap motifs pay for it.
// a wedge or triangle job whose fringe vectors have a multi-vertex anchor set (the
// full fc3 kernel variant)
bool needs_full(const JobDesc& j)
{
  if (j.cls != kWedge && j.cls != kTriangle) return false;
  bool multi = false;
  // fringes anchored on a pair of core vertices (those on all three have an inline form)
  for (int r = 0; r < j.rot; r++) multi = multi || (j.fr[r][3] | j.fr[r][5] | j.fr[r][6]) != 0;
  return multi;
}

int group_tier(const JobDesc& j)
{
  const bool multi = needs_full(j);
  if (j.cls == kEdge) return 0;
  // wedge jobs linear in the triple region need no c3 pass, like the specialised ones
  if (j.cls == kWedge) return (j.kind == kWedgeSpecial || j.p[11]) ? 0 : (multi ? 2 : 1);
  if (j.cls == kTriangle) return j.kind != kTriGeneral ? 1 : (multi ? 3 : 2);
  return 0;
}

}  // namespace

std::vector<Group> make_groups(const std::vector<Prepared>& ps, bool fast, int32_t maxdeg)
{
  std::vector<Group> out;
  Group perroot{kPersistent, {}};
  for (size_t i = 0; i < ps.size(); i++)
    if (!ps[i].trivial && schedule_of(ps[i], fast, maxdeg) == kPersistent) perroot.idx.push_back(i);
  if (!perroot.idx.empty()) out.push_back(perroot);
  if (!fast) return out;
  // each fast-schedule launch takes up to 8 motifs that share one enumeration
  for (int sched : {kPivot, kRank, kEdgePass, kClique}) {
    std::vector<Group> groups;
    for (size_t i = 0; i < ps.size(); i++) {
      if (ps[i].trivial || schedule_of(ps[i], fast, maxdeg) != sched) continue;
      // tailed 4-cliques with different tail counts go to separate launches
      const int key = (ps[i].job.cls == kTriangle && ps[i].job.kind == kTriTailedClique) ? ps[i].job.p[3] : -1;
      const int tier = group_tier(ps[i].job);
      bool placed = false;
      for (auto& gr : groups) {
        const auto& j0 = ps[gr.idx[0]].job;
        const int k0 = (j0.cls == kTriangle && j0.kind == kTriTailedClique) ? j0.p[3] : -1;
        // pivot launches of the full fc3 carry at most two motifs: their per-wedge
        // formula dominates, and the 8-motif variant of that kernel spills to scratch
        const int cap = (j0.cls == kWedge && needs_full(j0)) ? 2 : kMaxPivotJobs;
        if ((int)gr.idx.size() < cap && group_tier(j0) == tier && (key < 0 || k0 < 0 || key == k0)) {
          gr.idx.push_back(i);
          placed = true;
          break;
        }
      }
      if (!placed) groups.push_back(Group{sched, {i}});
    }
    out.insert(out.end(), groups.begin(), groups.end());
  }
  return out;
}

int plan_group(fsgc_graph* g, const std::vector<Prepared>& ps, const Group& gr, const fsgc_tune* tune, int sample,
               void* stream, LaunchPlan* lp, std::string* err)
{
  Tune t = tune_for(g, tune, gr.sched);
  t.sample = sample;
  std::vector<JobDesc> jobs;
  for (size_t k : gr.idx) jobs.push_back(ps[k].job);
  switch (gr.sched) {
    case kPivot: return plan_wedge_pivot(jobs, &g->dg, t, stream, lp, err);
    case kRank: return plan_triangles(jobs, &g->dg, t, stream, lp, err);
    case kEdgePass: return plan_edgepass(jobs, &g->dg, t, stream, lp, err);
    case kClique: return plan_cliques(jobs, &g->dg, t, stream, lp, err);
    default: {
      std::vector<const DeviceGraph*> graphs(jobs.size(), &g->dg);
      return plan_jobs(jobs, graphs, t, stream, lp, err);
    }
  }
}

int run_prepared(fsgc_graph* g, std::vector<Prepared>& ps, const fsgc_tune* tune, std::vector<JobResult>* res,
                 double* kernel_ms, void* stream, std::string* err, int sample)
{
  std::vector<JobDesc> jobs;
  std::vector<const DeviceGraph*> graphs;
  std::vector<DeviceGraph> mgs(ps.size());
  std::vector<int> data_job(ps.size(), -1), aut_job(ps.size(), -1);
  const bool fast = !(tune && tune->schedule == 1);
  DevScratch msc((hipStream_t)stream);  // motif graphs of the automorphism runs
  int rc = 0;
  for (size_t i = 0; i < ps.size(); i++) {
    Prepared& p = ps[i];
    if (p.need_aut && !p.aut_cached && sample <= 1) {
      if (p.aut.hn > g->dg.n || p.aut.hm > g->dg.m) continue;  // count is 0 regardless
      rc = device_graph_upload_async(p.mgraph, &mgs[i], &msc, err);
      if (rc) break;
      JobDesc hj = p.job;
      hj.smod = 1;
      hj.srem = 0;
      aut_job[i] = (int)jobs.size();
      jobs.push_back(hj);
      graphs.push_back(&mgs[i]);
    }
    if (!p.trivial && schedule_of(p, fast, g->dg.maxdeg) == kPersistent) {
      data_job[i] = (int)jobs.size();
      jobs.push_back(p.job);
      graphs.push_back(&g->dg);
    }
  }
  std::vector<JobResult> jr;
  double ms_total = 0, ms = 0;
  if (!rc) {
    Tune t = tune_for(g, tune, kPersistent);
    t.sample = sample;
    rc = device_run_jobs(jobs, graphs, t, stream, &jr, &ms, err);
  }
  ms_total += ms;
  if (log_on() && !jobs.empty()) fprintf(stderr, "[fsgc] per-root launch: %zu jobs, %.1f ms\n", jobs.size(), ms);
  std::vector<JobResult> sres(ps.size());
  if (!rc) {
    for (const Group& gr : make_groups(ps, fast, g->dg.maxdeg)) {
      if (gr.sched == kPersistent) continue;  // ran above, with the automorphism jobs
      LaunchPlan lp;
      std::vector<JobResult> r;
      rc = plan_group(g, ps, gr, tune, sample, stream, &lp, err);
      if (!rc) rc = launch_run(lp, stream, &r, &ms, err);
      if (rc) break;
      ms_total += ms;
      if (log_on()) fprintf(stderr, "[fsgc] %s launch: %zu motifs, %.1f ms\n", schedule_name(gr.sched), gr.idx.size(), ms);
      for (size_t k = 0; k < r.size(); k++) sres[gr.idx[k]] = r[k];
    }
  }
  if (kernel_ms) *kernel_ms = ms_total;
  if (rc) return rc;
  res->assign(ps.size(), JobResult());
  for (size_t i = 0; i < ps.size(); i++) {
    if (data_job[i] >= 0) (*res)[i] = jr[data_job[i]];
    else if (!ps[i].trivial) (*res)[i] = sres[i];
    if (aut_job[i] >= 0) {
      Prepared& p = ps[i];
      p.aut.sum = jr[aut_job[i]].sum;
      p.aut.flags = jr[aut_job[i]].flags;
      p.aut_cached = true;
      std::lock_guard<std::mutex> lk(g_aut_mu);
      g_aut[p.key] = p.aut;
    }
  }
  return 0;
}

}  // namespace fsgc
