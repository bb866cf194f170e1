// C ABI of libfsgc.so (include/fsgc.h): graph residency, motif planning, counting.
#include "../../include/fsgc.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "fsgc_plan.h"

struct fsgc_graph {
  fsgc::DeviceGraph dg;
};

namespace {

thread_local std::string t_err;

int fail(int code, const std::string& msg)
{
  t_err = msg;
  return code;
}

std::mutex g_mu;
bool g_ready = false;
int g_device = 0;

int ensure_init()
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ready) return 0;
  std::string err;
  const int rc = fsgc::device_init(g_device, &err);
  if (rc) return fail(rc, err);
  g_ready = true;
  return 0;
}

// Automorphism sums depend only on the matching order: cache them per motif.
struct AutEntry { fsgc::u128 sum; uint32_t flags; int32_t hn; int64_t hm; };
std::mutex g_aut_mu;
std::map<std::string, AutEntry> g_aut;

fsgc::Tune to_tune(const fsgc_tune* t)
{
  fsgc::Tune r;
  if (!t) return r;
  if (t->block > 0) r.block = t->block;
  if (t->grid > 0) r.grid = t->grid;
  if (t->chunk > 0) r.chunk = t->chunk;
  if (t->hub_degree > 0) r.hub_degree = t->hub_degree;
  if (t->lds_cap > 0) r.lds_cap = t->lds_cap;
  return r;
}

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Prepared {
  fsgc::MatchingOrder mo;
  fsgc::JobDesc job;
  int qf = 1;
  bool trivial = false;       // vertex count / edge count: no kernel
  bool need_aut = false;
  bool aut_cached = false;
  AutEntry aut{};
  std::string key;
  fsgc::HostGraph mgraph;
};

int prepare(const uint8_t* mob, const fsgc_tune* tune, Prepared* p)
{
  std::memcpy(&p->mo, mob, sizeof(p->mo));
  std::string err;
  int rc = fsgc::make_job(p->mo, &p->job, &err);
  if (rc) return fail(rc, err);
  if (tune && tune->slice_mod > 1) { p->job.smod = tune->slice_mod; p->job.srem = tune->slice_rem; }
  p->trivial = (p->job.cls == fsgc::kVertex && p->job.kind != fsgc::kVertexStars);
  p->need_aut = p->job.cls >= fsgc::kWedge;
  p->qf = fsgc::general_quirk_factor(p->mo);
  if (p->need_aut) {
    p->key.assign((const char*)mob, FSGC_MO_BYTES);
    // only the live fields determine the motif
    fsgc::MatchingOrder z;
    std::memset(&z, 0, sizeof(z));
    z.nodes = p->mo.nodes; z.anchorsets = p->mo.anchorsets; z.startdeg = p->mo.startdeg;
    for (int i = 0; i + 1 < p->mo.nodes; i++) { z.destdeg[i] = p->mo.destdeg[i]; z.list[i] = p->mo.list[i]; }
    for (int i = 0; i < p->mo.anchorsets; i++) { z.anch[i] = p->mo.anch[i]; z.fcnt[i] = p->mo.fcnt[i]; }
    p->key.assign((const char*)&z, sizeof(z));
    std::lock_guard<std::mutex> lk(g_aut_mu);
    auto it = g_aut.find(p->key);
    if (it != g_aut.end()) { p->aut = it->second; p->aut_cached = true; }
    else { p->mgraph = fsgc::motif_graph(p->mo); p->aut.hn = p->mgraph.n; p->aut.hm = p->mgraph.m; }
  }
  return 0;
}

void finish(const Prepared& p, const fsgc::JobResult& r, int32_t n, int64_t m, fsgc_result* out)
{
  typedef fsgc::u128 u128;
  out->motif_class = p.job.cls;
  out->core_matches = r.matches;
  uint32_t flags = r.flags;
  u128 sg = r.sum, aut = 1;
  if (p.trivial) sg = (p.job.kind == fsgc::kVertexCount) ? (u128)n : (u128)(m / 2);
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

enum Schedule { kPersistent = 0, kPivot = 1, kRank = 2, kEdgePass = 3 };

// FSGC_LOG=1 in the environment: one line per kernel launch on stderr
bool log_on()
{
  static const bool on = [] { const char* e = getenv("FSGC_LOG"); return e && *e && *e != '0'; }();
  return on;
}

// Which kernel counts a motif on the data graph: the pivot schedule for wedge cores,
// the rank schedule for triangle cores and the triangle kinds of the edge core, the
// per-root persistent kernel otherwise (or for everything with schedule == 1).
Schedule schedule_of(const Prepared& p, bool fast)
{
  if (!fast) return kPersistent;
  if (p.job.cls == fsgc::kWedge) return kPivot;
  if (fsgc::triangle_schedule_ok(p.job)) return kRank;
  if (fsgc::edgepass_schedule_ok(p.job)) return kEdgePass;
  return kPersistent;
}

// Runs the data-graph jobs of all prepared motifs, plus the motif-graph jobs of the
// motifs whose automorphism sum is not cached yet, in one persistent launch.
int run_prepared(fsgc_graph* g, std::vector<Prepared>& ps, const fsgc_tune* tune, std::vector<fsgc::JobResult>* res,
                 double* kernel_ms)
{
  std::string err;
  std::vector<fsgc::JobDesc> jobs;
  std::vector<const fsgc::DeviceGraph*> graphs;
  std::vector<fsgc::DeviceGraph> mgs(ps.size());
  std::vector<int> data_job(ps.size(), -1), aut_job(ps.size(), -1);
  const bool pivot = !(tune && tune->schedule == 1);
  int rc = 0;
  for (size_t i = 0; i < ps.size(); i++) {
    Prepared& p = ps[i];
    if (p.need_aut && !p.aut_cached) {
      if (p.aut.hn > g->dg.n || p.aut.hm > g->dg.m) continue;  // count is 0 regardless
      rc = fsgc::device_graph_upload(p.mgraph, &mgs[i], &err);
      if (rc) break;
      fsgc::JobDesc hj = p.job;
      hj.smod = 1;
      hj.srem = 0;
      aut_job[i] = (int)jobs.size();
      jobs.push_back(hj);
      graphs.push_back(&mgs[i]);
    }
    if (!p.trivial && schedule_of(p, pivot) == kPersistent) {
      data_job[i] = (int)jobs.size();
      jobs.push_back(p.job);
      graphs.push_back(&g->dg);
    }
  }
  std::vector<fsgc::JobResult> jr;
  double ms_total = 0, ms = 0;
  if (!rc) rc = fsgc::device_run_jobs(jobs, graphs, to_tune(tune), nullptr, &jr, &ms, &err);
  ms_total += ms;
  if (log_on() && !jobs.empty()) fprintf(stderr, "[fsgc] per-root launch: %zu jobs, %.1f ms\n", jobs.size(), ms);
  // wedge-class motifs: pivot schedule; triangle-based motifs: rank schedule; each
  // launch takes up to 8 motifs that share one enumeration
  std::vector<fsgc::JobResult> sres(ps.size());
  for (int sched : {kPivot, kRank, kEdgePass}) {
    if (rc || !pivot) break;
    std::vector<std::vector<size_t>> groups;
    for (size_t i = 0; i < ps.size(); i++) {
      if (ps[i].trivial || schedule_of(ps[i], pivot) != sched) continue;
      // tailed 4-cliques with different tail counts go to separate launches
      const int key = (ps[i].job.cls == fsgc::kTriangle && ps[i].job.kind == fsgc::kTriTailedClique) ? ps[i].job.p[3] : -1;
      bool placed = false;
      for (auto& gr : groups) {
        const auto& j0 = ps[gr[0]].job;
        const int k0 = (j0.cls == fsgc::kTriangle && j0.kind == fsgc::kTriTailedClique) ? j0.p[3] : -1;
        if ((int)gr.size() < fsgc::kMaxPivotJobs && (key < 0 || k0 < 0 || key == k0)) { gr.push_back(i); placed = true; break; }
      }
      if (!placed) groups.push_back({i});
    }
    for (auto& gr : groups) {
      std::vector<fsgc::JobDesc> wj;
      for (size_t k : gr) wj.push_back(ps[k].job);
      std::vector<fsgc::JobResult> r;
      if (sched == kPivot) rc = fsgc::device_run_wedge_pivot(wj, &g->dg, to_tune(tune), nullptr, &r, &ms, &err);
      else if (sched == kRank) rc = fsgc::device_run_triangles(wj, &g->dg, to_tune(tune), nullptr, &r, &ms, &err);
      else rc = fsgc::device_run_edgepass(wj, &g->dg, to_tune(tune), nullptr, &r, &ms, &err);
      if (rc) break;
      ms_total += ms;
      if (log_on())
        fprintf(stderr, "[fsgc] %s launch: %zu motifs, %.1f ms\n",
                sched == kPivot ? "pivot" : (sched == kRank ? "rank" : "edge-pass"), gr.size(), ms);
      for (size_t k = 0; k < r.size(); k++) sres[gr[k]] = r[k];
    }
  }
  if (kernel_ms) *kernel_ms = ms_total;
  for (auto& mg : mgs) fsgc::device_graph_free(&mg);
  if (rc) return fail(rc, err);
  res->assign(ps.size(), fsgc::JobResult());
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

}  // namespace

extern "C" {

int fsgc_init(int device)
{
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ready && device != g_device) g_ready = false;
    g_device = device;
  }
  return ensure_init();
}

int fsgc_graph_load_egr(const char* path, fsgc_graph** out)
{
  if (!path || !out) return fail(-22, "null argument");
  fsgc::HostGraph hg;
  std::string err;
  int rc = fsgc::read_egr(path, &hg, &err);
  if (rc) return fail(rc, err);
  return fsgc_graph_from_csr(hg.n, (int32_t)hg.m, hg.nindex.data(), hg.nlist.data(), out);
}

int fsgc_graph_from_csr(int32_t nodes, int32_t edges, const int32_t* nindex, const int32_t* nlist, fsgc_graph** out)
{
  if (!out || !nindex || (edges > 0 && !nlist)) return fail(-22, "null argument");
  if (nodes < 1 || edges < 0) return fail(-22, "ERROR: node or edge count too low");
  if (nindex[0] != 0 || nindex[nodes] != edges) return fail(-22, "nindex does not span nlist");
  int rc = ensure_init();
  if (rc) return rc;
  fsgc::HostGraph hg;
  hg.n = nodes;
  hg.m = edges;
  hg.nindex.assign(nindex, nindex + nodes + 1);
  if (edges) hg.nlist.assign(nlist, nlist + edges);
  fsgc_graph* g = new fsgc_graph();
  std::string err;
  rc = fsgc::device_graph_upload(hg, &g->dg, &err);
  if (rc) { fsgc::device_graph_free(&g->dg); delete g; return fail(rc, err); }
  *out = g;
  return 0;
}

int fsgc_graph_validate(const fsgc_graph* g)
{
  if (!g) return fail(-22, "null graph");
  fsgc::HostGraph hg;
  hg.n = g->dg.n;
  hg.m = g->dg.m;
  hg.nindex = g->dg.h_idx;
  hg.nlist.resize((size_t)hg.m);
  if (hg.m && hipMemcpy(hg.nlist.data(), g->dg.adj, sizeof(int32_t) * hg.m, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(-5, "copy of the neighbour list failed");
  std::string err;
  const int rc = fsgc::validate_csr(hg, &err);
  return rc ? fail(rc, err) : 0;
}

int fsgc_graph_info(const fsgc_graph* g, int32_t* nodes, int32_t* edges, int32_t* maxdeg)
{
  if (!g) return fail(-22, "null graph");
  if (nodes) *nodes = g->dg.n;
  if (edges) *edges = (int32_t)g->dg.m;
  if (maxdeg) *maxdeg = g->dg.maxdeg;
  return 0;
}

void fsgc_graph_free(fsgc_graph* g)
{
  if (!g) return;
  fsgc::device_graph_free(&g->dg);
  delete g;
}

int fsgc_count(fsgc_graph* g, const uint8_t mo[FSGC_MO_BYTES], const fsgc_tune* tune, fsgc_result* out)
{
  return fsgc_count_batch(g, mo, 1, tune, out, nullptr);
}

int fsgc_count_batch(fsgc_graph* g, const uint8_t* mos, int k, const fsgc_tune* tune, fsgc_result* out, double* batch_ms)
{
  if (!g || !mos || !out || k < 0) return fail(-22, "null argument");
  const double t0 = now_ms();
  int rc = ensure_init();
  if (rc) return rc;
  std::vector<Prepared> ps(k);
  for (int i = 0; i < k; i++) {
    rc = prepare(mos + (size_t)i * FSGC_MO_BYTES, tune, &ps[i]);
    if (rc) return rc;
  }
  std::vector<fsgc::JobResult> res;
  double ms = 0;
  rc = run_prepared(g, ps, tune, &res, &ms);
  if (rc) return rc;
  const double wall = now_ms() - t0;
  for (int i = 0; i < k; i++) {
    std::memset(&out[i], 0, sizeof(fsgc_result));
    finish(ps[i], res[i], g->dg.n, g->dg.m, &out[i]);
    out[i].kernel_ms = ms;
    out[i].wall_ms = wall;
    out[i].alg_bytes = (double)res[i].bytes;
    out[i].index_ms = res[i].index_ms;
  }
  if (batch_ms) *batch_ms = ms;
  return 0;
}

static void copy_text(const std::string& s, char* buf, size_t cap)
{
  if (!buf || cap == 0) return;
  const size_t n = std::min(cap - 1, s.size());
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
}

int fsgc_preprocess(int nodes, int nedges, const int* edges, uint8_t mo[FSGC_MO_BYTES], char* log, size_t logcap,
                    char* dot, size_t dotcap)
{
  if (!mo || (nedges > 0 && !edges) || nedges < 0) return fail(-22, "null argument");
  fsgc::Pattern p;
  std::string err, lg, dt;
  int rc = fsgc::pattern_from_edges(nodes, nedges, edges, &p, &err);
  if (rc) return fail(rc, err);
  fsgc::MatchingOrder m;
  rc = fsgc::preprocess(p, &m, &lg, &dt, &err);
  copy_text(lg, log, logcap);
  copy_text(dt, dot, dotcap);
  if (rc) return fail(rc, err);
  std::memcpy(mo, &m, FSGC_MO_BYTES);
  return 0;
}

int fsgc_preprocess_pattern(int number, uint8_t mo[FSGC_MO_BYTES], char* log, size_t logcap, char* dot, size_t dotcap)
{
  if (!mo) return fail(-22, "null argument");
  fsgc::Pattern p;
  if (fsgc::predefined_pattern(number, &p)) return fail(-22, "ERROR: unknown pattern");
  std::string err, lg, dt;
  fsgc::MatchingOrder m;
  int rc = fsgc::preprocess(p, &m, &lg, &dt, &err);
  copy_text(lg, log, logcap);
  copy_text(dt, dot, dotcap);
  if (rc) return fail(rc, err);
  std::memcpy(mo, &m, FSGC_MO_BYTES);
  return 0;
}

int fsgc_print_matching_order(const uint8_t mo[FSGC_MO_BYTES], char* out, size_t cap)
{
  if (!mo || !out) return fail(-22, "null argument");
  fsgc::MatchingOrder m;
  std::memcpy(&m, mo, sizeof(m));
  std::string s;
  fsgc::print_matching_order(m, &s);
  copy_text(s, out, cap);
  return 0;
}

static int generate_host(const char* spec, fsgc::HostGraph* hg)
{
  if (!spec) return fail(-22, "null spec");
  std::string s(spec);
  std::vector<std::string> f;
  size_t pos = 0;
  for (;;) {
    const size_t c = s.find(':', pos);
    f.push_back(s.substr(pos, c == std::string::npos ? std::string::npos : c - pos));
    if (c == std::string::npos) break;
    pos = c + 1;
  }
  int rc = -22;
  try {
    if (f[0] == "er" && f.size() == 4)
      rc = fsgc::gen_er(std::stoi(f[1]), std::stoll(f[2]), std::stoull(f[3]), hg);
    else if (f[0] == "rmat" && f.size() == 7)
      rc = fsgc::gen_rmat(std::stoi(f[1]), std::stoi(f[2]), std::stod(f[3]), std::stod(f[4]), std::stod(f[5]),
                          std::stoull(f[6]), hg);
    else if (f[0] == "chunglu" && f.size() == 6)
      rc = fsgc::gen_chung_lu(std::stoi(f[1]), std::stod(f[2]), std::stod(f[3]), std::stod(f[4]), std::stoull(f[5]), hg);
    else if (f[0] == "egr" && f.size() >= 2) {
      std::string err;
      rc = fsgc::read_egr(s.substr(4).c_str(), hg, &err);
      if (rc) return fail(rc, err);
    } else
      return fail(-22, "unknown graph spec: " + s);
  } catch (const std::exception& e) {
    return fail(-22, std::string("bad graph spec: ") + e.what());
  }
  if (rc) return fail(rc, "graph generation failed: " + s);
  if (hg->m > INT32_MAX) return fail(-75, "graph has more than 2^31-1 directed edges");
  return 0;
}

int fsgc_generate(const char* spec, int32_t* nodes, int32_t* edges, int32_t** nindex, int32_t** nlist)
{
  if (!nodes || !edges || !nindex || !nlist) return fail(-22, "null argument");
  fsgc::HostGraph hg;
  int rc = generate_host(spec, &hg);
  if (rc) return rc;
  *nodes = hg.n;
  *edges = (int32_t)hg.m;
  *nindex = (int32_t*)malloc(sizeof(int32_t) * (hg.nindex.size()));
  *nlist = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>(1, hg.nlist.size()));
  if (!*nindex || !*nlist) { free(*nindex); free(*nlist); return fail(-12, "out of host memory"); }
  std::memcpy(*nindex, hg.nindex.data(), sizeof(int32_t) * hg.nindex.size());
  if (hg.m) std::memcpy(*nlist, hg.nlist.data(), sizeof(int32_t) * hg.nlist.size());
  return 0;
}

int fsgc_graph_generate(const char* spec, fsgc_graph** out)
{
  fsgc::HostGraph hg;
  int rc = generate_host(spec, &hg);
  if (rc) return rc;
  return fsgc_graph_from_csr(hg.n, (int32_t)hg.m, hg.nindex.data(), hg.nlist.data(), out);
}

void fsgc_free_host(void* p) { free(p); }

int fsgc_random_motifs(uint64_t seed, int count, uint8_t* mos)
{
  if (!mos || count < 0) return fail(-22, "null argument");
  std::vector<fsgc::Pattern> ps;
  fsgc::gen_random_motifs(seed, count, 2, 5, 6, &ps);
  for (int i = 0; i < count; i++) {
    fsgc::MatchingOrder m;
    std::string err;
    const int rc = fsgc::preprocess(ps[i], &m, nullptr, nullptr, &err);
    if (rc) return fail(rc, err);
    std::memcpy(mos + (size_t)i * FSGC_MO_BYTES, &m, FSGC_MO_BYTES);
  }
  return 0;
}

int fsgc_random_patterns(uint64_t seed, int count, int* nodes, int* nedges, int* edges)
{
  if (!nodes || !nedges || !edges || count < 0) return fail(-22, "null argument");
  std::vector<fsgc::Pattern> ps;
  fsgc::gen_random_motifs(seed, count, 2, 5, 6, &ps);
  for (int i = 0; i < count; i++) {
    nodes[i] = ps[i].nodes;
    int k = 0;
    for (int a = 0; a < ps[i].nodes; a++)
      for (uint32_t m = ps[i].adj[a]; m; m &= m - 1) {
        const int b = __builtin_ctz(m);
        if (a < b && k < 64) { edges[i * 128 + 2 * k] = a; edges[i * 128 + 2 * k + 1] = b; k++; }
      }
    nedges[i] = k;
  }
  return 0;
}

const char* fsgc_last_error(void) { return t_err.c_str(); }

const char* fsgc_strerror(int code)
{
  switch (code) {
    case 0: return "success";
    case -2: return "not found";
    case -5: return "I/O or device error";
    case -7: return "argument list too long";
    case -12: return "out of memory";
    case -22: return "invalid argument";
    case -75: return "value too large";
    default: return "unknown error";
  }
}

}  // extern "C"
