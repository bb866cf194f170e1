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

#include "count.h"

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

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace fsgc {

int ensure_init(std::string* err)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ready) return 0;
  const int rc = device_init(g_device, err);
  if (rc) return rc;
  g_ready = true;
  return 0;
}

int current_device() { return g_device; }

void set_error(const std::string& msg) { t_err = msg; }

}  // namespace fsgc

namespace {

int ensure_init()
{
  std::string err;
  const int rc = fsgc::ensure_init(&err);
  return rc ? fail(rc, err) : 0;
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
  fsgc::HostGraph hg;
  hg.n = nodes;
  hg.m = edges;
  hg.nindex.assign(nindex, nindex + nodes + 1);
  if (edges) hg.nlist.assign(nlist, nlist + edges);
  std::string err;
  // no kernel or host index build may see a malformed graph: a non-monotone nindex or
  // an id out of range would index past the device arrays (checked before the GPU is
  // touched)
  int rc = fsgc::validate_csr(hg, &err);
  if (rc) return fail(rc, err);
  rc = ensure_init();
  if (rc) return rc;
  fsgc_graph* g = new fsgc_graph();
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

int fsgc_graph_relabel(const fsgc_graph* g, int32_t order, fsgc_graph** out, int32_t* map, double* relabel_ms)
{
  if (!g || !out) return fail(-22, "null argument");
  if (order != FSGC_RELABEL_DEGREE_DESC && order != FSGC_RELABEL_DEGREE_ASC) return fail(-22, "unknown relabel order");
  int rc = ensure_init();
  if (rc) return rc;
  fsgc_graph* r = new fsgc_graph();
  std::vector<int32_t> h_map;
  std::string err;
  double ms = 0;
  rc = fsgc::device_graph_relabel(g->dg, order, &r->dg, map ? &h_map : nullptr, &ms, &err);
  if (rc) { fsgc::device_graph_free(&r->dg); delete r; return fail(rc, err); }
  if (map) std::memcpy(map, h_map.data(), sizeof(int32_t) * h_map.size());
  if (relabel_ms) *relabel_ms = ms;
  *out = r;
  return 0;
}

int fsgc_graph_copy_csr(const fsgc_graph* g, int32_t* nindex, int32_t* nlist)
{
  if (!g || !nindex || (g->dg.m > 0 && !nlist)) return fail(-22, "null argument");
  std::memcpy(nindex, g->dg.h_idx.data(), sizeof(int32_t) * ((size_t)g->dg.n + 1));
  if (g->dg.m && hipMemcpy(nlist, g->dg.adj, sizeof(int32_t) * (size_t)g->dg.m, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(-5, "copy of the neighbour list failed");
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
  std::vector<fsgc::Prepared> ps(k);
  std::string err;
  for (int i = 0; i < k; i++) {
    rc = fsgc::prepare(mos + (size_t)i * FSGC_MO_BYTES, tune, &ps[i], &err);
    if (rc) return fail(rc, err);
  }
  std::vector<fsgc::JobResult> res;
  double ms = 0;
  rc = fsgc::run_prepared(g, ps, tune, &res, &ms, nullptr, &err);
  if (rc) return fail(rc, err);
  const double wall = now_ms() - t0;
  for (int i = 0; i < k; i++) {
    std::memset(&out[i], 0, sizeof(fsgc_result));
    fsgc::finish(ps[i], res[i], g->dg.n, g->dg.m, &out[i]);
    out[i].kernel_ms = ms;
    out[i].wall_ms = wall;
    out[i].alg_bytes = (double)res[i].bytes;
    out[i].index_ms = res[i].index_ms;
  }
  if (batch_ms) *batch_ms = ms;
  return 0;
}

int fsgc_count_noalloc(fsgc_graph* g, const uint8_t mo[FSGC_MO_BYTES], fsgc_result* out)
{
  if (!g || !mo || !out) return fail(-22, "null argument");
  const double t0 = now_ms();
  int rc = ensure_init();
  if (rc) return rc;
  if (hipSetDevice(fsgc::current_device()) != hipSuccess) return fail(-5, "cannot select the GPU");
  fsgc::Prepared p;
  std::string err;
  rc = fsgc::prepare(mo, nullptr, &p, &err);
  if (rc) return fail(rc, err);
  fsgc::JobResult r;
  double ms = 0;
  if (!p.trivial) {
    if (fsgc::schedule_of(p, true, g->dg.maxdeg) != fsgc::kEdgePass)
      return fail(-95, "no-allocation count: only vertex motifs and edge motifs of the closed-form kind "
                       "(counted by the edge pass) are supported");
    if (!g->dg.support_ready)
      return fail(-11, "no-allocation count: the graph's per-edge supports are not built yet (any fsgc_count "
                       "of an edge, wedge or triangle motif on the graph builds them)");
    rc = fsgc::edgepass_count_noalloc(&g->dg, p.job, &r, &ms, &err);
    if (rc) return fail(rc, err);
  }
  std::memset(out, 0, sizeof(fsgc_result));
  fsgc::finish(p, r, g->dg.n, g->dg.m, out);
  out->kernel_ms = ms;
  out->wall_ms = now_ms() - t0;
  out->alg_bytes = (double)r.bytes;
  return 0;
}

int fsgc_plan_schedule(const uint8_t mo[FSGC_MO_BYTES], int32_t maxdeg, int32_t* schedule, int32_t* motif_class)
{
  if (!mo || !schedule) return fail(-22, "null argument");
  fsgc::Prepared p;
  std::string err;
  const int rc = fsgc::prepare(mo, nullptr, &p, &err);
  if (rc) return fail(rc, err);
  *schedule = p.trivial ? -1 : (int32_t)fsgc::schedule_of(p, true, maxdeg);
  if (motif_class) *motif_class = p.job.cls;
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
