// ECLgraph .egr reader/writer and matching-order printing.
#include "fsgc_host.h"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <memory>

namespace fsgc {

namespace {
struct FileCloser { void operator()(FILE* f) const { if (f) fclose(f); } };
using File = std::unique_ptr<FILE, FileCloser>;
}

// ECLgraph.h:58-85: int32 nodes, int32 edges, int32 nindex[nodes + 1], int32 nlist[edges],
// optional int32 eweight[edges] (ignored). Same checks and messages, but errors are
// returned instead of exit(-1).
int read_egr(const char* path, HostGraph* g, std::string* err)
{
  File f(fopen(path, "rb"));
  if (!f) { if (err) *err = std::string("ERROR: could not open file ") + path; return -2; }
  int32_t hdr[2];
  if (fread(&hdr[0], 4, 1, f.get()) != 1) { if (err) *err = "ERROR: failed to read nodes"; return -5; }
  if (fread(&hdr[1], 4, 1, f.get()) != 1) { if (err) *err = "ERROR: failed to read edges"; return -5; }
  if (hdr[0] < 1 || hdr[1] < 0) { if (err) *err = "ERROR: node or edge count too low"; return -22; }
  g->n = hdr[0];
  g->m = hdr[1];
  g->nindex.resize((size_t)g->n + 1);
  g->nlist.resize((size_t)g->m);
  if (fread(g->nindex.data(), 4, g->nindex.size(), f.get()) != g->nindex.size()) {
    if (err) *err = "ERROR: failed to read neighbor index list";
    return -5;
  }
  if (g->m > 0 && fread(g->nlist.data(), 4, g->nlist.size(), f.get()) != g->nlist.size()) {
    if (err) *err = "ERROR: failed to read neighbor list";
    return -5;
  }
  return 0;
}

int write_egr(const char* path, const HostGraph& g, std::string* err)
{
  File f(fopen(path, "wb"));
  if (!f) { if (err) *err = std::string("ERROR: could not create file ") + path; return -2; }
  const int32_t hdr[2] = {g.n, (int32_t)g.m};
  bool ok = fwrite(hdr, 4, 2, f.get()) == 2;
  ok = ok && fwrite(g.nindex.data(), 4, g.nindex.size(), f.get()) == g.nindex.size();
  ok = ok && (g.m == 0 || fwrite(g.nlist.data(), 4, g.nlist.size(), f.get()) == g.nlist.size());
  if (!ok) { if (err) *err = "ERROR: write failed"; return -5; }
  return 0;
}

// The preconditions FC:50-56 states but never checks: no self loops, no duplicates,
// symmetric, sorted lists. Also structural sanity of nindex. Runs at every upload, so
// the per-vertex passes are parallel (O(m log d) in all); the message names the
// lowest failing vertex.
int validate_csr(const HostGraph& g, std::string* err)
{
  if (g.n < 1 || (int64_t)g.nindex.size() != (int64_t)g.n + 1 || (int64_t)g.nlist.size() != g.m) {
    if (err) *err = "invalid CSR sizes";
    return -22;
  }
  if (g.m > INT32_MAX) { if (err) *err = "more than 2^31-1 directed edges"; return -22; }
  if (g.nindex[0] != 0 || g.nindex[g.n] != g.m) { if (err) *err = "nindex does not span nlist"; return -22; }
  for (int32_t u = 0; u < g.n; u++)
    if (g.nindex[u + 1] < g.nindex[u]) { if (err) *err = "nindex not monotone at " + std::to_string(u); return -22; }
  const int32_t* nl = g.nlist.data();
  const int32_t* ni = g.nindex.data();
  // pass 1: ids in range, no self loop, strictly ascending lists
  int32_t bad = INT32_MAX;
#pragma omp parallel for schedule(dynamic, 4096) reduction(min : bad)
  for (int32_t u = 0; u < g.n; u++) {
    for (int32_t i = ni[u]; i < ni[u + 1]; i++) {
      const int32_t v = nl[i];
      if (v < 0 || v >= g.n || v == u || (i > ni[u] && nl[i - 1] >= v)) { bad = std::min(bad, u); break; }
    }
  }
  if (bad != INT32_MAX) {
    const int32_t u = bad;
    for (int32_t i = ni[u]; i < ni[u + 1]; i++) {
      const int32_t v = nl[i];
      if (v < 0 || v >= g.n) { if (err) *err = "neighbor out of range at vertex " + std::to_string(u); return -22; }
      if (v == u) { if (err) *err = "self loop at vertex " + std::to_string(u); return -22; }
    }
    if (err) *err = "unsorted or duplicate list at vertex " + std::to_string(u);
    return -22;
  }
  // pass 2 (lists are valid now): every (u, v) has (v, u)
  bad = INT32_MAX;
#pragma omp parallel for schedule(dynamic, 4096) reduction(min : bad)
  for (int32_t u = 0; u < g.n; u++) {
    for (int32_t i = ni[u]; i < ni[u + 1]; i++) {
      const int32_t v = nl[i];
      const int32_t* b = nl + ni[v];
      const int32_t* e = nl + ni[v + 1];
      const int32_t* p = std::lower_bound(b, e, u);
      if (p == e || *p != u) { bad = std::min(bad, u); break; }
    }
  }
  if (bad != INT32_MAX) {
    if (err) *err = "graph is not symmetric at vertex " + std::to_string(bad);
    return -22;
  }
  return 0;
}

// Text of fringes.h:61-98 (printMatchingOrder), kept line-for-line compatible because
// scripts read the fringeCount output.
void print_matching_order(const MatchingOrder& mo, std::string* out)
{
  char buf[256];
  auto put = [&](const char* s) { out->append(s); };
  snprintf(buf, sizeof(buf), "starting from %c >= %d\n\n", 'A', mo.startdeg); put(buf);
  for (int i = 0; i < mo.nodes - 1; i++) {
    snprintf(buf, sizeof(buf), "find %c >= %d\n", 'A' + i + 1, mo.destdeg[i]); put(buf);
    put(" connecting to");
    for (int j = 0; j <= i; j++) if (mo.list[i] & (1 << j)) { snprintf(buf, sizeof(buf), " %c", 'A' + j); put(buf); }
    if (mo.list[i] != ((2 << i) - 1)) {
      put("\n skipping over");
      for (int j = 0; j <= i; j++) if (~mo.list[i] & (1 << j)) { snprintf(buf, sizeof(buf), " %c", 'A' + j); put(buf); }
    }
    put("\n\n");
  }
  int anchors = 0;
  for (int i = 0; i < mo.anchorsets; i++) anchors |= mo.anch[i];
  put("anchors:");
  for (int i = 0; i < mo.nodes; i++) if (anchors & (1 << i)) { snprintf(buf, sizeof(buf), " %c", 'A' + i); put(buf); }
  put("\n\n");
  put("anchor sets\n");
  for (int i = 0; i < mo.anchorsets; i++) {
    for (int j = 0; j < kMaxCore; j++) if (mo.anch[i] & (1 << j)) { snprintf(buf, sizeof(buf), " %c", 'A' + j); put(buf); }
    snprintf(buf, sizeof(buf), ": %d\n", mo.fcnt[i]); put(buf);
  }
  put("\n");
}

}  // namespace fsgc
