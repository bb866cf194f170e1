// Motif compiler: splits a pattern into core and fringe vertices and produces the
// 139-byte matching order consumed by the counting path. Behaviour (including every
// tie-break) follows the reference's fringePreprocess.cpp so that the same pattern
// yields the same .mo bytes on all live fields; unused slots are zero here.
#include "fsgc_host.h"

#include <cstdio>
#include <cstring>
#include <deque>

namespace fsgc {

namespace {

enum State : uint8_t { kUndecided = 0, kCore = 1, kFringe = 2 };  // fringes.h:46

struct PatternRow {
  int number;
  const char* name;
  int nodes;
  int nedges;
  int edges[64];
};

// The predefined motifs 0..30 of patterns.h:57-354, as edge lists.
const PatternRow kPatterns[] = {
#include "patterns_table.inc"
};

int lowest_neighbor(const Pattern& p, int v) { return p.adj[v] ? __builtin_ctz(p.adj[v]) : kMaxPattern; }

// fringePreprocess.cpp:157-186 (isConnected): BFS from vertex 0.
bool is_connected(const Pattern& p)
{
  uint32_t reached = 1u, frontier = 1u;
  while (frontier) {
    uint32_t next = 0;
    for (uint32_t f = frontier; f; f &= f - 1) next |= p.adj[__builtin_ctz(f)];
    frontier = next & ~reached;
    reached |= next;
  }
  const uint32_t all = (p.nodes == 32) ? 0xffffffffu : ((1u << p.nodes) - 1);
  return (reached & all) == all;
}

// fringePreprocess.cpp:189-257 (connectCore): while the core is disconnected, BFS from
// the reached core vertices through the rest of the pattern and promote the vertex from
// which the first unreached core vertex is seen. Queue order = ascending indices.
void connect_core(const Pattern& p, State* state, std::string* log)
{
  for (;;) {
    bool reached[kMaxPattern];
    int c = -1;
    for (int u = 0; u < p.nodes; u++) {
      reached[u] = (state[u] != kCore);
      if (state[u] == kCore) c = u;
    }
    if (c < 0) return;
    std::deque<int> q{c};
    while (!q.empty()) {
      const int u = q.front();
      q.pop_front();
      if (reached[u]) continue;
      reached[u] = true;
      for (uint32_t a = p.adj[u]; a; a &= a - 1) {
        const int v = __builtin_ctz(a);
        if (!reached[v]) q.push_back(v);
      }
    }
    bool all = true;
    for (int u = 0; u < p.nodes; u++) all = all && reached[u];
    if (all) return;
    bool visited[kMaxPattern];
    for (int u = 0; u < p.nodes; u++) {
      visited[u] = (state[u] == kCore) && reached[u];
      if (visited[u]) q.push_back(u);
    }
    bool promoted = false;
    while (!q.empty() && !promoted) {
      const int u = q.front();
      q.pop_front();
      for (uint32_t a = p.adj[u]; a; a &= a - 1) {
        const int v = __builtin_ctz(a);
        if (visited[v]) continue;
        visited[v] = true;
        if (!reached[v]) {
          char buf[64];
          snprintf(buf, sizeof(buf), "including %d in core vertices\n", u);
          log->append(buf);
          state[u] = kCore;
          promoted = true;
          break;
        }
        q.push_back(v);
      }
    }
    if (!promoted) return;  // cannot happen for a connected pattern
  }
}

// fringePreprocess.cpp:260-284 (determineCore): for x = 1, 2, ..., scanning vertices
// from the highest index down, turn each undecided degree-x vertex into a fringe and
// its neighbours into core vertices; then connect the core.
int determine_core(const Pattern& p, State* state, std::string* log, std::string* err)
{
  for (int i = 0; i < p.nodes; i++) state[i] = kUndecided;
  for (int x = 1; x < p.nodes; x++) {
    for (int u = p.nodes - 1; u >= 0; u--) {
      if (state[u] != kUndecided || p.degree(u) != x) continue;
      state[u] = kFringe;
      for (uint32_t a = p.adj[u]; a; a &= a - 1) {
        const int v = __builtin_ctz(a);
        if (state[v] == kFringe) { if (err) *err = "internal error: adjacent fringe vertices"; return -22; }
        state[v] = kCore;
      }
    }
  }
  connect_core(p, state, log);
  return 0;
}

// fringePreprocess.cpp:49-128 (genCoreMatchingOrder): greedy order starting at src,
// each step taking the unvisited core vertex with the most visited neighbours, then
// the higher degree, then the smaller lowest neighbour; anchor sets are the bit masks
// (in matching-order positions) of each fringe's core neighbours, sorted ascending.
MatchingOrder matching_order(int src, const Pattern& p, const State* state)
{
  MatchingOrder mo;
  memset(&mo, 0, sizeof(mo));
  int map[kMaxPattern];
  for (int i = 0; i < kMaxPattern; i++) map[i] = 0;
  uint32_t visited = 1u << src;
  map[src] = 0;
  mo.startdeg = (uint8_t)p.degree(src);
  int top = 0;
  while (top < p.nodes - 1) {
    int dst = -1, con = -1;
    for (int v = 0; v < p.nodes; v++) {
      if ((visited >> v) & 1u || state[v] != kCore) continue;
      const int cnt = __builtin_popcount(p.adj[v] & visited);
      bool take = con < cnt;
      if (!take && con == cnt) {
        const int dd = p.degree(dst), dv = p.degree(v);
        take = (dd < dv) || (dd == dv && lowest_neighbor(p, dst) > lowest_neighbor(p, v));
      }
      if (take) { dst = v; con = cnt; }
    }
    if (dst < 0) break;
    if (top >= kMaxCore - 1) break;  // guarded by the caller's core-size check
    map[dst] = top + 1;
    mo.destdeg[top] = (uint8_t)p.degree(dst);
    uint8_t lst = 0;
    for (uint32_t a = p.adj[dst] & visited; a; a &= a - 1) lst |= (uint8_t)(1u << map[__builtin_ctz(a)]);
    mo.list[top] = lst;
    visited |= 1u << dst;
    top++;
  }
  mo.nodes = (uint8_t)(top + 1);
  int as[1 << kMaxCore] = {0};
  for (int v = 0; v < p.nodes; v++) {
    if (state[v] != kFringe) continue;
    int bmp = 0;
    for (uint32_t a = p.adj[v]; a; a &= a - 1) bmp |= 1 << map[__builtin_ctz(a)];
    as[bmp]++;
  }
  int k = 0;
  for (int i = 1; i < (1 << kMaxCore); i++) {
    if (as[i] > 0) { mo.anch[k] = (uint8_t)i; mo.fcnt[k] = (uint8_t)as[i]; k++; }
  }
  mo.anchorsets = (uint8_t)k;
  return mo;
}

}  // namespace

int predefined_pattern(int number, Pattern* p)
{
  for (const PatternRow& r : kPatterns) {
    if (r.number != number) continue;
    *p = Pattern();
    p->name = r.name;
    p->nodes = r.nodes;
    for (int i = 0; i < r.nedges; i++) p->add_edge(r.edges[2 * i], r.edges[2 * i + 1]);
    return 0;
  }
  return -2;
}

// fringePreprocess.cpp:299-317: "n s d s d ..." from the command line.
int pattern_from_edges(int nodes, int nedges, const int* edges, Pattern* p, std::string* err)
{
  if (nodes < 2 || nodes > kMaxPattern) {
    if (err) *err = "ERROR: number_of_nodes must be between 2 and " + std::to_string(kMaxPattern);
    return -22;
  }
  *p = Pattern();
  p->name = "user defined";
  p->nodes = nodes;
  for (int i = 0; i < nedges; i++) {
    const int s = edges[2 * i], d = edges[2 * i + 1];
    if (s < 0 || s >= nodes) { if (err) *err = "ERROR: edge source must be between 0 and " + std::to_string(nodes - 1); return -22; }
    if (d < 0 || d >= nodes) { if (err) *err = "ERROR: edge destination must be between 0 and " + std::to_string(nodes - 1); return -22; }
    if (s == d) { if (err) *err = "ERROR: self loops are not supported in motifs"; return -22; }
    p->add_edge(s, d);
  }
  return 0;
}

// fringePreprocess.cpp:319-360: report, reject disconnected motifs and cores larger
// than 6, start at the (first) highest-degree core vertex, build the order.
int preprocess(const Pattern& p, MatchingOrder* mo, std::string* log, std::string* dot, std::string* err)
{
  std::string local;
  if (!log) log = &local;
  char buf[256];
  snprintf(buf, sizeof(buf), "motif: %s\n", p.name.c_str());
  log->append(buf);
  int m = 0;
  for (int i = 0; i < p.nodes; i++) m += p.degree(i);
  snprintf(buf, sizeof(buf), "%d nodes and %d edges\n", p.nodes, m / 2);
  log->append(buf);
  if (!is_connected(p)) { if (err) *err = "ERROR: motif is disconnected"; return -22; }
  State state[kMaxPattern];
  int rc = determine_core(p, state, log, err);
  if (rc) return rc;
  if (dot) {
    dot->assign("graph {\n");
    for (int i = 0; i < p.nodes; i++) {
      snprintf(buf, sizeof(buf), "  %d [style=filled, fillcolor=\"%s\", label=\"%d\"]\n", i,
               state[i] == kCore ? "#9090ff" : "#ffffff", i);
      dot->append(buf);
    }
    for (int i = 0; i < p.nodes; i++)
      for (uint32_t a = p.adj[i]; a; a &= a - 1) {
        const int j = __builtin_ctz(a);
        if (i < j) { snprintf(buf, sizeof(buf), "  %d -- %d\n", i, j); dot->append(buf); }
      }
    dot->append("}\n");
  }
  int c = 0, src = -1, hideg = -1;
  for (int i = 0; i < p.nodes; i++) {
    if (state[i] != kCore) continue;
    c++;
    if (hideg < p.degree(i)) { hideg = p.degree(i); src = i; }
  }
  snprintf(buf, sizeof(buf), "%d core nodes\n\n", c);
  log->append(buf);
  if (c > kMaxCore) {
    if (err) *err = "ERROR: only motifs with up to " + std::to_string(kMaxCore) + " core nodes supported";
    return -22;
  }
  *mo = matching_order(src, p, state);
  print_matching_order(*mo, log);
  return 0;
}

}  // namespace fsgc
