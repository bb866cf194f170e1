// Deterministic synthetic inputs: Erdos-Renyi, R-MAT, Chung-Lu graphs and random motifs.
//
// RNG: splitmix64 (Steele, Lea, Flood 2014; the public-domain reference constants of
// Vigna's splitmix64.c). A uniform double in [0, 1) is the top 53 bits of one draw.
// Parallel generators give item i the contiguous segment of one splitmix64 sequence
// that a sequential generator would use, so outputs do not depend on thread count.
#include "fsgc_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <omp.h>

namespace fsgc {

static constexpr u64 kGamma = 0x9E3779B97F4A7C15ULL;

static inline u64 mix64(u64 z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

u64 splitmix64_next(u64* state) { *state += kGamma; return mix64(*state); }

static inline double unit(u64 x) { return (double)(x >> 11) * 0x1.0p-53; }

// Builds a symmetric CSR with sorted, deduplicated lists from directed pairs
// (src << 32 | dst), both directions already present; self loops are dropped.
int csr_from_pairs(int32_t n, std::vector<u64>& keys, HostGraph* g)
{
  const size_t k = keys.size();
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  const int nt = omp_get_max_threads();
  std::vector<std::vector<int64_t>> tc(nt, std::vector<int64_t>((size_t)n, 0));
  #pragma omp parallel
  {
    std::vector<int64_t>& c = tc[omp_get_thread_num()];
    #pragma omp for schedule(static)
    for (size_t i = 0; i < k; i++) c[keys[i] >> 32]++;
  }
  // per-thread bucket offsets, thread-major within each source, so the fill is stable
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int32_t u = 0; u < n; u++) {
    int64_t s = 0;
    for (int t = 0; t < nt; t++) s += tc[t][u];
    off[u + 1] = off[u] + s;
  }
  #pragma omp parallel for schedule(static)
  for (int32_t u = 0; u < n; u++) {
    int64_t p = off[u];
    for (int t = 0; t < nt; t++) { const int64_t c = tc[t][u]; tc[t][u] = p; p += c; }
  }
  std::vector<int32_t> dst((size_t)k);
  #pragma omp parallel
  {
    std::vector<int64_t>& c = tc[omp_get_thread_num()];
    #pragma omp for schedule(static)
    for (size_t i = 0; i < k; i++) dst[c[keys[i] >> 32]++] = (int32_t)(keys[i] & 0xffffffffu);
  }
  keys.clear();
  keys.shrink_to_fit();
  std::vector<int32_t> deg((size_t)n, 0);
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t u = 0; u < n; u++) {
    int32_t* b = dst.data() + off[u];
    int32_t* e = dst.data() + off[u + 1];
    std::sort(b, e);
    int32_t* w = b;
    for (int32_t* r = b; r < e; r++) {
      if (*r == u) continue;
      if (w > b && w[-1] == *r) continue;
      *w++ = *r;
    }
    deg[u] = (int32_t)(w - b);
  }
  g->n = n;
  g->nindex.assign((size_t)n + 1, 0);
  for (int32_t u = 0; u < n; u++) g->nindex[u + 1] = g->nindex[u] + deg[u];
  g->m = g->nindex[n];
  if (g->m > INT32_MAX) return -75;
  g->nlist.resize((size_t)g->m);
  #pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t u = 0; u < n; u++)
    std::memcpy(g->nlist.data() + g->nindex[u], dst.data() + off[u], sizeof(int32_t) * deg[u]);
  return 0;
}

// G(n, m): m distinct unordered pairs drawn uniformly (rejecting self pairs and repeats)
// from one sequential splitmix64 stream seeded with `seed`; a pair is (a, b) with
// a = draw % n, b = next draw % n.
int gen_er(int32_t n, int64_t m_undirected, u64 seed, HostGraph* g)
{
  if (n < 2 || m_undirected > (int64_t)n * (n - 1) / 2) return -22;
  std::unordered_set<u64> seen;
  seen.reserve((size_t)m_undirected * 2);
  std::vector<u64> keys;
  keys.reserve((size_t)m_undirected * 2);
  u64 s = seed;
  while ((int64_t)seen.size() < m_undirected) {
    const u64 a = splitmix64_next(&s) % (u64)n;
    const u64 b = splitmix64_next(&s) % (u64)n;
    if (a == b) continue;
    const u64 lo = std::min(a, b), hi = std::max(a, b);
    if (!seen.insert((lo << 32) | hi).second) continue;
    keys.push_back((a << 32) | b);
    keys.push_back((b << 32) | a);
  }
  return csr_from_pairs(n, keys, g);
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004) without noise: edge i takes draws
// i*scale .. i*scale+scale-1 of the splitmix64 sequence seeded with `seed`; draw l picks
// the quadrant for bit (scale-1-l) of (src, dst): [0,a) -> (0,0), [a,a+b) -> (0,1),
// [a+b,a+b+c) -> (1,0), else (1,1). Self loops dropped, symmetrised, deduplicated.
int gen_rmat(int scale, int edge_factor, double a, double b, double c, u64 seed, HostGraph* g)
{
  if (scale < 1 || scale > 30 || edge_factor < 1) return -22;
  const int32_t n = (int32_t)1 << scale;
  const int64_t E = (int64_t)edge_factor * n;
  std::vector<u64> keys((size_t)E * 2);
  const double ab = a + b, abc = a + b + c;
  #pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < E; i++) {
    u64 st = seed + (u64)i * (u64)scale * kGamma;
    u64 src = 0, dst = 0;
    for (int l = 0; l < scale; l++) {
      const double r = unit(splitmix64_next(&st));
      const u64 sb = r >= ab, db = (r >= a && r < ab) || r >= abc;
      src = (src << 1) | sb;
      dst = (dst << 1) | db;
    }
    keys[2 * i] = (src << 32) | dst;
    keys[2 * i + 1] = (dst << 32) | src;  // self pairs removed in csr_from_pairs
  }
  return csr_from_pairs(n, keys, g);
}

// Chung-Lu with expected degrees w_i = W (i0 / (i + i0))^(1/(gamma-1)), W = max_deg and
// i0 chosen by bisection so the mean of w is mean_deg. Pairs i < j are kept with
// probability min(1, w_i w_j / S), S = sum w, using the O(n + m) skipping sampler of
// Miller and Hagberg (2011) for weights sorted in decreasing order. Vertex i uses the
// splitmix64 sequence seeded with mix64(seed) ^ mix64(i + 1).
int gen_chung_lu(int32_t n, double gamma, double mean_deg, double max_deg, u64 seed, HostGraph* g)
{
  if (n < 2 || gamma <= 1.0 || mean_deg <= 0 || max_deg <= mean_deg) return -22;
  const double alpha = 1.0 / (gamma - 1.0);
  std::vector<double> w((size_t)n);
  auto mean_for = [&](double i0) {
    double s = 0;
    #pragma omp parallel for reduction(+ : s) schedule(static)
    for (int32_t i = 0; i < n; i++) s += max_deg * std::pow(i0 / (i + i0), alpha);
    return s / n;
  };
  double lo = 1e-9, hi = (double)n;
  for (int it = 0; it < 200; it++) {
    const double mid = std::sqrt(lo * hi);
    if (mean_for(mid) < mean_deg) lo = mid; else hi = mid;
  }
  const double i0 = std::sqrt(lo * hi);
  double S = 0;
  for (int32_t i = 0; i < n; i++) { w[i] = max_deg * std::pow(i0 / (i + i0), alpha); S += w[i]; }
  std::vector<std::vector<u64>> parts(omp_get_max_threads());
  const u64 sm = mix64(seed);
  #pragma omp parallel
  {
    std::vector<u64>& out = parts[omp_get_thread_num()];
    #pragma omp for schedule(dynamic, 256)
    for (int32_t u = 0; u < n - 1; u++) {
      u64 st = sm ^ mix64((u64)u + 1);
      int64_t v = u + 1;
      double p = std::min(w[u] * w[v] / S, 1.0);
      while (v < n && p > 0) {
        if (p != 1.0) {
          const double r = unit(splitmix64_next(&st));
          v += (int64_t)std::floor(std::log(r > 0 ? r : 0x1.0p-53) / std::log1p(-p));
        }
        if (v < n) {
          const double q = std::min(w[u] * w[v] / S, 1.0);
          const double r = unit(splitmix64_next(&st));
          if (r < q / p) { out.push_back(((u64)u << 32) | (u64)v); out.push_back(((u64)v << 32) | (u64)u); }
          p = q;
          v++;
        }
      }
    }
  }
  size_t total = 0;
  for (auto& p : parts) total += p.size();
  std::vector<u64> keys;
  keys.reserve(total);
  for (auto& p : parts) { keys.insert(keys.end(), p.begin(), p.end()); p.clear(); p.shrink_to_fit(); }
  return csr_from_pairs(n, keys, g);
}

// Random motifs: core size uniform in [core_min, core_max], edge density uniform in
// [0.4, 1.0) with each core pair kept independently, resampled until connected; then a
// fringe count uniform in [0, fringe_max], each fringe joined to one uniformly chosen
// core vertex. One sequential splitmix64 stream from `seed`.
void gen_random_motifs(u64 seed, int count, int core_min, int core_max, int fringe_max, std::vector<Pattern>* out)
{
  u64 s = seed;
  out->clear();
  for (int t = 0; t < count; t++) {
    Pattern p;
    const int k = core_min + (int)(splitmix64_next(&s) % (u64)(core_max - core_min + 1));
    const double dens = 0.4 + 0.6 * unit(splitmix64_next(&s));
    for (;;) {
      p = Pattern();
      p.nodes = k;
      for (int i = 0; i < k; i++)
        for (int j = i + 1; j < k; j++)
          if (unit(splitmix64_next(&s)) < dens) p.add_edge(i, j);
      // connected?
      uint32_t reached = 1u, frontier = 1u;
      while (frontier) {
        uint32_t nx = 0;
        for (uint32_t f = frontier; f; f &= f - 1) nx |= p.adj[__builtin_ctz(f)];
        frontier = nx & ~reached;
        reached |= nx;
      }
      if (k == 1 || reached == ((1u << k) - 1)) break;
    }
    const int f = (int)(splitmix64_next(&s) % (u64)(fringe_max + 1));
    for (int i = 0; i < f; i++) {
      const int a = (int)(splitmix64_next(&s) % (u64)k);
      p.add_edge(k + i, a);
    }
    p.nodes = k + f;
    p.name = "random motif " + std::to_string(t);
    out->push_back(p);
  }
}

}  // namespace fsgc
