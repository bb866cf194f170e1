// Oracle for the Fringe-SGC counting path.
//
// TEST INFRASTRUCTURE ONLY. This file is a plain C++/OpenMP restatement of the
// reference's core+fringe counting method (src/fringeCount.cu, "FC" below) and a
// brute-force non-induced subgraph counter. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, and only as the checker or as the timed
// CPU baseline. The product (libfsgc.so) never links or calls it.
//
// Arithmetic is exact: every count is an unsigned 128-bit integer, every product and
// sum is overflow-checked (flag FSGC_OVF128), and the reference's own u64 behaviour is
// reproduced on the side: the reference-compatible value is (S_G mod 2^64) / aut,
// which is what FC computes when none of its nCk intermediates wrap
// (FC:195 `r * (n + 1 - i)` overflowing is reported as FSGC_REF_NCK_INEXACT).
//
// Parity pins (what checks each function; see tests/test_oracle.py):
//   ncr / fc3 / functionx     brute-force enumeration of disjoint fringe placements
//   vertex/edge/wedge/triangle/general classes   brute-force monomorphism counts,
//                             networkx golden counts, K_n closed forms,
//                             count(H in H) == 1, relabel invariance
//   decode of .mo             reference fringePreprocess outputs (tests/golden/mo_golden.json)

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#include <omp.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint8_t byte;

enum : uint32_t {
  FSGC_OVF64 = 1u << 0,            // exact count >= 2^64
  FSGC_OVF128 = 1u << 1,           // some exact intermediate >= 2^128 (count invalid)
  FSGC_REF_NCK_INEXACT = 1u << 2,  // the reference's u64 nCk would have wrapped
  FSGC_AUT_REMAINDER = 1u << 3,    // S_G not divisible by the automorphism sum
  FSGC_REF_AUT_TRUNC = 1u << 4,    // the reference's `int automorphisms` (FC:2213) truncates
};

// fringes.h:50-58 — the 139-byte matching order, read field by field.
struct MO {
  int nodes, anchorsets, startdeg;
  int destdeg[5], list[5];
  int anch[63], fcnt[63];
};

static MO decode_mo(const byte* b)
{
  MO mo;
  memset(&mo, 0, sizeof(mo));
  mo.nodes = b[0];
  mo.anchorsets = b[1];
  mo.startdeg = b[2];
  for (int i = 0; i < 5; i++) { mo.destdeg[i] = b[3 + i]; mo.list[i] = b[8 + i]; }
  for (int i = 0; i < 63; i++) { mo.anch[i] = b[13 + i]; mo.fcnt[i] = b[76 + i]; }
  return mo;
}

struct G {
  int n;
  long long m;
  const int* idx;
  const int* adj;
  int deg(int u) const { return idx[u + 1] - idx[u]; }
  int beg(int u) const { return idx[u]; }
  int end(int u) const { return idx[u + 1]; }
};

// ---------------------------------------------------------------- exact arithmetic

struct Acc {
  u128 sum = 0;
  uint32_t flags = 0;
  u64 matches = 0;
  void add(u128 x) { u128 s = sum + x; if (s < sum) flags |= FSGC_OVF128; sum = s; }
};

static inline u128 mul(u128 a, u128 b, uint32_t& fl)
{
  u128 r;
  if (__builtin_mul_overflow(a, b, &r)) fl |= FSGC_OVF128;
  return r;
}

static inline u128 add(u128 a, u128 b, uint32_t& fl)
{
  u128 r = a + b;
  if (r < a) fl |= FSGC_OVF128;
  return r;
}

// FC:189-197. nCk(n, k): k = min(k, n - k); k < 0 -> 0 (also for n < 0, k == 0);
// k == 0 -> 1; else C(n, k). Exact here. The reference iterates r = r * (n + 1 - i) / i
// in u64; its intermediate at step i equals i * C(n, i), so it wraps iff
// k * C(n, k) >= 2^64 (the largest step), which sets FSGC_REF_NCK_INEXACT.
static u128 ncr(int n, int k, uint32_t& fl)
{
  k = std::min(k, n - k);
  if (k < 0) return 0;
  if (k == 0) return 1;
  u128 r = 1;
  for (int i = 1; i <= k; i++) {
    r = mul(r, (u128)(unsigned)(n - k + i), fl);
    r /= (unsigned)i;
  }
  u128 top;
  if (__builtin_mul_overflow(r, (u128)(unsigned)k, &top) || (top >> 64) != 0) fl |= FSGC_REF_NCK_INEXACT;
  return r;
}

// FC:200-212 — membership in the sorted half-open range [l, r) of adj.
static inline bool contains(const G& g, int v, int l, int r)
{
  if (l >= r) return false;
  const int* p = std::lower_bound(g.adj + l, g.adj + r, v);
  return (p != g.adj + r) && (*p == v);
}

static inline bool has_edge(const G& g, int a, int b) { return contains(g, b, g.beg(a), g.end(a)); }

static int common(const G& g, int a, int b)
{
  int i = g.beg(a), ie = g.end(a), j = g.beg(b), je = g.end(b), c = 0;
  while (i < ie && j < je) {
    const int x = g.adj[i], y = g.adj[j];
    if (x < y) i++; else if (x > y) j++; else { c++; i++; j++; }
  }
  return c;
}

// Venn-region indices FC:61-67.
enum { iu = 1, iv = 2, iw = 4, iuv = 3, iuw = 5, ivw = 6, iuvw = 7 };

// FC:233-327 — number of ways to place the fringes of the 7 anchor sets of a 3-vertex
// core into the 7 Venn regions (disjointly), with the reference's loop order
// (uvw, uv, uw, vw, then u split over uvw/uv/uw/u, v and w merged by Vandermonde).
static u128 fc3(const int fr[8], const int venn[8], uint32_t& fl)
{
  const int fr_uvw = fr[iuvw], fr_uv = fr[iuv], fr_uw = fr[iuw], fr_vw = fr[ivw];
  const int fr_u = fr[iu], fr_v = fr[iv], fr_w = fr[iw];
  int vd_uvw = venn[iuvw], vd_uv = venn[iuv], vd_uw = venn[iuw], vd_vw = venn[ivw];
  const int vd_u = venn[iu], vd_v = venn[iv], vd_w = venn[iw];
  u128 count = 0;
  if (vd_uvw < fr_uvw) return 0;
  const u128 cuvw = ncr(vd_uvw, fr_uvw, fl);
  vd_uvw -= fr_uvw;
  for (int i = std::min(vd_uvw, fr_uv); i >= 0; i--) {
    if (vd_uv < fr_uv - i) break;
    const u128 c1 = mul(mul(cuvw, ncr(vd_uvw, i, fl), fl), ncr(vd_uv, fr_uv - i, fl), fl);
    vd_uvw -= i; vd_uv -= fr_uv - i;
    for (int j = std::min(vd_uvw, fr_uw); j >= 0; j--) {
      if (vd_uw < fr_uw - j) break;
      const u128 c2 = mul(mul(c1, ncr(vd_uvw, j, fl), fl), ncr(vd_uw, fr_uw - j, fl), fl);
      vd_uvw -= j; vd_uw -= fr_uw - j;
      for (int k = std::min(vd_uvw, fr_vw); k >= 0; k--) {
        if (vd_vw < fr_vw - k) break;
        const u128 c3 = mul(mul(c2, ncr(vd_uvw, k, fl), fl), ncr(vd_vw, fr_vw - k, fl), fl);
        vd_uvw -= k; vd_vw -= fr_vw - k;
        for (int a = std::min(vd_uvw, fr_u); a >= 0; a--) {
          const u128 ta = mul(c3, ncr(vd_uvw, a, fl), fl);
          vd_uvw -= a;
          int vd_uvw_vw = vd_uvw + vd_vw;
          for (int b = std::min(vd_uv, fr_u - a); b >= 0; b--) {
            const u128 tb = mul(ta, ncr(vd_uv, b, fl), fl);
            vd_uv -= b;
            for (int c = std::min(vd_uw, fr_u - a - b); c >= 0; c--) {
              if (vd_u < fr_u - a - b - c) break;
              const u128 tc = mul(mul(tb, ncr(vd_uw, c, fl), fl), ncr(vd_u, fr_u - a - b - c, fl), fl);
              vd_uw -= c;
              for (int d = std::min(vd_uvw_vw, fr_v); d >= 0; d--) {
                if (vd_uv + vd_v < fr_v - d) break;
                const u128 td = mul(mul(tc, ncr(vd_uvw_vw, d, fl), fl), ncr(vd_uv + vd_v, fr_v - d, fl), fl);
                vd_uvw_vw -= d;
                if (vd_uvw_vw + vd_uw + vd_w >= fr_w) {
                  count = add(count, mul(td, ncr(vd_uvw_vw + vd_uw + vd_w, fr_w, fl), fl), fl);
                }
                vd_uvw_vw += d;
              }
              vd_uw += c;
            }
            vd_uv += b;
          }
          vd_uvw += a;
        }
        vd_uvw += k; vd_vw += fr_vw - k;
      }
      vd_uvw += j; vd_uw += fr_uw - j;
    }
    vd_uvw += i; vd_uv += fr_uv - i;
  }
  return count;
}

// ---------------------------------------------------------------- vertex core

// FC:2152-2191 (vertexCore, k_stars)
static void vertex_core(const MO& mo, const G& g, Acc& acc, int smod, int srem)
{
  if (mo.anchorsets == 0) { acc.add((u128)g.n); return; }
  if (mo.fcnt[0] == 1) { acc.add((u128)(g.m / 2)); return; }
  const int t = mo.fcnt[0];
  #pragma omp parallel
  {
    Acc loc;
    #pragma omp for schedule(dynamic, 1024) nowait
    for (int u = 0; u < g.n; u++) {
      if (smod > 1 && (u % smod) != srem) continue;
      const int d = g.deg(u);
      if (d >= t) { loc.add(ncr(d, t, loc.flags)); loc.matches++; }
    }
    #pragma omp critical
    { acc.add(loc.sum); acc.flags |= loc.flags; acc.matches += loc.matches; }
  }
}

// ---------------------------------------------------------------- edge core

// FC:1999-2051 closed form for one edge (general_edge_core).
static u128 edge_formula(int du, int dv, int c_uv, int f_u, int f_uv, int f_v, uint32_t& fl)
{
  const int a_uv = c_uv - f_uv;
  if (a_uv < 0) return 0;
  const int c_u = du - 1 - c_uv, c_v = dv - 1 - c_uv;
  const int bound = std::min(f_u, a_uv);
  u128 acc = 0;
  for (int j = 0; j <= bound; j++) {
    const int k = f_u - j;
    u128 term = mul(ncr(c_u, k, fl), ncr(a_uv + c_v - j, f_v, fl), fl);
    if (f_u != f_v) term = add(term, mul(ncr(c_v, k, fl), ncr(a_uv + c_u - j, f_v, fl), fl), fl);
    acc = add(acc, mul(term, ncr(a_uv, j, fl), fl), fl);
  }
  return mul(acc, ncr(c_uv, f_uv, fl), fl);
}

// FC:2106-2144 (edgeCore) with FC:2054-2103 (init_wl_edge), FC:1864-1894 (triangle),
// FC:1931-1966 (tailed_triangle), FC:1999-2051 (general_edge_core).
static void edge_core(const MO& mo, const G& g, Acc& acc, int smod, int srem)
{
  int f[4] = {0, 0, 0, 0};
  for (int i = 0; i < mo.anchorsets; i++) f[mo.anch[i] & 3] = mo.fcnt[i];
  const int f_u = std::min(f[1], f[2]), f_uv = f[3], f_v = std::max(f[1], f[2]);
  const bool tri = (f_uv == 1) && ((f_u == 0) || (f_v == 0));
  const int lim1 = tri ? 1 : f_u + f_uv, lim2 = tri ? 1 : f_u + f_uv + f_v;
  #pragma omp parallel
  {
    Acc loc;
    #pragma omp for schedule(dynamic, 64) nowait
    for (int u = 0; u < g.n; u++) {
      if (smod > 1 && (u % smod) != srem) continue;
      const int du = g.deg(u);
      if (du <= lim1) continue;
      for (int e = g.beg(u); e < g.end(u); e++) {
        const int v = g.adj[e];
        if (v >= u) break;
        const int dv = g.deg(v);
        if (!((dv > lim1) && (du + dv > lim2))) continue;
        if (tri) {
          const u128 c12 = add(ncr(du - 2, f_v, loc.flags), ncr(dv - 2, f_v, loc.flags), loc.flags);
          int i = g.beg(u), ie = g.end(u), j = g.beg(v), je = g.end(v);
          while (i < ie && j < je) {
            const int x = g.adj[i], y = g.adj[j];
            if (x >= v || y >= v) break;
            if (x < y) i++; else if (x > y) j++; else {
              loc.matches++;
              if (f_u == 0 && f_v == 0) loc.add(1);
              else loc.add(add(c12, ncr(g.deg(x) - 2, f_v, loc.flags), loc.flags));
              i++; j++;
            }
          }
        } else {
          loc.matches++;
          loc.add(edge_formula(du, dv, common(g, u, v), f_u, f_uv, f_v, loc.flags));
        }
      }
    }
    #pragma omp critical
    { acc.add(loc.sum); acc.flags |= loc.flags; acc.matches += loc.matches; }
  }
}

// ---------------------------------------------------------------- rotations

// FC:1291-1317 (triangleRotations)
static int triangle_rotations(const int f[8], int fr[6][8])
{
  int p = 0;
  auto put = [&](int r, int a, int b, int c, int d, int e, int ff, int gg) {
    fr[r][iuvw] = a; fr[r][iuv] = b; fr[r][iuw] = c; fr[r][ivw] = d; fr[r][iu] = e; fr[r][iv] = ff; fr[r][iw] = gg; fr[r][0] = 0;
  };
  put(p++, f[iuvw], f[iuv], f[iuw], f[ivw], f[iu], f[iv], f[iw]);
  if ((f[iu] != f[iv]) || (f[iu] != f[iw]) || (f[iuv] != f[iuw]) || (f[iuv] != f[ivw])) {
    int q = p - 1;
    put(p++, fr[q][iuvw], fr[q][ivw], fr[q][iuv], fr[q][iuw], fr[q][iv], fr[q][iw], fr[q][iu]);
    put(p++, fr[q][iuvw], fr[q][iuw], fr[q][ivw], fr[q][iuv], fr[q][iw], fr[q][iu], fr[q][iv]);
    if (((f[iv] != f[iw]) || (f[iuv] != f[iuw])) && ((f[iu] != f[iw]) || (f[iuv] != f[ivw])) && ((f[iu] != f[iv]) || (f[iuw] != f[ivw]))) {
      put(p++, f[iuvw], f[iuw], f[iuv], f[ivw], f[iu], f[iw], f[iv]);
      q = p - 1;
      put(p++, fr[q][iuvw], fr[q][ivw], fr[q][iuv], fr[q][iuw], fr[q][iv], fr[q][iw], fr[q][iu]);
      put(p++, fr[q][iuvw], fr[q][iuw], fr[q][ivw], fr[q][iuv], fr[q][iw], fr[q][iu], fr[q][iv]);
    }
  }
  return p;
}

// FC:1754-1780 (wedgeRotations)
static int wedge_rotations(const int f[8], int fr[6][8])
{
  int p = 0;
  auto put = [&](int r, int a, int b, int c, int d, int e, int ff, int gg) {
    fr[r][iuvw] = a; fr[r][iuv] = b; fr[r][iuw] = c; fr[r][ivw] = d; fr[r][iu] = e; fr[r][iv] = ff; fr[r][iw] = gg; fr[r][0] = 0;
  };
  put(p++, f[iuvw], f[iuv], f[iuw], f[ivw], f[iu], f[iv], f[iw]);
  int q = p - 1;
  put(p++, fr[q][iuvw], fr[q][ivw], fr[q][iuv], fr[q][iuw], fr[q][iv], fr[q][iw], fr[q][iu]);
  put(p++, fr[q][iuvw], fr[q][iuw], fr[q][ivw], fr[q][iuv], fr[q][iw], fr[q][iu], fr[q][iv]);
  if ((f[iv] != f[iw]) || (f[iuv] != f[iuw])) {
    put(p++, f[iuvw], f[iuw], f[iuv], f[ivw], f[iu], f[iw], f[iv]);
    q = p - 1;
    put(p++, fr[q][iuvw], fr[q][ivw], fr[q][iuv], fr[q][iuw], fr[q][iv], fr[q][iw], fr[q][iu]);
    put(p++, fr[q][iuvw], fr[q][iuw], fr[q][ivw], fr[q][iuv], fr[q][iw], fr[q][iu], fr[q][iv]);
  }
  return p;
}

// Seven Venn regions of N(a), N(b), N(c) by a walk over N(c) (FC:934-964 / FC:1480-1510),
// before removing the core vertices themselves.
static void venn3_raw(const G& g, int a, int b, int c, int p_ab, int raw[8])
{
  int t13 = 0, t23 = 0, t123 = 0;
  for (int e = g.beg(c); e < g.end(c); e++) {
    const int q = g.adj[e];
    const bool b13 = contains(g, q, g.beg(a), g.end(a));
    const bool b23 = contains(g, q, g.beg(b), g.end(b));
    t13 += b13; t23 += b23; t123 += (b13 && b23);
  }
  raw[iuvw] = t123;
  raw[iuv] = p_ab;      // |N(a) & N(b)|, not yet split
  raw[iuw] = t13;
  raw[ivw] = t23;
  raw[0] = 0;
}

// ---------------------------------------------------------------- wedge core

// FC:1381-1406 (wedge): endpoints (a, b) of a wedge whose centre has no fringes.
static u128 wedge_formula(const G& g, int a, int b, int f_u, int f_uv, int f_v, int c_uv, uint32_t& fl)
{
  const int a_uv = c_uv - f_uv;
  if (a_uv < 0) return 0;
  const int corr = has_edge(g, b, a) ? 2 : 1;
  const int a_u = g.deg(a) - (corr + c_uv), a_v = g.deg(b) - (corr + c_uv);
  const int bound = std::min(f_u, a_uv);
  u128 sum = 0;
  for (int j = 0; j <= bound; j++) {
    const int k = f_u - j;
    u128 term = mul(ncr(a_u, k, fl), ncr(a_uv + a_v - j, f_v, fl), fl);
    if (f_u != f_v) term = add(term, mul(ncr(a_v, k, fl), ncr(a_uv + a_u - j, f_v, fl), fl), fl);
    sum = add(sum, mul(term, ncr(a_uv, j, fl), fl), fl);
  }
  return mul(sum, ncr(c_uv, f_uv, fl), fl);
}

// FC:1783-1825 (wedgeCore), FC:1409-1477 (wedgecore_specialized), FC:1513-1650
// (general_wedge_core), FC:1653-1751 (init_wl_specialized_wedge, init_wl_wedge).
static void wedge_core(const MO& mo, const G& g, Acc& acc, int smod, int srem)
{
  int f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < mo.anchorsets; i++) f[mo.anch[i] & 7] = mo.fcnt[i];
  if (mo.list[1] == 2) { std::swap(f[iu], f[iv]); std::swap(f[iuw], f[ivw]); }
  const int min_u = f[iuvw] + f[iuv] + f[iuw] + f[iu] + 2;
  const int min_v = f[iuvw] + f[iuv] + f[ivw] + f[iv] + 1;
  const int min_w = f[iuvw] + f[iuw] + f[ivw] + f[iw] + 1;
  const int sum3 = min_u + min_v + min_w;
  const int sum2 = sum3 - std::max(std::max(min_u, min_v), min_w);
  const int sum1 = std::min(std::min(min_u, min_v), min_w);
  const int lim = std::min(min_v, min_w);
  const bool special = (f[iuvw] == 0) && (f[iuv] == 0) && (f[iuw] == 0) && (f[iu] == 0);
  int fr[6][8];
  const int rot = special ? 0 : wedge_rotations(f, fr);

  #pragma omp parallel
  {
    Acc loc;
    #pragma omp for schedule(dynamic, 16) nowait
    for (int u = 0; u < g.n; u++) {
      if (smod > 1 && (u % smod) != srem) continue;
      const int du = g.deg(u);
      if (special) {
        if (du < 2) continue;
        const int smin_v = f[iv] + f[ivw] + 1, smin_w = f[ivw] + f[iw] + 1;
        const int slim = std::min(smin_v, smin_w);
        for (int e = g.beg(u); e < g.end(u); e++) {
          const int v = g.adj[e];
          const int dv = g.deg(v);
          if (dv < lim) continue;  // init_wl_specialized_wedge uses the outer lim (FC:1808)
          const int thres = std::max(slim, smin_v + smin_w - dv);
          for (int e2 = g.beg(u); e2 < g.end(u); e2++) {
            const int w = g.adj[e2];
            if (w >= v) break;
            if (g.deg(w) < thres) continue;
            loc.matches++;
            const int c_uv = common(g, v, w) - 1;
            loc.add(wedge_formula(g, v, w, f[iv], f[ivw], f[iw], c_uv, loc.flags));
          }
        }
        continue;
      }
      if (du < sum1) continue;
      const int thres1w = std::max(sum1, sum2 - du);
      for (int e = g.beg(u); e < g.end(u); e++) {
        const int v = g.adj[e];
        const int dv = g.deg(v);
        if (dv < thres1w) continue;  // init_wl_wedge
        const int p12 = common(g, u, v);
        const bool cond = (du >= min_u) && (dv >= lim);
        const int thres1 = std::max(lim, sum2 - dv);
        const int thres2 = std::max(std::max(sum1, sum3 - du - dv), std::max(sum2 - du, sum2 - dv));
        for (int e2 = g.beg(u); e2 < g.end(u); e2++) {
          const int w = g.adj[e2];
          if (w >= v) break;
          const int dw = g.deg(w);
          if (dw < thres2) continue;
          const bool tri = has_edge(g, v, w);
          if (tri ? !(u < w) : !(cond && dw >= thres1)) continue;
          loc.matches++;
          int venn[8];
          venn3_raw(g, u, v, w, p12, venn);
          const int c123 = venn[iuvw];
          const int c12 = p12 - c123, c13 = venn[iuw] - c123, c23 = venn[ivw] - c123;
          venn[iu] = du - (p12 + c13);
          venn[iv] = dv - (p12 + c23);
          venn[iw] = dw - (c13 + c23 + c123);
          venn[iuv] = c12; venn[iuw] = c13; venn[ivw] = c23; venn[iuvw] = c123;
          if (!tri) {
            venn[iu] -= 2; venn[ivw] -= 1;  // FC:1598-1602
            loc.add(fc3(fr[0], venn, loc.flags));
            if (rot > 3) loc.add(fc3(fr[3], venn, loc.flags));
          } else {
            venn[iuv] -= 1; venn[iuw] -= 1; venn[ivw] -= 1;  // FC:1631-1636
            for (int r = 0; r < rot; r++) loc.add(fc3(fr[r], venn, loc.flags));
          }
        }
      }
    }
    #pragma omp critical
    { acc.add(loc.sum); acc.flags |= loc.flags; acc.matches += loc.matches; }
  }
}

// ---------------------------------------------------------------- triangle core

// FC:1320-1373 (triangleCore), FC:1167-1218 (init_wl_triangle), FC:774-931 (clique,
// tailed_clique_triangle), FC:934-1062 (com3, common3_triangle, general_triangle_core).
// The degree relabel of FC:1342-1352 is an isomorphism and does not change any sum.
static void triangle_core(const MO& mo, const G& g, Acc& acc, int smod, int srem)
{
  int f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < mo.anchorsets; i++) f[mo.anch[i] & 7] = mo.fcnt[i];
  const int min_u = f[iuvw] + f[iuv] + f[iuw] + f[iu];
  const int min_v = f[iuvw] + f[iuv] + f[ivw] + f[iv];
  const int min_w = f[iuvw] + f[iuw] + f[ivw] + f[iw];
  const int sum3 = min_u + min_v + min_w + 6;
  const int sum2 = sum3 - (std::max(std::max(min_u, min_v), min_w) + 2);
  const int sum1 = std::min(std::min(min_u, min_v), min_w) + 2;
  int fr[6][8];
  const int rot = triangle_rotations(f, fr);
  const bool cliq = (f[iuvw] == 1) && (f[iuv] + f[iuw] + f[ivw] == 0) && ((f[iu] == 0) + (f[iv] == 0) + (f[iw] == 0) >= 2);
  const int tails = f[iu] + f[iv] + f[iw];

  #pragma omp parallel
  {
    Acc loc;
    std::vector<int> com;
    #pragma omp for schedule(dynamic, 16) nowait
    for (int u = 0; u < g.n; u++) {
      if (smod > 1 && (u % smod) != srem) continue;
      const int du = g.deg(u);
      if (du < sum1) continue;
      const int thres1 = std::max(sum1, sum2 - du);
      for (int e = g.beg(u); e < g.end(u); e++) {
        const int v = g.adj[e];
        if (v >= u) break;
        const int dv = g.deg(v);
        if (dv < thres1) continue;
        // common neighbours w of u and v with w < v
        com.clear();
        int p12 = 0;
        {
          int i = g.beg(u), ie = g.end(u), j = g.beg(v), je = g.end(v);
          while (i < ie && j < je) {
            const int x = g.adj[i], y = g.adj[j];
            if (x < y) i++; else if (x > y) j++; else { p12++; if (x < v) com.push_back(x); i++; j++; }
          }
        }
        if (cliq) {
          const u128 c12 = add(ncr(du - 3, tails, loc.flags), ncr(dv - 3, tails, loc.flags), loc.flags);
          for (int w : com) {
            const int dw = g.deg(w);
            if (dw < 3) continue;
            loc.matches++;
            const u128 c123 = add(c12, ncr(dw - 3, tails, loc.flags), loc.flags);
            for (int q : com) {
              if (q >= w) break;
              if (!has_edge(g, w, q)) continue;
              if (tails == 0) loc.add(1);
              else loc.add(add(c123, ncr(g.deg(q) - 3, tails, loc.flags), loc.flags));
            }
          }
          continue;
        }
        const int thres2 = std::max(std::max(sum1, sum3 - du - dv), std::max(sum2 - du, sum2 - dv));
        for (int w : com) {
          const int dw = g.deg(w);
          if (dw < thres2) continue;
          loc.matches++;
          int venn[8];
          venn3_raw(g, u, v, w, p12, venn);
          const int c123 = venn[iuvw];
          const int c12 = p12 - (c123 + 1);
          const int c13 = venn[iuw] - (c123 + 1);
          const int c23 = venn[ivw] - (c123 + 1);
          venn[iu] = du - (c12 + c13 + (c123 + 2));
          venn[iv] = dv - (c12 + c23 + (c123 + 2));
          venn[iw] = dw - (c13 + c23 + (c123 + 2));
          venn[iuv] = c12; venn[iuw] = c13; venn[ivw] = c23; venn[iuvw] = c123;
          for (int r = 0; r < rot; r++) loc.add(fc3(fr[r], venn, loc.flags));
        }
      }
    }
    #pragma omp critical
    { acc.add(loc.sum); acc.flags |= loc.flags; acc.matches += loc.matches; }
  }
}

// ---------------------------------------------------------------- general core

// FC:408-486 (functionX), restated as a recursion over (anchor set, superset region).
// Anchor sets are processed from the last to the first; the fringes of anchor set S are
// split over the regions R that contain S (within the anchor universe `top`), in the
// order of the superset walk idx -> ((idx | nonanchors) + 1 | S) & ~nonanchors, with
// the last region (top) taking the remainder. The reference's repeat of its first
// level when anch[last] == top and anchorsets > 1 (FC:428 sets `a` without the top
// test of FC:480) multiplies every sum by (fcnt[last] + 1); that constant factor is
// applied by the caller to S_G and to the automorphism sum alike.
struct FX {
  const MO* mo;
  int top, nonanch;
  uint32_t* fl;
  int venn[64];
  int fr[64];
  u128 rec(int asi, int idx, u128 prod)
  {
    const int anch = mo->anch[asi];
    if (idx == top) {
      const int need = fr[anch];
      const u128 p = mul(prod, ncr(venn[idx], need, *fl), *fl);
      if (p == 0) return 0;
      if (asi == 0) return p;
      const int sv = venn[idx], sf = fr[anch];
      venn[idx] -= need; fr[anch] = 0;
      const u128 r = rec(asi - 1, mo->anch[asi - 1], p);
      venn[idx] = sv; fr[anch] = sf;
      return r;
    }
    u128 sum = 0;
    const int nidx = (((idx | nonanch) + 1) | anch) & ~nonanch;
    for (int a = std::min(venn[idx], fr[anch]); a >= 0; a--) {
      const u128 p = mul(prod, ncr(venn[idx], a, *fl), *fl);
      if (p == 0) continue;
      venn[idx] -= a; fr[anch] -= a;
      sum = add(sum, rec(asi, nidx, p), *fl);
      venn[idx] += a; fr[anch] += a;
    }
    return sum;
  }
};

// FC:345-405 (commonX): Venn regions over the anchor core vertices, keyed by the exact
// set of anchors a non-core vertex is adjacent to.
static void commonx(const G& g, int nodes, const int* stack, int nonanch, int venn[64])
{
  for (int i = 0; i < 64; i++) venn[i] = 0;
  // Every vertex adjacent to at least one anchor: compute its anchor set directly.
  // Walk each anchor's list; count x at its lowest adjacent anchor only.
  for (int i = 0; i < nodes; i++) {
    if ((1 << i) & nonanch) continue;
    const int a = stack[i];
    for (int e = g.beg(a); e < g.end(a); e++) {
      const int x = g.adj[e];
      int bmp = 1 << i;
      bool first = true;
      for (int k = 0; k < nodes; k++) {
        if (k == i || ((1 << k) & nonanch)) continue;
        if (has_edge(g, stack[k], x)) { bmp |= 1 << k; if (k < i) first = false; }
      }
      if (first) venn[bmp]++;
    }
  }
  // core vertices are not fringe candidates (FC:391-403)
  for (int c = 0; c < nodes; c++) {
    const int x = stack[c];
    int bmp = 0;
    for (int k = 0; k < nodes; k++) {
      if ((1 << k) & nonanch) continue;
      if (has_edge(g, stack[k], x)) bmp |= 1 << k;
    }
    if (bmp) venn[bmp]--;
  }
}

// FC:489-564 (general_core_level<L>): candidates for core vertex L+1 are the
// neighbours of the first vertex in list[L]; degree >= destdeg[L]; distinct from the
// earlier vertices it need not be adjacent to; adjacent to all vertices in list[L].
static void general_level(const MO& mo, const G& g, int level, int* stack, int nonanch, int top, Acc& loc)
{
  const int list = mo.list[level];
  const int first = __builtin_ctz(list);
  const int wn = stack[first];
  for (int e = g.beg(wn); e < g.end(wn); e++) {
    const int x = g.adj[e];
    if (g.deg(x) < mo.destdeg[level]) continue;
    bool ok = true;
    int skip = list ^ ((2 << level) - 1);
    while (skip) { const int z = stack[__builtin_ctz(skip)]; skip &= skip - 1; if (x == z) { ok = false; break; } }
    if (!ok) continue;
    int need = list & (list - 1);
    while (need) { const int z = stack[__builtin_ctz(need)]; need &= need - 1; if (!has_edge(g, z, x)) { ok = false; break; } }
    if (!ok) continue;
    stack[level + 1] = x;
    if (level == mo.nodes - 2) {
      loc.matches++;
      FX fx;
      fx.mo = &mo; fx.top = top; fx.nonanch = nonanch; fx.fl = &loc.flags;
      commonx(g, mo.nodes, stack, nonanch, fx.venn);
      for (int i = 0; i < 64; i++) fx.fr[i] = 0;
      for (int i = 0; i < mo.anchorsets; i++) fx.fr[mo.anch[i]] = mo.fcnt[i];
      const int last = mo.anchorsets - 1;
      loc.add(fx.rec(last, mo.anch[last], 1));
    } else {
      general_level(mo, g, level + 1, stack, nonanch, top, loc);
    }
  }
}

// FC:666-692 (generalCore), FC:617-663 (init_wl_general), FC:567-614 (general_core_level0)
static void general_core(const MO& mo, const G& g, Acc& acc, int smod, int srem)
{
  int anchors = 0;
  for (int i = 0; i < mo.anchorsets; i++) anchors |= mo.anch[i];
  const int nonanch = anchors ^ ((1 << mo.nodes) - 1);
  const int top = ((1 << mo.nodes) - 1) ^ nonanch;
  #pragma omp parallel
  {
    Acc loc;
    int stack[8];
    #pragma omp for schedule(dynamic, 4) nowait
    for (int u = 0; u < g.n; u++) {
      if (smod > 1 && (u % smod) != srem) continue;
      if (g.deg(u) < mo.startdeg) continue;
      for (int e = g.beg(u); e < g.end(u); e++) {
        const int v = g.adj[e];
        if (g.deg(v) < mo.destdeg[0]) continue;
        stack[0] = u; stack[1] = v;
        general_level(mo, g, 1, stack, nonanch, top, loc);
      }
    }
    #pragma omp critical
    { acc.add(loc.sum); acc.flags |= loc.flags; acc.matches += loc.matches; }
  }
}

// ---------------------------------------------------------------- motif graph

// FC:121-186 (matchingOrder2ECLgraph): core vertices 0..nodes-1 in matching order,
// then the fringes of each anchor set in order.
static void mo_graph(const MO& mo, std::vector<int>& idx, std::vector<int>& adj)
{
  int n = mo.nodes;
  for (int i = 0; i < mo.anchorsets; i++) n += mo.fcnt[i];
  std::vector<std::vector<int>> nb(n);
  for (int i = 0; i < mo.nodes - 1; i++)
    for (int j = 0; j <= i; j++)
      if (mo.list[i] & (1 << j)) { nb[i + 1].push_back(j); nb[j].push_back(i + 1); }
  int v = mo.nodes;
  for (int i = 0; i < mo.anchorsets; i++)
    for (int k = 0; k < mo.fcnt[i]; k++, v++)
      for (int j = 0; j < 6; j++)
        if (mo.anch[i] & (1 << j)) { nb[j].push_back(v); nb[v].push_back(j); }
  idx.assign(n + 1, 0);
  adj.clear();
  for (int i = 0; i < n; i++) {
    std::sort(nb[i].begin(), nb[i].end());
    adj.insert(adj.end(), nb[i].begin(), nb[i].end());
    idx[i + 1] = (int)adj.size();
  }
}

static int motif_class(const MO& mo)
{
  if (mo.nodes == 1) return 0;
  if (mo.nodes == 2) return 1;
  if (mo.nodes == 3) return (__builtin_popcount(mo.list[1]) == 1) ? 2 : 3;
  return 4;
}

static bool general_quirk(const MO& mo, int* factor)
{
  int anchors = 0;
  for (int i = 0; i < mo.anchorsets; i++) anchors |= mo.anch[i];
  if (mo.nodes >= 4 && mo.anchorsets > 1 && mo.anch[mo.anchorsets - 1] == anchors) {
    *factor = mo.fcnt[mo.anchorsets - 1] + 1;
    return true;
  }
  *factor = 1;
  return false;
}

static void run_class(const MO& mo, const G& g, Acc& acc, int smod, int srem)
{
  switch (motif_class(mo)) {
    case 0: vertex_core(mo, g, acc, smod, srem); break;
    case 1: edge_core(mo, g, acc, smod, srem); break;
    case 2: wedge_core(mo, g, acc, smod, srem); break;
    case 3: triangle_core(mo, g, acc, smod, srem); break;
    default: general_core(mo, g, acc, smod, srem); break;
  }
}

// ---------------------------------------------------------------- brute force

// Non-induced occurrences of H in G = (# injective homomorphisms H -> G) / |Aut(H)|,
// by backtracking over H's vertices in BFS order. Exact u128.
struct BF {
  int hn;
  std::vector<std::vector<int>> hadj;
  std::vector<int> order, parent;
  const G* g;
  std::vector<int> map;
  std::vector<char> used;
  u128 count;
  void go(int i)
  {
    if (i == hn) { count++; return; }
    const int h = order[i];
    const int p = parent[i];
    auto try_x = [&](int x) {
      if (used[x]) return;
      for (int hb : hadj[h]) {
        const int y = map[hb];
        if (y >= 0 && !has_edge(*g, x, y)) return;
      }
      used[x] = 1; map[h] = x;
      go(i + 1);
      used[x] = 0; map[h] = -1;
    };
    if (p < 0) { for (int x = 0; x < g->n; x++) try_x(x); }
    else { const int px = map[p]; for (int e = g->beg(px); e < g->end(px); e++) try_x(g->adj[e]); }
  }
  u128 run(const G& gg)
  {
    g = &gg; map.assign(hn, -1); used.assign(gg.n, 0); count = 0;
    go(0);
    return count;
  }
  void setup(int n, int m, const int* edges)
  {
    hn = n; hadj.assign(n, {});
    for (int i = 0; i < m; i++) {
      const int a = edges[2 * i], b = edges[2 * i + 1];
      if (a == b) continue;
      if (std::find(hadj[a].begin(), hadj[a].end(), b) == hadj[a].end()) { hadj[a].push_back(b); hadj[b].push_back(a); }
    }
    // BFS order from the highest-degree vertex; parent = earliest mapped neighbour
    std::vector<char> seen(n, 0);
    order.clear(); parent.clear();
    int s = 0;
    for (int i = 1; i < n; i++) if (hadj[i].size() > hadj[s].size()) s = i;
    std::vector<int> pos(n, -1);
    std::vector<int> q{s}; seen[s] = 1;
    for (size_t qi = 0; qi < q.size(); qi++) {
      const int x = q[qi];
      for (int y : hadj[x]) if (!seen[y]) { seen[y] = 1; q.push_back(y); }
    }
    for (int x : q) {
      int p = -1;
      for (int y : hadj[x]) if (pos[y] >= 0 && (p < 0 || pos[y] < pos[p])) p = y;
      pos[x] = (int)order.size();
      order.push_back(x); parent.push_back(p);
    }
  }
};

// ---------------------------------------------------------------- C ABI for tests

extern "C" {

struct oracle_result {
  u64 exact_lo, exact_hi;   // floor(S_G / aut), exact
  u64 ref_compat;           // what the reference prints (as u64): (S_G mod 2^64) / aut
  u64 sg_lo, sg_hi;         // S_G: the class sum on the data graph (before / aut)
  u64 aut;                  // automorphism sum S_H (as the reference prints it)
  u64 core_matches;         // core embeddings for which a fringe formula was evaluated
  uint32_t flags;
  int32_t motif_class;      // 0 vertex, 1 edge, 2 wedge, 3 triangle, 4 general
  double seconds;           // wall time of the counting (S_G only)
  int32_t threads;
};

// Count motif `mo` (139 bytes) in the CSR graph. smod > 1 restricts S_G to root
// vertices u with u % smod == srem (a deterministic slice); exact/ref_compat are then
// not meaningful and only sg_*/core_matches are compared.
int oracle_count(const byte* mob, int n, long long m, const int* idx, const int* adj, int threads,
                 int smod, int srem, oracle_result* out)
{
  if (!mob || !idx || !adj || !out || n < 1) return -22;
  if (threads > 0) omp_set_num_threads(threads);
  memset(out, 0, sizeof(*out));
  const MO mo = decode_mo(mob);
  if (mo.nodes < 1 || mo.nodes > 6) return -22;
  G g{n, m, idx, adj};
  const int cls = motif_class(mo);
  out->motif_class = cls;
  out->threads = omp_get_max_threads();
  u128 aut = 1;
  uint32_t flags = 0;
  int qf = 1;
  if (cls >= 2) {
    std::vector<int> hidx, hadj;
    mo_graph(mo, hidx, hadj);
    const int hn = (int)hidx.size() - 1;
    if (hn > n || (long long)hadj.size() > m) {  // FC:2216-2219
      out->aut = 0;
      return 0;
    }
    G h{hn, (long long)hadj.size(), hidx.data(), hadj.data()};
    Acc ha;
    run_class(mo, h, ha, 1, 0);
    general_quirk(mo, &qf);
    aut = ha.sum * (u128)qf;
    flags |= ha.flags;
    if (aut == 0) return -5;
  }
  const auto t0 = std::chrono::steady_clock::now();
  Acc acc;
  run_class(mo, g, acc, smod, srem);
  const auto t1 = std::chrono::steady_clock::now();
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  flags |= acc.flags;
  const u128 sg = acc.sum * (u128)qf;
  const u128 q = sg / aut;
  if (smod <= 1 && (sg % aut) != 0) flags |= FSGC_AUT_REMAINDER;
  if ((q >> 64) != 0) flags |= FSGC_OVF64;
  out->exact_lo = (u64)q; out->exact_hi = (u64)(q >> 64);
  out->sg_lo = (u64)acc.sum; out->sg_hi = (u64)(acc.sum >> 64);
  out->aut = (u64)aut;
  // FC:2213, 2230: the automorphism sum is stored in an int and converted back for the
  // u64 division, so the reference divides by the sign-extended low 32 bits.
  const u64 ref_aut = (u64)(int64_t)(int32_t)(uint32_t)(u64)aut;
  if ((u128)ref_aut != aut) flags |= FSGC_REF_AUT_TRUNC;
  out->ref_compat = ref_aut ? (u64)sg / ref_aut : 0;
  out->core_matches = acc.matches;
  out->flags = flags;
  return 0;
}

// Brute-force non-induced count of the pattern (hn vertices, hm edges as pairs) in G.
int oracle_bruteforce(int hn, int hm, const int* hedges, int n, long long m, const int* idx, const int* adj,
                      u64* out_lo, u64* out_hi)
{
  BF bf;
  bf.setup(hn, hm, hedges);
  std::vector<int> hidx(hn + 1, 0), hadjv;
  for (int i = 0; i < hn; i++) {
    std::vector<int> s = bf.hadj[i];
    std::sort(s.begin(), s.end());
    hadjv.insert(hadjv.end(), s.begin(), s.end());
    hidx[i + 1] = (int)hadjv.size();
  }
  G h{hn, (long long)hadjv.size(), hidx.data(), hadjv.data()};
  const u128 aut = bf.run(h);
  if (aut == 0) return -5;
  G g{n, m, idx, adj};
  const u128 inj = bf.run(g);
  const u128 q = inj / aut;
  *out_lo = (u64)q; *out_hi = (u64)(q >> 64);
  return (inj % aut) ? 1 : 0;
}

// fc3 on explicit vectors (for the fringe-placement brute-force check).
void oracle_fc3(const int* fr8, const int* venn8, u64* lo, u64* hi, uint32_t* flags)
{
  uint32_t fl = 0;
  const u128 r = fc3(fr8, venn8, fl);
  *lo = (u64)r; *hi = (u64)(r >> 64); *flags = fl;
}

// functionX on explicit vectors: anchor sets (bitmasks, ascending), counts, nonanchor mask.
void oracle_functionx(int nodes, int anchorsets, const int* anch, const int* fcnt, const int* venn64,
                      u64* lo, u64* hi, uint32_t* flags)
{
  MO mo;
  memset(&mo, 0, sizeof(mo));
  mo.nodes = nodes; mo.anchorsets = anchorsets;
  int anchors = 0;
  for (int i = 0; i < anchorsets; i++) { mo.anch[i] = anch[i]; mo.fcnt[i] = fcnt[i]; anchors |= anch[i]; }
  const int nonanch = anchors ^ ((1 << nodes) - 1);
  FX fx;
  uint32_t fl = 0;
  fx.mo = &mo; fx.top = ((1 << nodes) - 1) ^ nonanch; fx.nonanch = nonanch; fx.fl = &fl;
  for (int i = 0; i < 64; i++) { fx.venn[i] = venn64[i]; fx.fr[i] = 0; }
  for (int i = 0; i < anchorsets; i++) fx.fr[anch[i]] = fcnt[i];
  const u128 r = fx.rec(anchorsets - 1, anch[anchorsets - 1], 1);
  *lo = (u64)r; *hi = (u64)(r >> 64); *flags = fl;
}

// ncr with the reference semantics (n < 0 or k outside [0, n] -> 0 except k == 0 with n >= 0).
void oracle_ncr(int n, int k, u64* lo, u64* hi, uint32_t* flags)
{
  uint32_t fl = 0;
  const u128 r = ncr(n, k, fl);
  *lo = (u64)r; *hi = (u64)(r >> 64); *flags = fl;
}

}  // extern "C"

// ---------------------------------------------------------------- pivot schedule bytes
//
// Not a restatement of the reference: the unit counts of the library's pivot schedule
// for wedge cores (its algorithmic-bytes model), computed from the CSR alone, so that
// the bytes a pivot launch reports can be recomputed from units. The schedule's work
// items (pivot w with a range of its centres), its c3 windows and its LDS hubs follow
// the library's rules:
//   items: pivot w of cost c = sum of its neighbours' degrees is cut into
//          k = max(1, min(d_w, c / max(budget, c / 8 + 1))) centre ranges; every item
//          repeats the pivot's 2-hop scatter (phase A);
//   windows: a centre u of pivot w (w at position rq of N(u)) walks positions rq+1..d_u-1,
//          with c3 (a launch that needs it and t(u,w) > 0) in ceil((d_u-rq-1)/ecap) windows;
//   hubs: in launches without c3, the hub_slots vertices of highest degree (ties by the
//          smaller id) have LDS counters.
// Units (out[0..12]): 2-hop paths to D, to LDS hubs; adjacency marks to D, to LDS hubs;
// pivot neighbours visited and the sum of their bitlen(d_x + 1); centres and the sum of
// bitlen(d_u + 1); c3 entries (x of T(u,w), per window) and the sum of their search
// lengths bitlen(t_ux + 1) (counted twice in a window that is not the centre's last);
// list elements scattered (= sum over directed edges of C(t_e, 2)); walk positions to
// D, to LDS hubs.
static int bitlen(long long x) { return 64 - __builtin_clzll((unsigned long long)x); }

extern "C" int oracle_pivot_units(int n, long long m, const int* idx, const int* adj, int ecap, int hub_slots,
                                  long long budget, int c3, int threads, u64* out)
{
  G g{n, m, idx, adj};
  if (threads > 0) omp_set_num_threads(threads);
  // supports t_e per directed edge (merge intersection)
  std::vector<int> tsup((size_t)std::max<long long>(1, m));
#pragma omp parallel for schedule(dynamic, 256)
  for (int u = 0; u < n; u++) {
    for (int e = g.beg(u); e < g.end(u); e++) {
      const int x = adj[e];
      int i = g.beg(u), j = g.beg(x), c = 0;
      while (i < g.end(u) && j < g.end(x)) {
        if (adj[i] < adj[j]) i++;
        else if (adj[i] > adj[j]) j++;
        else { c++; i++; j++; }
      }
      tsup[e] = c;
    }
  }
  std::vector<char> hub((size_t)n, 0);
  if (hub_slots > 0) {
    std::vector<int> ord(n);
    for (int v = 0; v < n; v++) ord[v] = v;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return g.deg(a) > g.deg(b); });
    for (int r = 0; r < std::min(n, hub_slots); r++) hub[ord[r]] = 1;
  }
  u64 U[13] = {0};
#pragma omp parallel
  {
    u64 L[13] = {0};
#pragma omp for schedule(dynamic, 64)
    for (int w = 0; w < n; w++) {
      const int dw = g.deg(w);
      if (dw == 0) continue;
      long long cost = 0;
      for (int e = g.beg(w); e < g.end(w); e++) cost += g.deg(adj[e]);
      const long long k = std::max<long long>(1, std::min<long long>(dw, cost / std::max(budget, cost / 8 + 1)));
      // phase A, once per item
      u64 pd = 0, ph = 0, md = 0, mh = 0, xb = 0;
      for (int e = g.beg(w); e < g.end(w); e++) {
        const int x = adj[e];
        xb += bitlen((long long)g.deg(x) + 1);
        if (x > w) { if (hub[x]) mh++; else md++; }
        const int s = (int)(std::upper_bound(adj + g.beg(x), adj + g.end(x), w) - adj);
        for (int i = s; i < g.end(x); i++) { if (hub[adj[i]]) ph++; else pd++; }
      }
      long long items = 0;
      for (long long s = 0; s < k; s++) if (dw * (s + 1) / k > dw * s / k) items++;
      L[0] += pd * items; L[1] += ph * items; L[2] += md * items; L[3] += mh * items;
      L[4] += (u64)dw * items; L[5] += xb * items;
      // phase B: every centre once
      for (int e = g.beg(w); e < g.end(w); e++) {
        const int u = adj[e];
        const int ub = g.beg(u), du = g.deg(u);
        const int rq = (int)(std::lower_bound(adj + ub, adj + ub + du, w) - (adj + ub));
        if (rq + 1 >= du) continue;
        L[6]++;
        L[7] += bitlen((long long)du + 1);
        for (int q = rq + 1; q < du; q++) { if (hub[adj[ub + q]]) L[12]++; else L[11]++; }
        const int tuw = tsup[(size_t)ub + rq];
        if (!c3 || tuw == 0) continue;
        const long long nwin = (du - rq - 1 + ecap - 1) / ecap;
        // x in T(u,w): common neighbours of u and w, with t_ux
        int i = ub, j = g.beg(w);
        while (i < g.end(u) && j < g.end(w)) {
          if (adj[i] < adj[j]) i++;
          else if (adj[i] > adj[j]) j++;
          else {
            const int b = bitlen((long long)tsup[i] + 1);
            L[8] += nwin;
            L[9] += (u64)b * (u64)(2 * nwin - 1);
            i++; j++;
          }
        }
      }
    }
#pragma omp critical
    for (int q = 0; q < 13; q++) U[q] += L[q];
  }
  if (c3)
    for (long long e = 0; e < m; e++) U[10] += (u64)tsup[e] * (u64)(tsup[e] - 1) / 2;
  for (int q = 0; q < 13; q++) out[q] = U[q];
  return 0;
}
