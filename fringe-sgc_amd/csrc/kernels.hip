// Fringe-SGC counting kernels for gfx950 (CDNA4, wave64).
//
// One persistent kernel drains a device work queue of (job, root range) items; a job
// is one motif on one graph, so a single launch can count one motif (the plain path)
// or many motifs of any core class at once (the batched path). Within an item the
// wavefront takes 64 roots, filters them per lane, and then processes the surviving
// roots one at a time cooperatively (64-bit ballot loop): neighbour-list intersections
// are lane-strided over the shorter list with binary search in the longer one, and the
// fringe formulas (fc3, functionX) run per lane in parallel over the matches found.
// All counts are exact unsigned 128-bit integers with overflow flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "fsgc_plan.h"

namespace fsgc {
namespace dev {

typedef unsigned __int128 u128;
typedef unsigned long long ull;

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kVennSlots = 16;  // general-core Venn vectors staged per wave

enum { iu = 1, iv = 2, iw = 4, iuv = 3, iuw = 5, ivw = 6, iuvw = 7 };

// inverses of the odd numbers 1, 3, ..., 63 modulo 2^128 (for exact division)
__constant__ ull c_inv_lo[32];
__constant__ ull c_inv_hi[32];

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T bcast(T x, int src) { return __shfl(x, src, kWave); }

__device__ __forceinline__ int wave_sum(int x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}

// ------------------------------------------------------------------ exact arithmetic

__device__ __forceinline__ u128 mul_lo128(u128 a, u128 b)
{
  const ull al = (ull)a, ah = (ull)(a >> 64), bl = (ull)b, bh = (ull)(b >> 64);
  u128 p = (u128)al * bl;
  const ull cross = ah * bl + al * bh;
  return p + ((u128)cross << 64);
}

// a * b with overflow detection (flag kOvf128)
__device__ __forceinline__ u128 mulx(u128 a, u128 b, uint32_t& fl)
{
  const ull al = (ull)a, ah = (ull)(a >> 64), bl = (ull)b, bh = (ull)(b >> 64);
  const u128 p = (u128)al * bl;
  if ((ah | bh) == 0) return p;
  if (ah && bh) fl |= kOvf128;
  const u128 c = (u128)ah * bl + (u128)al * bh;
  if ((ull)(c >> 64)) fl |= kOvf128;
  const u128 r = p + ((u128)(ull)c << 64);
  if (r < p) fl |= kOvf128;
  return r;
}

__device__ __forceinline__ u128 addx(u128 a, u128 b, uint32_t& fl)
{
  const u128 r = a + b;
  if (r < a) fl |= kOvf128;
  return r;
}

// x / i for 1 <= i <= 63 when the division is exact: shift out the power of two, then
// multiply by the inverse of the odd part modulo 2^128.
__device__ __forceinline__ u128 div_exact(u128 x, int i)
{
  const int s = __builtin_ctz(i);
  const int o = i >> s;
  x >>= s;
  if (o == 1) return x;
  const u128 inv = ((u128)c_inv_hi[o >> 1] << 64) | c_inv_lo[o >> 1];
  return mul_lo128(x, inv);
}

// FC:189-197 semantics: k = min(k, n - k); k < 0 -> 0; k == 0 -> 1; otherwise the exact
// binomial. The reference's u64 loop wraps iff k * C(n, k) >= 2^64 (kRefNckInexact).
__device__ u128 ncr(int n, int k, uint32_t& fl)
{
  k = min(k, n - k);
  if (k < 0) return 0;
  if (k == 0) return 1;
  if (k == 1) return (u128)(unsigned)n;
  // fast path: all intermediates i * C(n-k+i, i) <= k * C(n, k) fit 64 bits
  ull r = 1;
  int i = 1;
  for (; i <= k; i++) {
    const ull f = (ull)(unsigned)(n - k + i);
    if (__umul64hi(r, f) != 0) break;
    r = (r * f) / (ull)i;
  }
  u128 R = r;
  if (i <= k) {
    fl |= kRefNckInexact;
    for (; i <= k; i++) {
      R = mulx(R, (u128)(unsigned)(n - k + i), fl);
      R = div_exact(R, i);
    }
    return R;
  }
  if (__umul64hi(r, (ull)k) != 0) fl |= kRefNckInexact;
  return R;
}

// ------------------------------------------------------------------ graph helpers

struct G {
  const int* __restrict__ idx;
  const int* __restrict__ adj;
  __device__ __forceinline__ int beg(int u) const { return idx[u]; }
  __device__ __forceinline__ int end(int u) const { return idx[u + 1]; }
  __device__ __forceinline__ int deg(int u) const { return idx[u + 1] - idx[u]; }
};

// membership of x in the sorted range adj[l, r)
__device__ __forceinline__ bool contains(const int* __restrict__ a, int l, int r, int x)
{
  int lo = l, hi = r;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < r && a[lo] == x;
}

// |N(a) & N(b)|, lane-strided over the shorter list (wave-cooperative, uniform args)
__device__ int coop_common(const G& g, int ab, int ae, int bb, int be)
{
  if (ae - ab > be - bb) { int t = ab; ab = bb; bb = t; t = ae; ae = be; be = t; }
  int c = 0;
  for (int i = ab + lane_id(); i < ae; i += kWave) c += contains(g.adj, bb, be, g.adj[i]);
  return wave_sum(c);
}

// owner of CSR position j, searching vertices [lo, hi]
__device__ __forceinline__ int owner_of(const int* __restrict__ idx, int lo, int hi, long long j)
{
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (idx[mid] <= j) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------------ accumulation

struct LaneAcc {
  ull lo = 0, hi = 0, matches = 0;
  uint32_t fl = 0;
  __device__ __forceinline__ void add(u128 x)
  {
    const u128 s = (((u128)hi << 64) | lo);
    const u128 r = s + x;
    if (r < s) fl |= kOvf128;
    lo = (ull)r;
    hi = (ull)(r >> 64);
  }
};

__device__ void flush(LaneAcc& a, Accum* out)
{
  ull lo = a.lo, hi = a.hi, mt = a.matches;
  uint32_t fl = a.fl;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const ull olo = __shfl_xor(lo, o, kWave), ohi = __shfl_xor(hi, o, kWave);
    const ull nlo = lo + olo;
    const ull c = nlo < lo;
    const u128 nh = (u128)hi + ohi + c;
    if ((ull)(nh >> 64)) fl |= kOvf128;
    lo = nlo;
    hi = (ull)nh;
    mt += __shfl_xor(mt, o, kWave);
    fl |= __shfl_xor(fl, o, kWave);
  }
  if (lane_id() == 0 && (lo | hi | mt | fl)) {
    const ull old = atomicAdd((ull*)&out->lo, lo);
    const ull carry = (old + lo) < old;
    const ull add_hi = hi + carry;
    if (add_hi < hi) fl |= kOvf128;
    if (add_hi) {
      const ull oh = atomicAdd((ull*)&out->hi, add_hi);
      if (oh + add_hi < oh) fl |= kOvf128;
    }
    if (mt) atomicAdd((ull*)&out->matches, mt);
    if (fl) atomicOr(&out->flags, fl);
  }
  a = LaneAcc();
}

// ------------------------------------------------------------------ fringe formulas

// FC:233-327 (fc3): ways to place the fringes of the 7 anchor sets of a 3-vertex core
// disjointly into the 7 Venn regions (loop order uvw, uv, uw, vw, u; v and w merged).
__device__ u128 fc3(const int* fr, const int* venn, uint32_t& fl)
{
  const int fr_uvw = fr[iuvw], fr_uv = fr[iuv], fr_uw = fr[iuw], fr_vw = fr[ivw];
  const int fr_u = fr[iu], fr_v = fr[iv], fr_w = fr[iw];
  int vd_uvw = venn[iuvw], vd_uv = venn[iuv], vd_uw = venn[iuw], vd_vw = venn[ivw];
  const int vd_u = venn[iu], vd_v = venn[iv], vd_w = venn[iw];
  if (vd_uvw < fr_uvw) return 0;
  u128 count = 0;
  const u128 cuvw = ncr(vd_uvw, fr_uvw, fl);
  vd_uvw -= fr_uvw;
  for (int i = min(vd_uvw, fr_uv); i >= 0; i--) {
    if (vd_uv < fr_uv - i) break;
    const u128 c1 = mulx(mulx(cuvw, ncr(vd_uvw, i, fl), fl), ncr(vd_uv, fr_uv - i, fl), fl);
    vd_uvw -= i; vd_uv -= fr_uv - i;
    for (int j = min(vd_uvw, fr_uw); j >= 0; j--) {
      if (vd_uw < fr_uw - j) break;
      const u128 c2 = mulx(mulx(c1, ncr(vd_uvw, j, fl), fl), ncr(vd_uw, fr_uw - j, fl), fl);
      vd_uvw -= j; vd_uw -= fr_uw - j;
      for (int k = min(vd_uvw, fr_vw); k >= 0; k--) {
        if (vd_vw < fr_vw - k) break;
        const u128 c3 = mulx(mulx(c2, ncr(vd_uvw, k, fl), fl), ncr(vd_vw, fr_vw - k, fl), fl);
        vd_uvw -= k; vd_vw -= fr_vw - k;
        for (int a = min(vd_uvw, fr_u); a >= 0; a--) {
          const u128 ta = mulx(c3, ncr(vd_uvw, a, fl), fl);
          vd_uvw -= a;
          int vd_uvw_vw = vd_uvw + vd_vw;
          for (int b = min(vd_uv, fr_u - a); b >= 0; b--) {
            const u128 tb = mulx(ta, ncr(vd_uv, b, fl), fl);
            vd_uv -= b;
            for (int c = min(vd_uw, fr_u - a - b); c >= 0; c--) {
              if (vd_u < fr_u - a - b - c) break;
              const u128 tc = mulx(mulx(tb, ncr(vd_uw, c, fl), fl), ncr(vd_u, fr_u - a - b - c, fl), fl);
              vd_uw -= c;
              for (int d = min(vd_uvw_vw, fr_v); d >= 0; d--) {
                if (vd_uv + vd_v < fr_v - d) break;
                const u128 td = mulx(mulx(tc, ncr(vd_uvw_vw, d, fl), fl), ncr(vd_uv + vd_v, fr_v - d, fl), fl);
                vd_uvw_vw -= d;
                if (vd_uvw_vw + vd_uw + vd_w >= fr_w)
                  count = addx(count, mulx(td, ncr(vd_uvw_vw + vd_uw + vd_w, fr_w, fl), fl), fl);
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

// FC:2030-2048 (general_edge_core) and FC:1381-1406 (wedge): the two-vertex closed
// form; `corr` is how many core vertices each list holds besides the shared ones
// (1 for an edge core, 1 or 2 for wedge endpoints).
__device__ u128 pair_formula(int da, int db, int c_ab, int corr, int f_a, int f_ab, int f_b, uint32_t& fl)
{
  const int a_ab = c_ab - f_ab;
  if (a_ab < 0) return 0;
  const int ca = da - (corr + c_ab), cb = db - (corr + c_ab);
  const int bound = min(f_a, a_ab);
  u128 acc = 0;
  for (int j = 0; j <= bound; j++) {
    const int k = f_a - j;
    u128 term = mulx(ncr(ca, k, fl), ncr(a_ab + cb - j, f_b, fl), fl);
    if (f_a != f_b) term = addx(term, mulx(ncr(cb, k, fl), ncr(a_ab + ca - j, f_b, fl), fl), fl);
    acc = addx(acc, mulx(term, ncr(a_ab, j, fl), fl), fl);
  }
  return mulx(acc, ncr(c_ab, f_ab, fl), fl);
}

// FC:408-486 (functionX) as a per-lane depth-first walk over a precomputed plan of
// (anchor set, region) levels: each level places some of the set's fringes into the
// region (all the remaining ones at the set's last region).
__device__ u128 functionx(const JobDesc* __restrict__ J, int* venn, uint32_t& fl)
{
  const int L = J->nlev;
  int fr[64];
  for (int i = 0; i < 64; i++) fr[i] = J->fcnt_by_mask[i];
  int acur[kMaxPlanLevels];
  u128 prod[kMaxPlanLevels];
  u128 sum = 0;
  int lvl = 0;
  prod[0] = 1;
  {
    const int S = J->lev_set[0], R = J->lev_reg[0];
    acur[0] = J->lev_last[0] ? fr[S] : min(venn[R], fr[S]);
  }
  while (lvl >= 0) {
    const int S = J->lev_set[lvl], R = J->lev_reg[lvl];
    const int a = acur[lvl];
    if (a < 0) {
      lvl--;
      if (lvl >= 0) {
        const int pS = J->lev_set[lvl], pR = J->lev_reg[lvl];
        venn[pR] += acur[lvl];
        fr[pS] += acur[lvl];
        acur[lvl] = J->lev_last[lvl] ? -1 : acur[lvl] - 1;
      }
      continue;
    }
    const u128 p = mulx(prod[lvl], ncr(venn[R], a, fl), fl);
    if (p == 0) { acur[lvl] = J->lev_last[lvl] ? -1 : a - 1; continue; }
    if (J->lev_final[lvl]) { sum = addx(sum, p, fl); acur[lvl] = -1; continue; }
    venn[R] -= a;
    fr[S] -= a;
    lvl++;
    prod[lvl] = p;
    const int nS = J->lev_set[lvl], nR = J->lev_reg[lvl];
    acur[lvl] = J->lev_last[lvl] ? fr[nS] : min(venn[nR], fr[nS]);
  }
  return sum;
}

// ------------------------------------------------------------------ classes

__device__ __forceinline__ bool in_slice(int u, int smod, int srem) { return smod <= 1 || (u % smod) == srem; }

// FC:2152-2169 (k_stars)
__device__ void do_vertex(const JobDesc* __restrict__ J, const WorkItem& it, LaneAcc& acc)
{
  const G g{J->idx, J->adj};
  const int t = J->p[0];
  const int smod = J->smod, srem = J->srem;
  for (long long u = it.begin + lane_id(); u < it.end; u += kWave) {
    if (!in_slice((int)u, smod, srem)) continue;
    const int d = g.deg((int)u);
    if (d >= t) { acc.add(ncr(d, t, acc.fl)); acc.matches++; }
  }
}

// Loads 64 consecutive CSR positions of an edge-rooted item: root (u, v) per lane.
struct Root {
  bool ok;
  int u, v, du, dv;
  long long j;
};

__device__ __forceinline__ Root load_root(const G& g, const WorkItem& it, long long base)
{
  Root r;
  r.j = base + lane_id();
  r.ok = r.j < it.end;
  r.u = r.v = r.du = r.dv = 0;
  if (r.ok) {
    r.u = owner_of(g.idx, it.u0, it.u1, r.j);
    r.v = g.adj[r.j];
    r.du = g.deg(r.u);
    r.dv = g.deg(r.v);
  }
  return r;
}

// FC:2054-2103 (init_wl_edge filters) + FC:1864-1966, 1999-2051 (edge kernels)
__device__ void do_edge(const JobDesc* __restrict__ J, const WorkItem& it, LaneAcc& acc)
{
  const G g{J->idx, J->adj};
  const int kind = J->kind;
  const int f_u = J->p[0], f_uv = J->p[1], f_v = J->p[2], lim1 = J->p[3], lim2 = J->p[4];
  const int smod = J->smod, srem = J->srem, n = J->n;
  for (long long base = it.begin; base < it.end; base += kWave) {
    const Root r = load_root(g, it, base);
    const bool act = r.ok && r.v < r.u && r.du > lim1 && r.dv > lim1 && r.du + r.dv > lim2 && in_slice(r.u, smod, srem);
    ull mask = __ballot(act);
    while (mask) {
      const int who = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int U = bcast(r.u, who), V = bcast(r.v, who), DU = bcast(r.du, who), DV = bcast(r.dv, who);
      if (kind == kEdgeGeneral) {
        const int c = coop_common(g, g.beg(U), g.end(U), g.beg(V), g.end(V));
        if (lane_id() == who) { acc.add(pair_formula(DU, DV, c, 1, f_u, f_uv, f_v, acc.fl)); acc.matches++; }
      } else {
        u128 c12 = 0;
        if (kind == kEdgeTailed) c12 = addx(ncr(DU - 2, f_v, acc.fl), ncr(DV - 2, f_v, acc.fl), acc.fl);
        int sb = g.beg(U), se = g.end(U), lb = g.beg(V), le = g.end(V);
        if (DU >= DV) { sb = g.beg(V); se = g.end(V); lb = g.beg(U); le = g.end(U); }
        for (int i = sb + lane_id(); i < se; i += kWave) {
          const int x = g.adj[i];
          if (x >= V) break;
          if (!contains(g.adj, lb, le, x)) continue;
          acc.matches++;
          if (kind == kEdgeTriangle) acc.add(1);
          else acc.add(addx(c12, ncr(g.deg(x) - 2, f_v, acc.fl), acc.fl));
        }
      }
    }
  }
}

// raw three-list counts for the Venn regions of N(a), N(b), N(c) by a walk over N(c)
// (FC:934-964, FC:1480-1510); every lane receives the wave totals.
__device__ __forceinline__ void coop_venn3(const G& g, int a, int b, int c, int& t13, int& t23, int& t123)
{
  const int ab = g.beg(a), ae = g.end(a), bb = g.beg(b), be = g.end(b);
  int x13 = 0, x23 = 0, x123 = 0;
  for (int i = g.beg(c) + lane_id(); i < g.end(c); i += kWave) {
    const int q = g.adj[i];
    const bool b13 = contains(g.adj, ab, ae, q);
    const bool b23 = contains(g.adj, bb, be, q);
    x13 += b13; x23 += b23; x123 += (b13 && b23);
  }
  t13 = wave_sum(x13); t23 = wave_sum(x23); t123 = wave_sum(x123);
}

// FC:1653-1751 (wedge worklists) + FC:1409-1477 (wedgecore_specialized) +
// FC:1513-1650 (general_wedge_core). The candidates w < v of root (u, v) are the
// positions of N(u) before v.
__device__ void do_wedge(const JobDesc* __restrict__ J, const WorkItem& it, LaneAcc& acc)
{
  const G g{J->idx, J->adj};
  const int min_u = J->p[0], lim = J->p[1], sum1 = J->p[2], sum2 = J->p[3], sum3 = J->p[4];
  const int smod = J->smod, srem = J->srem, n = J->n;
  const bool special = (J->kind == kWedgeSpecial);
  const int rot = J->rot;
  for (long long base = it.begin; base < it.end; base += kWave) {
    const Root r = load_root(g, it, base);
    bool act = r.ok && in_slice(r.u, smod, srem);
    if (special) act = act && r.du >= 2 && r.dv >= lim;
    else act = act && r.du >= sum1 && r.dv >= max(sum1, sum2 - r.du);
    ull mask = __ballot(act);
    while (mask) {
      const int who = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int U = bcast(r.u, who), V = bcast(r.v, who), DU = bcast(r.du, who), DV = bcast(r.dv, who);
      const int JV = (int)bcast(r.j, who);
      const int ub = g.beg(U), vb = g.beg(V), ve = g.end(V);
      if (special) {
        const int f_v = J->p[5], f_vw = J->p[6], f_w = J->p[7];
        const int thres = max(J->p[10], J->p[8] + J->p[9] - DV);
        for (int wb = ub; wb < JV; wb += kWave) {
          const int e = wb + lane_id();
          int w = 0, dw = 0;
          bool a = e < JV;
          if (a) { w = g.adj[e]; dw = g.deg(w); a = dw >= thres; }
          ull wm = __ballot(a);
          int mine = 0;
          while (wm) {
            const int w2 = __builtin_ctzll(wm);
            wm &= wm - 1;
            const int W = bcast(w, w2);
            const int c = coop_common(g, vb, ve, g.beg(W), g.end(W));
            if (lane_id() == w2) mine = c - 1;
          }
          if (a) {
            const int corr = contains(g.adj, g.beg(w), g.end(w), V) ? 2 : 1;
            acc.add(pair_formula(DV, dw, mine, corr, f_v, f_vw, f_w, acc.fl));
            acc.matches++;
          }
        }
        continue;
      }
      const int p12 = coop_common(g, ub, g.end(U), vb, ve);
      const bool cond = (DU >= min_u) && (DV >= lim);
      const int thres1 = max(lim, sum2 - DV);
      const int thres2 = max(max(sum1, sum3 - DU - DV), max(sum2 - DU, sum2 - DV));
      for (int wb = ub; wb < JV; wb += kWave) {
        const int e = wb + lane_id();
        int w = 0, dw = 0;
        bool a = e < JV, tri = false;
        if (a) {
          w = g.adj[e];
          dw = g.deg(w);
          a = dw >= thres2;
          if (a) {
            tri = contains(g.adj, vb, ve, w);
            a = tri ? (U < w) : (cond && dw >= thres1);
          }
        }
        ull wm = __ballot(a);
        int m13 = 0, m23 = 0, m123 = 0;
        while (wm) {
          const int w2 = __builtin_ctzll(wm);
          wm &= wm - 1;
          int t13, t23, t123;
          coop_venn3(g, U, V, bcast(w, w2), t13, t23, t123);
          if (lane_id() == w2) { m13 = t13; m23 = t23; m123 = t123; }
        }
        if (a) {
          int venn[8];
          const int c123 = m123, c12 = p12 - c123, c13 = m13 - c123, c23 = m23 - c123;
          venn[0] = 0;
          venn[iu] = DU - (p12 + c13);
          venn[iv] = DV - (p12 + c23);
          venn[iw] = dw - (c13 + c23 + c123);
          venn[iuv] = c12; venn[iuw] = c13; venn[ivw] = c23; venn[iuvw] = c123;
          u128 s = 0;
          if (!tri) {
            venn[iu] -= 2; venn[ivw] -= 1;  // FC:1598-1602
            s = fc3(J->fr[0], venn, acc.fl);
            if (rot > 3) s = addx(s, fc3(J->fr[3], venn, acc.fl), acc.fl);
          } else {
            venn[iuv] -= 1; venn[iuw] -= 1; venn[ivw] -= 1;  // FC:1631-1636
            for (int q = 0; q < rot; q++) s = addx(s, fc3(J->fr[q], venn, acc.fl), acc.fl);
          }
          acc.add(s);
          acc.matches++;
        }
      }
    }
  }
}

// FC:1167-1218 (init_wl_triangle) + FC:700-931 (clique kernels) + FC:934-1062
// (com3, common3_triangle, general_triangle_core)
__device__ void do_triangle(const JobDesc* __restrict__ J, const WorkItem& it, LaneAcc& acc)
{
  const G g{J->idx, J->adj};
  const int sum1 = J->p[0], sum2 = J->p[1], sum3 = J->p[2], tails = J->p[3];
  const int smod = J->smod, srem = J->srem, n = J->n;
  const int kind = J->kind, rot = J->rot;
  for (long long base = it.begin; base < it.end; base += kWave) {
    const Root r = load_root(g, it, base);
    const bool act = r.ok && r.v < r.u && r.du >= sum1 && r.dv >= max(sum1, sum2 - r.du) && in_slice(r.u, smod, srem);
    ull mask = __ballot(act);
    while (mask) {
      const int who = __builtin_ctzll(mask);
      mask &= mask - 1;
      const int U = bcast(r.u, who), V = bcast(r.v, who), DU = bcast(r.du, who), DV = bcast(r.dv, who);
      const int ub = g.beg(U), ue = g.end(U), vb = g.beg(V), ve = g.end(V);
      int p12 = 0;
      if (kind == kTriGeneral) p12 = coop_common(g, ub, ue, vb, ve);
      const int thres2 = (kind == kTriGeneral) ? max(max(sum1, sum3 - DU - DV), max(sum2 - DU, sum2 - DV)) : 3;
      u128 c12 = 0;
      if (kind == kTriTailedClique) c12 = addx(ncr(DU - 3, tails, acc.fl), ncr(DV - 3, tails, acc.fl), acc.fl);
      int sb = ub, se = ue, lb = vb, le = ve;
      if (!(DU < DV)) { sb = vb; se = ve; lb = ub; le = ue; }
      for (int b0 = sb;; b0 += kWave) {
        const int e = b0 + lane_id();
        const int x = (e < se) ? g.adj[e] : 0x7fffffff;
        if (__ballot(x < V) == 0) break;
        bool a = false;
        int dx = 0;
        if (x < V && contains(g.adj, lb, le, x)) { dx = g.deg(x); a = dx >= thres2; }
        ull wm = __ballot(a);
        if (kind == kTriGeneral) {
          int m13 = 0, m23 = 0, m123 = 0;
          while (wm) {
            const int w2 = __builtin_ctzll(wm);
            wm &= wm - 1;
            int t13, t23, t123;
            coop_venn3(g, U, V, bcast(x, w2), t13, t23, t123);
            if (lane_id() == w2) { m13 = t13; m23 = t23; m123 = t123; }
          }
          if (a) {
            int venn[8];
            const int c123 = m123;
            const int cc12 = p12 - (c123 + 1), c13 = m13 - (c123 + 1), c23 = m23 - (c123 + 1);
            venn[0] = 0;
            venn[iu] = DU - (cc12 + c13 + (c123 + 2));
            venn[iv] = DV - (cc12 + c23 + (c123 + 2));
            venn[iw] = dx - (c13 + c23 + (c123 + 2));
            venn[iuv] = cc12; venn[iuw] = c13; venn[ivw] = c23; venn[iuvw] = c123;
            u128 s = 0;
            for (int q = 0; q < rot; q++) s = addx(s, fc3(J->fr[q], venn, acc.fl), acc.fl);
            acc.add(s);
            acc.matches++;
          }
        } else {
          if (a) acc.matches++;
          while (wm) {
            const int w2 = __builtin_ctzll(wm);
            wm &= wm - 1;
            const int W = bcast(x, w2), DW = bcast(dx, w2);
            u128 c123 = 0;
            if (kind == kTriTailedClique) c123 = addx(c12, ncr(DW - 3, tails, acc.fl), acc.fl);
            // walk the shortest of the three lists below W, probe the other two
            int l0 = ub, e0 = ue, l1 = vb, e1 = ve, l2 = g.beg(W), e2 = g.end(W);
            if (e1 - l1 < e0 - l0) { int t = l0; l0 = l1; l1 = t; t = e0; e0 = e1; e1 = t; }
            if (e2 - l2 < e0 - l0) { int t = l0; l0 = l2; l2 = t; t = e0; e0 = e2; e2 = t; }
            for (int i = l0 + lane_id(); i < e0; i += kWave) {
              const int q = g.adj[i];
              if (q >= W) break;
              if (contains(g.adj, l1, e1, q) && contains(g.adj, l2, e2, q)) {
                if (kind == kTriClique) acc.add(1);
                else acc.add(addx(c123, ncr(g.deg(q) - 3, tails, acc.fl), acc.fl));
              }
            }
          }
        }
      }
    }
  }
}

// FC:345-405 (commonX): venn[S] = number of non-core vertices whose set of adjacent
// anchors is exactly S. Each vertex is counted once, at its lowest adjacent anchor.
__device__ void coop_commonx(const G& g, int nodes, int nonanch, const int* st, int* venn)
{
  venn[lane_id()] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int i = 0; i < nodes; i++) {
    if ((nonanch >> i) & 1) continue;
    const int a = st[i];
    for (int e = g.beg(a) + lane_id(); e < g.end(a); e += kWave) {
      const int x = g.adj[e];
      bool first = true;
      for (int k = 0; k < i && first; k++) {
        if ((nonanch >> k) & 1) continue;
        if (contains(g.adj, g.beg(st[k]), g.end(st[k]), x)) first = false;
      }
      if (!first) continue;
      int bmp = 1 << i;
      for (int k = i + 1; k < nodes; k++) {
        if ((nonanch >> k) & 1) continue;
        if (contains(g.adj, g.beg(st[k]), g.end(st[k]), x)) bmp |= 1 << k;
      }
      atomicAdd(&venn[bmp], 1);
    }
  }
  if (lane_id() < nodes) {
    const int x = st[lane_id()];
    int bmp = 0;
    for (int k = 0; k < nodes; k++) {
      if ((nonanch >> k) & 1) continue;
      if (contains(g.adj, g.beg(st[k]), g.end(st[k]), x)) bmp |= 1 << k;
    }
    if (bmp) atomicSub(&venn[bmp], 1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

struct Stack6 { int v[6]; };

// FC:489-564 (general_core_level<L>)
template <int L>
__device__ void gen_level(const JobDesc* __restrict__ J, const G& g, Stack6 st, LaneAcc& acc, int* slots)
{
  const int nodes = J->p[0];
  const int list = J->p[7 + L];
  const int mindeg = J->p[2 + L];
  const int nonanch = J->p[12];
  const int wn = st.v[__builtin_ctz(list)];
  const int b = g.beg(wn), e = g.end(wn);
  const bool leaf = (L == nodes - 2);
  for (int cb = b; cb < e; cb += kWave) {
    const int i = cb + lane_id();
    int x = -1;
    bool a = i < e;
    if (a) {
      x = g.adj[i];
      a = g.deg(x) >= mindeg;
      int skip = list ^ ((2 << L) - 1);
      while (a && skip) { const int k = __builtin_ctz(skip); skip &= skip - 1; if (x == st.v[k]) a = false; }
      int need = list & (list - 1);
      while (a && need) {
        const int k = __builtin_ctz(need);
        need &= need - 1;
        if (!contains(g.adj, g.beg(st.v[k]), g.end(st.v[k]), x)) a = false;
      }
    }
    ull hm = __ballot(a);
    if (leaf) {
      while (hm) {
        // stage up to kVennSlots Venn vectors, then evaluate them in parallel
        int slot = 0, myslot = -1;
        while (hm && slot < kVennSlots) {
          const int w2 = __builtin_ctzll(hm);
          hm &= hm - 1;
          Stack6 s2 = st;
          s2.v[L + 1] = bcast(x, w2);
          coop_commonx(g, nodes, nonanch, s2.v, slots + slot * 64);
          if (lane_id() == w2) myslot = slot;
          slot++;
        }
        if (myslot >= 0) {
          int venn[64];
          for (int q = 0; q < 64; q++) venn[q] = slots[myslot * 64 + q];
          acc.add(functionx(J, venn, acc.fl));
          acc.matches++;
        }
        __builtin_amdgcn_wave_barrier();
      }
    } else {
      if constexpr (L < 4) {
        while (hm) {
          const int w2 = __builtin_ctzll(hm);
          hm &= hm - 1;
          Stack6 s2 = st;
          s2.v[L + 1] = bcast(x, w2);
          gen_level<L + 1>(J, g, s2, acc, slots);
        }
      }
    }
  }
}

// FC:617-663 (init_wl_general) + FC:567-614 (general_core_level0)
__device__ void do_general(const JobDesc* __restrict__ J, const WorkItem& it, LaneAcc& acc, int* slots)
{
  const G g{J->idx, J->adj};
  const int startdeg = J->p[1], dd0 = J->p[2];
  const int smod = J->smod, srem = J->srem, n = J->n;
  for (long long base = it.begin; base < it.end; base += kWave) {
    const Root r = load_root(g, it, base);
    const bool act = r.ok && r.du >= startdeg && r.dv >= dd0 && in_slice(r.u, smod, srem);
    ull mask = __ballot(act);
    while (mask) {
      const int who = __builtin_ctzll(mask);
      mask &= mask - 1;
      Stack6 st;
      for (int q = 0; q < 6; q++) st.v[q] = 0;
      st.v[0] = bcast(r.u, who);
      st.v[1] = bcast(r.v, who);
      gen_level<1>(J, g, st, acc, slots);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_count(const JobDesc* __restrict__ jobs, const WorkItem* __restrict__ items,
                                                  int nitems, int* __restrict__ head)
{
  __shared__ int s_venn[kWavesPerBlock][kVennSlots * 64];
  int* slots = s_venn[threadIdx.x / kWave];
  LaneAcc acc;
  for (;;) {
    int i = 0;
    if (lane_id() == 0) i = atomicAdd(head, 1);
    i = bcast(i, 0);
    if (i >= nitems) break;
    const WorkItem it = items[i];
    const JobDesc* J = jobs + it.job;
    switch (J->cls) {
      case kVertex: do_vertex(J, it, acc); break;
      case kEdge: do_edge(J, it, acc); break;
      case kWedge: do_wedge(J, it, acc); break;
      case kTriangle: do_triangle(J, it, acc); break;
      default: do_general(J, it, acc, slots); break;
    }
    flush(acc, J->out);
  }
}

}  // namespace dev

// ==================================================================== host driver

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      if (err) *err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
      return -5;                                                                      \
    }                                                                                 \
  } while (0)

static bool g_inv_ready = false;

int device_init(int device, std::string* err)
{
  HIP_TRY(hipSetDevice(device));
  if (!g_inv_ready) {
    typedef unsigned __int128 u128;
    unsigned long long lo[32], hi[32];
    for (int k = 0; k < 32; k++) {
      const u128 o = (u128)(2 * k + 1);
      u128 x = o;  // Newton iteration for the inverse modulo 2^128
      for (int it = 0; it < 8; it++) x = x * ((u128)2 - o * x);
      lo[k] = (unsigned long long)x;
      hi[k] = (unsigned long long)(x >> 64);
    }
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(dev::c_inv_lo), lo, sizeof(lo)));
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(dev::c_inv_hi), hi, sizeof(hi)));
    g_inv_ready = true;
  }
  return 0;
}

int device_graph_upload(const HostGraph& g, DeviceGraph* dg, std::string* err)
{
  dg->n = g.n;
  dg->m = g.m;
  dg->maxdeg = 0;
  for (int32_t u = 0; u < g.n; u++) dg->maxdeg = std::max(dg->maxdeg, g.nindex[u + 1] - g.nindex[u]);
  HIP_TRY(hipMalloc((void**)&dg->idx, sizeof(int32_t) * ((size_t)g.n + 1)));
  HIP_TRY(hipMalloc((void**)&dg->adj, sizeof(int32_t) * std::max<size_t>(1, (size_t)g.m)));
  HIP_TRY(hipMemcpy(dg->idx, g.nindex.data(), sizeof(int32_t) * ((size_t)g.n + 1), hipMemcpyHostToDevice));
  if (g.m) HIP_TRY(hipMemcpy(dg->adj, g.nlist.data(), sizeof(int32_t) * (size_t)g.m, hipMemcpyHostToDevice));
  dg->h_idx = g.nindex;
  return 0;
}

void device_graph_free(DeviceGraph* dg)
{
  if (dg->idx) (void)hipFree(dg->idx);
  if (dg->adj) (void)hipFree(dg->adj);
  dg->idx = dg->adj = nullptr;
}

// Cuts the roots of one job into work items. Edge-rooted classes: consecutive CSR
// positions, at most `chunk` per item, and for vertices whose per-position work grows
// with their degree, fewer positions per item (roughly chunk * 64 / degree).
static void cut_items(const JobDesc& j, const DeviceGraph& g, int jid, const Tune& t, std::vector<WorkItem>* out)
{
  if (j.cls == kVertex) {
    if (j.kind != kVertexStars) return;
    const int64_t step = std::max<int64_t>(64, (int64_t)t.chunk * 16);
    for (int64_t b = 0; b < g.n; b += step) out->push_back(WorkItem{jid, 0, 0, 0, b, std::min<int64_t>(g.n, b + step)});
    return;
  }
  const bool heavy = (j.cls == kWedge || j.cls == kGeneral);
  const std::vector<int32_t>& idx = g.h_idx;
  int64_t cur_b = -1, cur_n = 0;
  int32_t cur_u0 = 0, cur_u1 = 0;
  auto emit = [&]() {
    if (cur_b >= 0) out->push_back(WorkItem{jid, cur_u0, cur_u1, 0, cur_b, idx[cur_u1 + 1]});
    cur_b = -1;
    cur_n = 0;
  };
  for (int32_t u = 0; u < g.n; u++) {
    const int64_t b = idx[u], e = idx[u + 1];
    const int64_t d = e - b;
    if (d == 0) continue;
    int64_t per = t.chunk;
    if (heavy && d > 64) per = std::max<int64_t>(1, (int64_t)t.chunk * 64 / d);
    if (d > per || d > t.hub_degree) {
      emit();
      for (int64_t s = b; s < e; s += per) out->push_back(WorkItem{jid, u, u, 0, s, std::min(e, s + per)});
      continue;
    }
    if (cur_b < 0) { cur_b = b; cur_u0 = u; }
    cur_u1 = u;
    cur_n += d;
    if (cur_n >= per) emit();
  }
  emit();
}

int device_run_jobs(std::vector<JobDesc>& jobs, const std::vector<const DeviceGraph*>& graphs, const Tune& t,
                    void* stream_v, std::vector<JobResult>* out, double* kernel_ms, std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int nj = (int)jobs.size();
  out->assign(nj, JobResult());
  if (kernel_ms) *kernel_ms = 0;
  if (nj == 0) return 0;
  std::vector<WorkItem> items;
  for (int k = 0; k < nj; k++) {
    jobs[k].idx = graphs[k]->idx;
    jobs[k].adj = graphs[k]->adj;
    jobs[k].n = graphs[k]->n;
    jobs[k].m = (int32_t)graphs[k]->m;
    cut_items(jobs[k], *graphs[k], k, t, &items);
  }
  // most expensive items first: order by job, big graphs first
  Accum* d_acc = nullptr;
  JobDesc* d_jobs = nullptr;
  WorkItem* d_items = nullptr;
  int* d_head = nullptr;
  const size_t acc_bytes = sizeof(Accum) * nj;
  HIP_TRY(hipMalloc((void**)&d_acc, acc_bytes + 256));
  d_head = (int*)((char*)d_acc + acc_bytes + 128);
  for (int k = 0; k < nj; k++) jobs[k].out = d_acc + k;
  HIP_TRY(hipMalloc((void**)&d_jobs, sizeof(JobDesc) * nj));
  HIP_TRY(hipMalloc((void**)&d_items, sizeof(WorkItem) * std::max<size_t>(1, items.size())));
  HIP_TRY(hipMemsetAsync(d_acc, 0, acc_bytes + 256, stream));
  HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
  if (!items.empty())
    HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, stream));
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int grid = t.grid > 0 ? t.grid : cus * 8;
  const int nitems = (int)items.size();
  grid = std::max(1, std::min(grid, (nitems + dev::kWavesPerBlock - 1) / dev::kWavesPerBlock));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, stream));
  if (nitems > 0) {
    hipLaunchKernelGGL(dev::k_count, dim3(grid), dim3(dev::kBlock), 0, stream, d_jobs, d_items, nitems, d_head);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(e1, stream));
  std::vector<Accum> h_acc(nj);
  HIP_TRY(hipMemcpyAsync(h_acc.data(), d_acc, acc_bytes, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  if (kernel_ms) *kernel_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_items);
  (void)hipFree(d_jobs);
  (void)hipFree(d_acc);
  for (int k = 0; k < nj; k++) {
    JobResult& r = (*out)[k];
    r.sum = ((unsigned __int128)h_acc[k].hi << 64) | h_acc[k].lo;
    r.matches = h_acc[k].matches;
    r.flags = h_acc[k].flags;
  }
  return 0;
}

}  // namespace fsgc
