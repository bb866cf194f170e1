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

#include "devmath.h"
#include "devscratch.h"
#include "fsgc_plan.h"

namespace fsgc {
namespace dev {

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
// With sub_core = false the core vertices are left in (the prefix counts of
// coop_leaf_venn).
__device__ void coop_commonx(const G& g, int nodes, int nonanch, const int* st, int* venn, bool sub_core = true)
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
  if (sub_core && lane_id() < nodes) {
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

// coop_commonx of a full match from the counts P of its prefix (core vertices
// 0 .. nodes - 2, core vertices left in) and the last core vertex x: only N(x) is
// classified, by its prefix pattern p into cnt[p]; then venn[S | x] = cnt[S] and
// venn[S] = P[S] - cnt[S] for S within the prefix (S != 0), and the core vertices are
// subtracted as in coop_commonx. When x anchors no fringe, venn = P minus the core.
__device__ void coop_leaf_venn(const G& g, int nodes, int nonanch, const int* st, const int* P, int* cnt, int* venn)
{
  const int lane = lane_id();
  const int leaf = nodes - 1, lb = 1 << leaf;
  const bool leaf_anch = !((nonanch >> leaf) & 1);
  if (leaf_anch) {
    cnt[lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int a = st[leaf];
    for (int e = g.beg(a) + lane; e < g.end(a); e += kWave) {
      const int y = g.adj[e];
      int pm = 0;
      for (int k = 0; k < leaf; k++) {
        if ((nonanch >> k) & 1) continue;
        if (contains(g.adj, g.beg(st[k]), g.end(st[k]), y)) pm |= 1 << k;
      }
      atomicAdd(&cnt[pm], 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // lane S: the region of anchor set S (nodes <= 6, so every region has a lane)
  int v = 0;
  if (lane != 0 && lane < (1 << nodes)) {
    if (lane & lb) v = leaf_anch ? cnt[lane & ~lb] : 0;
    else v = P[lane] - (leaf_anch ? cnt[lane] : 0);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  venn[lane] = v;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (lane < nodes) {
    const int x = st[lane];
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

// LDS staging of a prefix's anchor lists: at most kPrefixCap ints per wave, after
// kPrefixMeta ints of ranges
constexpr int kPrefixCap = 2048;
constexpr int kPrefixMeta = 16;

// Copies the neighbour lists of the prefix's anchored core vertices (0 .. pn - 1) to
// `lst` when together they fit kPrefixCap ints: list k at [meta[2k], meta[2k + 1]).
// Lane c < pn gets in `ppc` the prefix anchors adjacent to core vertex c.
// Wave-cooperative, uniform arguments; false (nothing staged) when the lists do not fit.
__device__ bool stage_prefix(const G& g, int pn, int nonanch, const int* st, int* meta, int* lst, int& ppc)
{
  int tot = 0;
  for (int k = 0; k < pn; k++)
    if (!((nonanch >> k) & 1)) tot += g.deg(st[k]);
  if (tot > kPrefixCap) return false;
  int off = 0;
  for (int k = 0; k < pn; k++) {
    int b = 0, d = 0;
    if (!((nonanch >> k) & 1)) { b = g.beg(st[k]); d = g.end(st[k]) - b; }
    for (int i = lane_id(); i < d; i += kWave) lst[off + i] = g.adj[b + i];
    if (lane_id() == 0) { meta[2 * k] = off; meta[2 * k + 1] = off + d; }
    off += d;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  ppc = 0;
  if (lane_id() < pn) {
    const int c = st[lane_id()];
    for (int k = 0; k < pn; k++)
      if (!((nonanch >> k) & 1) && contains(lst, meta[2 * k], meta[2 * k + 1], c)) ppc |= 1 << k;
  }
  return true;
}

// coop_leaf_venn with the prefix's anchor lists staged in LDS (stage_prefix): N(x) is
// classified by LDS searches, and x's adjacency to the prefix's core vertices (for the
// core subtraction) is read off the same walk; ppc (lane c) is core vertex c's
// prefix pattern. Same result as coop_leaf_venn.
__device__ void coop_leaf_venn_lds(const G& g, int nodes, int nonanch, const int* st, const int* P, int* cnt,
                                   int* venn, const int* meta, const int* lst, int ppc)
{
  const int lane = lane_id();
  const int leaf = nodes - 1, lb = 1 << leaf;
  const bool leaf_anch = !((nonanch >> leaf) & 1);
  const int pmask = ~nonanch & (lb - 1);
  const int x = st[leaf];
  int adjx = 0;  // bit k: x adjacent to core vertex k (k < leaf)
  if (leaf_anch) {
    cnt[lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    int mine = 0;
    for (int e = g.beg(x) + lane; e < g.end(x); e += kWave) {
      const int y = g.adj[e];
      int pm = 0;
      for (int k = 0; k < leaf; k++) {
        if (y == st[k]) mine |= 1 << k;
        if (((pmask >> k) & 1) && contains(lst, meta[2 * k], meta[2 * k + 1], y)) pm |= 1 << k;
      }
      atomicAdd(&cnt[pm], 1);
    }
    for (int k = 0; k < leaf; k++)
      if (__ballot((mine >> k) & 1)) adjx |= 1 << k;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    for (int k = 0; k < leaf; k++)
      if (((pmask >> k) & 1) && contains(lst, meta[2 * k], meta[2 * k + 1], x)) adjx |= 1 << k;
  }
  int v = 0;
  if (lane != 0 && lane < (1 << nodes)) {
    if (lane & lb) v = leaf_anch ? cnt[lane & ~lb] : 0;
    else v = P[lane] - (leaf_anch ? cnt[lane] : 0);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  venn[lane] = v;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // each core vertex leaves the region of its adjacent anchors
  int bmp = 0;
  if (lane < leaf) bmp = ppc | ((leaf_anch && ((adjx >> lane) & 1)) ? lb : 0);
  else if (lane == leaf) bmp = adjx & pmask;
  if (bmp) atomicSub(&venn[bmp], 1);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ ull dp_mul(ull a, ull b, uint32_t&) { return a * b; }
__device__ __forceinline__ ull dp_add(ull a, ull b, uint32_t&) { return a + b; }
__device__ __forceinline__ u128 dp_mul(u128 a, u128 b, uint32_t& fl) { return mulx(a, b, fl); }
__device__ __forceinline__ u128 dp_add(u128 a, u128 b, uint32_t& fl) { return addx(a, b, fl); }

// The region DP of functionx_dp in words of type T (ull only where no value can wrap):
// `old` and `bin` are 64 T of LDS each.
template <typename T>
__device__ __noinline__ T region_dp(const JobDesc* __restrict__ J, const int* venn, T* old, T* bin, uint32_t& fl)
{
  const int lane = lane_id();
  const int ns = J->dp_ns, nreg = J->dp_nreg;
  constexpr ull kGuard = 0x8080808080808080ull;
  const ull dig_t = lane < ns ? J->dp_dig[lane] : 0;
  T cur = lane == 0 ? (T)1 : (T)0;
  for (int r = 0; r < nreg; r++) {
    const int nr = venn[J->dp_reg[r]];
    if (nr == 0) continue;
    const int maxj = J->dp_maxj[r];
    old[lane] = cur;
    bin[lane] = lane <= maxj ? (T)ncr(nr, lane, fl) : (T)0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const ull allow = J->dp_allow[r];
    T acc = 0;
    for (int d = 0; d < ns; d++) {
      const ull dd = J->dp_dig[d];
      if (dd & ~allow) continue;
      if (J->dp_sum[d] > maxj) continue;
      if (lane < ns && (((dig_t | kGuard) - dd) & kGuard) == kGuard) {
        const T coef = dp_mul(bin[J->dp_sum[d]], (T)J->dp_mult[d], fl);
        acc = dp_add(acc, dp_mul(old[lane - d], coef, fl), fl);
      }
    }
    cur = acc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  old[lane] = cur;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const T res = old[ns - 1];
  __builtin_amdgcn_wave_barrier();
  return res;
}

// The fringe count of one general-core match by the region DP of the job's tables,
// wave-cooperative: lane t holds state t. `venn` is the match's Venn vector and `tab`
// 128 u128 of LDS per wave. Returns the same value as functionx on every lane; a
// negative region that an anchor set may draw from makes it 0, as the reference's
// nCk(n < 0, k) does. With N vertices in the regions drawn from and F fringes, every
// DP value, binomial and product is at most N^F (a count of choices of at most F of
// the N vertices), so the DP runs in u64 when N^F < 2^64; the binomials are the same
// calls either way, so the reference-compatibility flags are too.
__device__ u128 functionx_dp(const JobDesc* __restrict__ J, const int* venn, u128* tab, uint32_t& fl)
{
  const int ns = J->dp_ns, nreg = J->dp_nreg;
  bool neg = false;
  ull n_all = 0;
  for (int r = 0; r < nreg; r++) {
    const int nr = venn[J->dp_reg[r]];
    neg = neg || nr < 0;
    n_all += nr > 0 ? (ull)nr : 0;
  }
  if (neg) return 0;
  const int nf = J->dp_sum[ns - 1];
  bool fits = true;
  ull pw = 1;
  for (int i = 0; i < nf && fits; i++) {
    if (__umul64hi(pw, n_all) != 0) fits = false;
    else pw *= n_all;
  }
  if (fits) return (u128)region_dp<ull>(J, venn, (ull*)tab, (ull*)tab + 64, fl);
  return region_dp<u128>(J, venn, tab, tab + 64, fl);
}

struct Stack6 { int v[6]; };

// FC:489-564 (general_core_level<L>)
template <int L>
__device__ void gen_level(const JobDesc* __restrict__ J, const G& g, Stack6 st, LaneAcc& acc, int* slots,
                          int parts = 0, int part = 0)
{
  const int nodes = J->p[0];
  const int list = J->p[7 + L];
  const int mindeg = J->p[2 + L];
  const int nonanch = J->p[12];
  const int wn = st.v[__builtin_ctz(list)];
  const int b = g.beg(wn), e = g.end(wn);
  const bool leaf = (L == nodes - 2);
  // the prefix counts of coop_leaf_venn and the staged prefix lists, set up on the
  // first match of this prefix
  int* P = slots + kVennSlots * 64 + 512;
  int* cnt = P + 64;
  int* meta = cnt + 64;
  int* lst = meta + kPrefixMeta;
  bool have_p = false, staged = false;
  int ppc = 0;
  for (int cb = b; cb < e; cb += kWave) {
    const int i = cb + lane_id();
    int x = -1;
    bool a = i < e;
    // an item that holds one part of the first level's candidates (WorkItem::split)
    if constexpr (L == 1) a = a && (parts <= 1 || (i - b) % parts == part);
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
          if (!have_p) {
            coop_commonx(g, nodes - 1, nonanch, st.v, P, false);
            staged = stage_prefix(g, nodes - 1, nonanch, st.v, meta, lst, ppc);
            have_p = true;
          }
          if (staged) coop_leaf_venn_lds(g, nodes, nonanch, s2.v, P, cnt, slots + slot * 64, meta, lst, ppc);
          else coop_leaf_venn(g, nodes, nonanch, s2.v, P, cnt, slots + slot * 64);
          if (lane_id() == w2) myslot = slot;
          slot++;
        }
        if (J->dp_ns > 0) {
          // polynomial region DP, one staged match at a time across the wave
          u128* tab = (u128*)(slots + kVennSlots * 64);
          for (int q = 0; q < slot; q++) {
            const u128 v = functionx_dp(J, slots + q * 64, tab, acc.fl);
            if (lane_id() == 0) { acc.add(v); acc.matches++; }
          }
        } else if (myslot >= 0) {
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
  const int parts = it.split >> 8, part = it.split & 255;
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
      gen_level<1>(J, g, st, acc, slots, parts, part);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_count(const JobDesc* __restrict__ jobs, const WorkItem* __restrict__ items,
                                                  int nitems, int* __restrict__ head)
{
  // per wave: kVennSlots staged Venn vectors, 128 u128 of region-DP state, a prefix's
  // counts and leaf histogram (64 + 64 ints), then its staged anchor lists
  __shared__ __attribute__((aligned(16))) int s_venn[kWavesPerBlock][kVennSlots * 64 + 128 * 4 + 128 + kPrefixMeta +
                                                                     kPrefixCap];
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

// zeroes n 64-bit words (the accumulators of a launch; a kernel rather than a memset
// so that a captured graph holds no memset node)
__global__ void k_zero64(ull* __restrict__ p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
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

int device_init(int device, std::string* err)
{
  HIP_TRY(hipSetDevice(device));
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    if (err) *err = "no such GPU";
    return -19;
  }
  // keep freed call temporaries in the device's default pool for the next call
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
    uint64_t keep = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  return 0;
}

LaunchPlan& LaunchPlan::operator=(LaunchPlan&& o) noexcept
{
  if (this != &o) {
    if (h_acc) (void)hipHostFree(h_acc);
    nj = o.nj; d_acc = o.d_acc; acc_bytes = o.acc_bytes; h_acc = o.h_acc; index_ms = o.index_ms;
    nitems = o.nitems; head = o.head; neg = o.neg;
    kernel = std::move(o.kernel);
    post = std::move(o.post);
    scratch = std::move(o.scratch);
    o.h_acc = nullptr;
    o.d_acc = nullptr;
  }
  return *this;
}

LaunchPlan::~LaunchPlan()
{
  if (h_acc) (void)hipHostFree(h_acc);
}

int plan_accumulators(LaunchPlan* lp, int nj, void* stream, std::string* err, bool neg)
{
  if (!lp->scratch) lp->scratch = std::make_shared<DevScratch>((hipStream_t)stream);
  lp->nj = nj;
  lp->neg = neg;
  const int nacc = neg ? 2 * nj : nj;
  lp->acc_bytes = sizeof(Accum) * nacc + 256;  // the work-queue head lives past the accumulators
  HIP_TRY(lp->scratch->get(&lp->d_acc, lp->acc_bytes));
  lp->head = (int*)((char*)lp->d_acc + sizeof(Accum) * nacc + 128);
  HIP_TRY(hipHostMalloc((void**)&lp->h_acc, sizeof(Accum) * std::max(1, nacc), hipHostMallocDefault));
  return 0;
}

static size_t acc_copy_bytes(const LaunchPlan& lp) { return sizeof(Accum) * (lp.neg ? 2 * lp.nj : lp.nj); }

int launch_enqueue(const LaunchPlan& lp, void* stream_v, void* ev_begin, void* ev_end, std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int words = (int)(lp.acc_bytes / 8);
  hipLaunchKernelGGL(dev::k_zero64, dim3((words + 255) / 256), dim3(256), 0, stream, (dev::ull*)lp.d_acc, words);
  HIP_TRY(hipGetLastError());
  if (ev_begin) HIP_TRY(hipEventRecord((hipEvent_t)ev_begin, stream));
  if (lp.kernel && lp.nitems > 0) {
    lp.kernel(stream, 0, lp.nitems, lp.head);
    HIP_TRY(hipGetLastError());
  }
  if (lp.post) {
    lp.post(stream);
    HIP_TRY(hipGetLastError());
  }
  if (ev_end) HIP_TRY(hipEventRecord((hipEvent_t)ev_end, stream));
  HIP_TRY(hipMemcpyAsync(lp.h_acc, lp.d_acc, acc_copy_bytes(lp), hipMemcpyDeviceToHost, stream));
  return 0;
}

void launch_collect(const LaunchPlan& lp, std::vector<JobResult>* out)
{
  out->assign(lp.nj, JobResult());
  for (int k = 0; k < lp.nj; k++) {
    JobResult& r = (*out)[k];
    r.sum = ((unsigned __int128)lp.h_acc[k].hi << 64) | lp.h_acc[k].lo;
    r.matches = lp.h_acc[k].matches;
    r.flags = lp.h_acc[k].flags;
    r.bytes = lp.h_acc[k].bytes;
    if (lp.neg) {
      const Accum& a = lp.h_acc[lp.nj + k];
      const unsigned __int128 ng = ((unsigned __int128)a.hi << 64) | a.lo;
      r.flags |= a.flags;
      // the negative part never exceeds the sum unless a part wrapped at 2^128
      if (r.sum < ng) r.flags |= kOvf128;
      r.sum -= ng;
    }
    r.index_ms = (k == 0) ? lp.index_ms : 0;
  }
}

static int segment_items()
{
  static const int k = [] { const char* e = getenv("FSGC_SEGMENT_ITEMS"); return e ? atoi(e) : 0; }();
  return k;
}

int launch_run(const LaunchPlan& lp, void* stream_v, std::vector<JobResult>* out, double* kernel_ms, std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int seg = segment_items();
  if (seg <= 0 || lp.nitems <= seg || !lp.kernel) {
    EventPair ev;
    HIP_TRY(ev.create());
    int rc = launch_enqueue(lp, stream_v, ev.a, ev.b, err);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    if (kernel_ms) *kernel_ms = ms;
    launch_collect(lp, out);
    return 0;
  }
  // the same items in consecutive launches, each with its own queue head
  const int nseg = std::min(32, (lp.nitems + seg - 1) / seg);
  const int words = (int)(lp.acc_bytes / 8);
  hipLaunchKernelGGL(dev::k_zero64, dim3((words + 255) / 256), dim3(256), 0, stream, (dev::ull*)lp.d_acc, words);
  HIP_TRY(hipGetLastError());
  double total = 0;
  for (int s = 0; s < nseg; s++) {
    const int first = (int)((long long)lp.nitems * s / nseg), last = (int)((long long)lp.nitems * (s + 1) / nseg);
    EventPair ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.a, stream));
    lp.kernel(stream, first, last - first, lp.head + s);
    HIP_TRY(hipGetLastError());
    if (s == nseg - 1 && lp.post) {
      lp.post(stream);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ev.b, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    total += ms;
    const char* lg = getenv("FSGC_LOG");
    if (lg && *lg && *lg != '0') fprintf(stderr, "[fsgc]   segment %d/%d: %d items, %.1f ms\n", s + 1, nseg, last - first, ms);
  }
  HIP_TRY(hipMemcpyAsync(lp.h_acc, lp.d_acc, acc_copy_bytes(lp), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (kernel_ms) *kernel_ms = total;
  launch_collect(lp, out);
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

// A small graph (a motif graph for the automorphism run) in call temporaries on `stream`.
int device_graph_upload_async(const HostGraph& g, DeviceGraph* dg, DevScratch* sc, std::string* err)
{
  dg->n = g.n;
  dg->m = g.m;
  dg->maxdeg = 0;
  for (int32_t u = 0; u < g.n; u++) dg->maxdeg = std::max(dg->maxdeg, g.nindex[u + 1] - g.nindex[u]);
  HIP_TRY(sc->get(&dg->idx, sizeof(int32_t) * ((size_t)g.n + 1)));
  HIP_TRY(sc->get(&dg->adj, sizeof(int32_t) * std::max<size_t>(1, (size_t)g.m)));
  HIP_TRY(hipMemcpyAsync(dg->idx, g.nindex.data(), sizeof(int32_t) * ((size_t)g.n + 1), hipMemcpyHostToDevice,
                         sc->stream));
  if (g.m)
    HIP_TRY(hipMemcpyAsync(dg->adj, g.nlist.data(), sizeof(int32_t) * (size_t)g.m, hipMemcpyHostToDevice, sc->stream));
  dg->h_idx = g.nindex;
  return 0;
}

void device_graph_free(DeviceGraph* dg)
{
  device_derived_free(dg);
  if (dg->idx) (void)hipFree(dg->idx);
  if (dg->adj) (void)hipFree(dg->adj);
  dg->idx = dg->adj = nullptr;
}

// Cuts the roots of one job into work items. Edge-rooted classes: consecutive CSR
// positions, at most `chunk` per item, and for vertices whose per-position work grows
// with their degree, fewer positions per item (roughly chunk * 64 / degree).
// a general core's launch is cut into about this many items of equal estimated work
constexpr int64_t kGeneralItems = 1 << 20;

template <class Emit>
static void cut_items_with(const JobDesc& j, const DeviceGraph& g, int jid, const Tune& t, Emit&& put)
{
  if (j.cls == kVertex) {
    if (j.kind != kVertexStars) return;
    const int64_t step = std::max<int64_t>(64, (int64_t)t.chunk * 16);
    for (int64_t b = 0; b < g.n; b += step) put(WorkItem{jid, 0, 0, 0, b, std::min<int64_t>(g.n, b + step)});
    return;
  }
  const bool heavy = (j.cls == kWedge || j.cls == kGeneral);
  const std::vector<int32_t>& idx = g.h_idx;
  // a general core's root edge (u, v) extends over about d_u^2 further core matches, each
  // with its own Venn vector: items hold about `wt` such units, so that the longest item
  // is a small share of the launch (a hub root edge's first-level candidates are split
  // over up to 64 items)
  int64_t wt = 0;
  if (j.cls == kGeneral) {
    double total = 0;
    for (int32_t u = 0; u < g.n; u++) {
      const double d = idx[u + 1] - idx[u];
      total += d * d * d;
    }
    wt = std::max<int64_t>((int64_t)t.chunk * 4, (int64_t)(total / kGeneralItems));
  }
  int64_t cur_b = -1, cur_n = 0;
  int32_t cur_u0 = 0, cur_u1 = 0;
  auto emit = [&]() {
    if (cur_b >= 0) put(WorkItem{jid, cur_u0, cur_u1, 0, cur_b, idx[cur_u1 + 1]});
    cur_b = -1;
    cur_n = 0;
  };
  for (int32_t u = 0; u < g.n; u++) {
    const int64_t b = idx[u], e = idx[u + 1];
    const int64_t d = e - b;
    if (d == 0) continue;
    int64_t per = t.chunk;
    if (heavy && d > 64) per = std::max<int64_t>(1, (int64_t)t.chunk * 64 / d);
    int parts = 1;
    if (j.cls == kGeneral) {
      if (d * d <= wt) per = std::max<int64_t>(1, std::min<int64_t>(per, wt / (d * d)));
      else {
        per = 1;
        parts = (int)std::min<int64_t>(std::min<int64_t>(64, d), (d * d + wt - 1) / wt);
      }
    }
    if (d > per || d > t.hub_degree) {
      emit();
      for (int64_t s = b; s < e; s += per)
        for (int q = 0; q < parts; q++)
          put(WorkItem{jid, u, u, parts > 1 ? (parts << 8) | q : 0, s, std::min(e, s + per)});
      continue;
    }
    if (cur_b < 0) { cur_b = b; cur_u0 = u; }
    cur_u1 = u;
    cur_n += d;
    if (cur_n >= per) emit();
  }
  emit();
}

static void cut_items(const JobDesc& j, const DeviceGraph& g, int jid, const Tune& t, std::vector<WorkItem>* out)
{
  cut_items_with(j, g, jid, t, [out](const WorkItem& w) { out->push_back(w); });
}

size_t perroot_scratch_bytes(const DeviceGraph& g, const Tune& t)
{
  // work items: at most one per `chunk` CSR positions (per 4) plus one per split hub
  // piece, or the items of a general core (cut by work), counted here
  const size_t per = (size_t)std::max(1, std::min(t.chunk, 4));
  JobDesc gj{};
  gj.cls = kGeneral;
  size_t ngen = 0;
  cut_items_with(gj, g, 0, t, [&ngen](const WorkItem&) { ngen++; });
  return sizeof(WorkItem) * std::max((size_t)g.m / per + (size_t)g.n + 64, ngen) + 4096;
}

int plan_jobs(std::vector<JobDesc>& jobs, const std::vector<const DeviceGraph*>& graphs, const Tune& t,
              void* stream_v, LaunchPlan* lp, std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int nj = (int)jobs.size();
  std::vector<WorkItem> items;
  for (int k = 0; k < nj; k++) {
    jobs[k].idx = graphs[k]->idx;
    jobs[k].adj = graphs[k]->adj;
    jobs[k].n = graphs[k]->n;
    jobs[k].m = (int32_t)graphs[k]->m;
    cut_items(jobs[k], *graphs[k], k, t, &items);
  }
  sample_items(&items, t.sample);
  int rc = plan_accumulators(lp, nj, stream, err);
  if (rc) return rc;
  JobDesc* d_jobs = nullptr;
  WorkItem* d_items = nullptr;
  for (int k = 0; k < nj; k++) jobs[k].out = lp->d_acc + k;
  HIP_TRY(lp->scratch->get(&d_jobs, sizeof(JobDesc) * std::max(1, nj)));
  HIP_TRY(lp->scratch->get(&d_items, sizeof(WorkItem) * std::max<size_t>(1, items.size())));
  if (nj) HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
  if (!items.empty())
    HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, stream));
  // the host arrays may go away before the copies run
  HIP_TRY(hipStreamSynchronize(stream));
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int grid = t.grid > 0 ? t.grid : cus * 8;
  const int nitems = (int)items.size();
  grid = std::max(1, std::min(grid, (nitems + dev::kWavesPerBlock - 1) / dev::kWavesPerBlock));
  lp->nitems = nitems;
  lp->kernel = [=](void* s, int first, int count, int* head) {
    hipLaunchKernelGGL(dev::k_count, dim3(grid), dim3(dev::kBlock), 0, (hipStream_t)s, d_jobs, d_items + first, count,
                       head);
  };
  return 0;
}

int device_run_jobs(std::vector<JobDesc>& jobs, const std::vector<const DeviceGraph*>& graphs, const Tune& t,
                    void* stream, std::vector<JobResult>* out, double* kernel_ms, std::string* err)
{
  out->assign(jobs.size(), JobResult());
  if (kernel_ms) *kernel_ms = 0;
  if (jobs.empty()) return 0;
  LaunchPlan lp;
  int rc = plan_jobs(jobs, graphs, t, stream, &lp, err);
  return rc ? rc : launch_run(lp, stream, out, kernel_ms, err);
}

}  // namespace fsgc
