// Wedge-core counting on the pivot schedule (gfx950).
//
// A wedge of the reference (FC:1513-1650, FC:1409-1477) is a centre u with two
// neighbours v > w; its fringe formula needs d_u, d_v, d_w, the edge supports
// t_uv = |N(u) & N(v)| and t_uw, the codegree c_vw = |N(v) & N(w)|, the triple count
// c3 = |N(u) & N(v) & N(w)| and whether v ~ w. The reference walks N(w) per wedge
// with two binary searches per element (sum over wedges of d_w probes). Here:
//   * t_e for every directed edge is computed once per graph (support pass), and so
//     are the per-edge common-neighbour lists T(u,x) as positions in N(u);
//   * a workgroup takes a pivot w (the smaller endpoint) and scatters the 2-hop
//     counts c(w, y) for y > w into a dense per-workgroup array D (plus an adjacency
//     bit, and the item's epoch so that D needs no clearing pass), so every wedge
//     with endpoint w reads c_vw and v ~ w with one load;
//   * for each centre u of w, c3(u, v, w) for all v is the number of x in T(u,w)
//     with v in T(u,x): a scatter over the precomputed lists into LDS counters E
//     indexed by position in N(u), one window of kECap positions at a time (hub
//     centres take several windows; each x's list is cut to the window by two
//     binary searches); the cut lists of a wave's 64 x's are streamed as one
//     sequence, so short and long lists alike keep eight loads per lane in flight;
//   * a wave then walks the window and evaluates every wedge job per lane; the walk
//     is software-pipelined (list loads two steps ahead, 2-hop counter loads one);
//   * the positions of w in N(x) and N(u) come from the per-graph reverse-edge index;
//   * within a workgroup the waves share the work of one pivot through LDS task
//     lists: neighbours x with long lists (2-hop scatter) and centres u with several
//     windows (c3 scatter and walk) are cut into chunks that any wave takes.
// The sum over wedges is the reference's, with the same per-wedge filters, formulas
// and counted matches; only the order of evaluation differs. Jobs whose count is
// linear in the triple count (f fringes on the centre, one on each endpoint) skip c3:
// the kernel sums the rest of their formula per wedge and k_wedge_lin_centres adds the
// c3 term per centre (eval_wedge_lin).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "devmath.h"
#include "devscratch.h"
#include "fsgc_plan.h"

namespace fsgc {
namespace dev {

// launch shapes of the pivot kernel: waves per workgroup and LDS c3 counters per wave
// (positions of N(u)); 4 x 4096 (64 KB, two workgroups per CU), 8 x 4096 and
// 16 x 2048 (128 KB, one workgroup per CU with 8 or 16 waves sharing one D array);
// launches without c3 use a 16-wave shape whose LDS holds hub counters instead
constexpr int kPivWavesMax = 16;
constexpr int kMaxPivotJobs = 8;
// Hubs: the kHubSlots highest-degree vertices, where most 2-hop paths end (about half
// of them on R-MAT s20). Launches without c3 need no LDS windows, so their 16-wave
// shape keeps the hub counters in LDS (160 KB) and only the other vertices' counters
// in the global per-workgroup array D. (Launches with c3 keep their 128 KB of windows:
// 8184 hubs in the remaining LDS measured 2% slower than none, the lanes of a wave
// diverging between the two kinds of counter.)
constexpr int kHubSlots = 39552;
// Phase B defers centres whose walk spans several windows (kECap positions with c3,
// kWalkChunk without) to a task list shared by the workgroup's waves, up to kHeavyCap
// per item; the hub counters above leave LDS room for that list (4 KB)
constexpr int kHeavyCap = 512;
constexpr int kWalkChunk = 2048;
// phase A defers an x whose list past w is longer than kAChunk positions the same way
constexpr int kAChunk = 2048;

struct PivotItem {
  int32_t w;   // pivot
  int32_t p0;  // centre positions [p0, p1) within N(w)
  int32_t p1;
  int32_t pad;
};

// lower_bound of x in adj[l, r)
__device__ __forceinline__ int lower_bound_pos(const int* __restrict__ a, int l, int r, int x)
{
  while (l < r) {
    const int mid = (l + r) >> 1;
    if (a[mid] < x) l = mid + 1; else r = mid;
  }
  return l;
}

// ---------------------------------------------------------------- support pass

// Per-edge common neighbours, built once per graph: the supports t(u,x) = |N(u) & N(x)|
// (count pass) and the lists T(u,x) of the positions in N(u) of those common
// neighbours, ascending (list pass, after a scan of the supports gives the offsets).
//
// Each undirected edge {u, x} is intersected once, from its endpoint of higher rank
// (degree, then id), u: a wave stages N(u) in LDS (every stride-th id of it when it is
// longer than kComStage, the search then ending in one block of the global list) and
// streams the lists N(x) of u's lower-ranked neighbours x, packed one after another
// across its 64 lanes, so short lists share a wave and long ones are read coalesced.
// Each element y at position j of N(x) is searched in the staged N(u); a hit at
// position i is entry r of both T(x,u) (= j) and T(u,x) (= i), r being the number of
// earlier hits of the pair: both lists ascend with y. Cost: the shorter list of every
// edge streamed once, with log2(d_u) LDS probes per element.
constexpr int kComStage = 2048;  // ids of N(u) staged per wave (8 KB)
constexpr int kComBlock = 256;
constexpr int kComWaves = kComBlock / kWave;

struct ComItem {
  int32_t u;
  int32_t p0;  // positions [p0, p1) of N(u) whose lower-ranked neighbours this item takes
  int32_t p1;
  int32_t pad;
};

template <bool kList>
__global__ __launch_bounds__(kComBlock) void k_common(const int* __restrict__ idx, const int* __restrict__ adj,
                                                      const int* __restrict__ adjdeg,
                                                      const ComItem* __restrict__ items, int nitems,
                                                      int* __restrict__ head, int* __restrict__ tsup,
                                                      const long long* __restrict__ tl_off, int* __restrict__ tl,
                                                      int* __restrict__ rev)
{
  __shared__ int s_stage[kComWaves][kComStage];
  // per wave, one entry per lane's x: offset of its list in the packed stream (entry 64
  // = total), list start in adj, hits so far, and (list pass) the output bases
  __shared__ int s_off[kComWaves][kWave + 1];
  __shared__ int s_xb[kComWaves][kWave];
  __shared__ int s_cnt[kComWaves][kWave];
  __shared__ long long s_bux[kComWaves][kList ? kWave : 1];
  __shared__ long long s_bxu[kComWaves][kList ? kWave : 1];
  __shared__ int s_lim[kComWaves][kList ? kWave : 1];  // the pair's support from the count pass
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  int* __restrict__ S = s_stage[wave];
  int* __restrict__ off = s_off[wave];
  int* __restrict__ sxb = s_xb[wave];
  int* __restrict__ cnt = s_cnt[wave];
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(head, 1);
    it = bcast(it, 0);
    if (it >= nitems) break;
    const ComItem I = items[it];
    const int u = I.u, ub = idx[u], du = idx[u + 1] - ub;
    const int stride = (du + kComStage - 1) / kComStage;  // 1 when N(u) fits
    const int ns = (du + stride - 1) / stride;
    __builtin_amdgcn_wave_barrier();  // the previous item's readers of S are done
    for (int k = lane; k < ns; k += kWave) S[k] = adj[ub + k * stride];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // position of y in N(u), or -1
    auto find = [&](int y) -> int {
      if (S[0] > y) return -1;
      int lo = 0, hi = ns;  // S[lo] <= y < S[hi] (S[ns] = +inf)
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S[mid] <= y) lo = mid; else hi = mid;
      }
      if (stride == 1) return S[lo] == y ? lo : -1;
      int l = ub + lo * stride + 1, r = min(ub + (lo + 1) * stride, ub + du);
      if (S[lo] == y) return lo * stride;
      const int r0 = r;
      while (l < r) {
        const int mid = (l + r) >> 1;
        if (adj[mid] < y) l = mid + 1; else r = mid;
      }
      return (l < r0 && adj[l] == y) ? l - ub : -1;
    };
    for (int q0 = I.p0; q0 < I.p1; q0 += kWave) {
      const int p = q0 + lane;
      int x = 0, dx = 0, xb = 0, exu = 0;
      bool take = false;
      if (p < I.p1) {
        x = adj[ub + p];
        dx = adjdeg[ub + p];
        take = dx < du || (dx == du && x < u);  // u ranks higher: the edge is intersected here
      }
      if (take) {
        xb = idx[x];
        int l = xb, r = xb + dx;  // position of u in N(x): the reverse edge (x, u)
        while (l < r) {
          const int mid = (l + r) >> 1;
          if (adj[mid] < u) l = mid + 1; else r = mid;
        }
        exu = l;
      }
      const int len = take ? dx : 0;
      int incl = len;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += t;
      }
      const int tot = __shfl(incl, kWave - 1, kWave);
      off[lane] = incl - len;
      if (lane == kWave - 1) off[kWave] = tot;
      sxb[lane] = xb;
      cnt[lane] = 0;
      if constexpr (kList) {
        // the reverse-edge index: the CSR index of (x, u) for (u, x) and back
        if (take) {
          rev[ub + p] = exu;
          rev[exu] = ub + p;
        }
        s_bux[wave][lane] = take ? tl_off[ub + p] : 0;
        s_bxu[wave][lane] = take ? tl_off[exu] : 0;
        s_lim[wave][lane] = take ? tsup[ub + p] : 0;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // the packed stream: element g belongs to the last lane k with off[k] <= g (an x
      // with an empty list shares its offset with the next one, which wins)
      int k = 0;
      {
        int lo = 0, hi = kWave;  // off[lo] <= g < off[hi]
        const int g0 = min(lane, max(tot - 1, 0));
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (off[mid] <= g0) lo = mid; else hi = mid;
        }
        k = lo;
      }
      for (int c = 0; c < tot; c += kWave) {
        const int gi = c + lane;
        const bool act = gi < tot;
        int i = -1, j = 0;
        if (act) {
          while (off[k + 1] <= gi) k++;
          j = gi - off[k];
          i = find(adj[sxb[k] + j]);
        }
        const bool hit = i >= 0;
        const unsigned long long hm = __ballot(hit);
        // lanes of segment k in this chunk: [first, last]
        const int first = act ? max(off[k] - c, 0) : 0;
        const int last = act ? min(off[k + 1] - c, kWave) - 1 : 0;
        const unsigned long long below = (lane == 0 ? 0ull : (~0ull >> (kWave - lane))) & ~(first == 0 ? 0ull : (~0ull >> (kWave - first)));
        if (kList && hit) {
          const int r = cnt[k] + __popcll(hm & below);
          // the count pass found the same hits, so r < the support; the test keeps every
          // store inside the pair's two lists whatever happens
          if (r < s_lim[wave][k]) {
            tl[s_bux[wave][k] + r] = i;
            tl[s_bxu[wave][k] + r] = j;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (act && lane == last) cnt[k] += __popcll(hm & (below | (1ull << lane)));
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (!kList && take) {
        const int t = cnt[lane];
        tsup[ub + p] = t;
        tsup[exu] = t;
      }
      __builtin_amdgcn_wave_barrier();  // cnt, off and sxb are reused by the next batch
    }
  }
}

struct Widen {
  __host__ __device__ long long operator()(const int& x) const { return (long long)x; }
};

// ---------------------------------------------------------------- pivot kernel

struct WedgeEval {
  int kind, rot, smod, srem;
  int p[11];
};

// One wedge (centre u, endpoints v > w) for one wedge job: filters of FC:1558-1582
// (general) or FC:1445-1457 (specialised), Venn corrections FC:1598-1602 / 1631-1636.
template <bool kFull>
__device__ __forceinline__ u128 fc3_any(const int* fr, const int* venn, uint32_t& fl)
{
  if constexpr (kFull) {
    if (!venn_nonneg(venn)) return fc3(fr, venn, fl);  // the reference's loop, its conventions
    if (!fc3_is_singleton(fr)) {
      // one pair anchor set: the inline placement sum; several: the peeled loops
      const int np = (fr[iuv] != 0) + (fr[iuw] != 0) + (fr[ivw] != 0);
      if (np == 1) return fc3_one_pair(fr, venn, fr[iuv] ? iuv : (fr[iuw] ? iuw : ivw), fl);
      if (np > 1) return fc3_peeled(fr, venn, fl);
    }
  }
  // single-vertex anchor sets, and fringes anchored on all three core vertices (the
  // regions are non-negative for every wedge and triangle that passes the filters)
  return fc3_singletons_uvw(fr, venn, fl);
}

__device__ __forceinline__ bool eval_wedge_special(const JobDesc* __restrict__ J, int du, int dv, int dw, int cvw,
                                                   bool tri, u128& val, uint32_t& fl)
{
  const int* p = J->p;
  if (du < 2 || dv < p[1]) return false;
  const int thres = max(p[10], p[8] + p[9] - dv);
  if (dw < thres) return false;
  val = pair_formula(dv, dw, cvw - 1, tri ? 2 : 1, p[5], p[6], p[7], fl);
  return true;
}

template <bool kFull>
__device__ __forceinline__ bool eval_wedge(const JobDesc* __restrict__ J, int u, int w, int du, int dv, int dw,
                                           int tuv, int tuw, int cvw, int c3, bool tri, u128& val, uint32_t& fl)
{
  const int* p = J->p;
  if (J->kind == kWedgeSpecial) return eval_wedge_special(J, du, dv, dw, cvw, tri, val, fl);
  const int min_u = p[0], lim = p[1], sum1 = p[2], sum2 = p[3], sum3 = p[4];
  if (du < sum1 || dv < max(sum1, sum2 - du)) return false;
  const int thres2 = max(max(sum1, sum3 - du - dv), max(sum2 - du, sum2 - dv));
  if (dw < thres2) return false;
  if (tri) {
    if (!(u < w)) return false;
  } else {
    const bool cond = (du >= min_u) && (dv >= lim);
    if (!(cond && dw >= max(lim, sum2 - dv))) return false;
  }
  int venn[8];
  const int c13 = tuw - c3, c23 = cvw - c3;
  venn[0] = 0;
  venn[iuvw] = c3;
  venn[iuv] = tuv - c3;
  venn[iuw] = c13;
  venn[ivw] = c23;
  venn[iu] = du - (tuv + c13);
  venn[iv] = dv - (tuv + c23);
  venn[iw] = dw - (c13 + c23 + c3);
  u128 s = 0;
  if (!tri) {
    venn[iu] -= 2;
    venn[ivw] -= 1;
    s = fc3_any<kFull>(J->fr[0], venn, fl);
    if (J->rot > 3) s = addx(s, fc3_any<kFull>(J->fr[3], venn, fl), fl);
  } else {
    venn[iuv] -= 1;
    venn[iuw] -= 1;
    venn[ivw] -= 1;
    for (int q = 0; q < J->rot; q++) s = addx(s, fc3_any<kFull>(J->fr[q], venn, fl), fl);
  }
  val = s;
  return true;
}

// A wedge job whose count is linear in the triple region (JobDesc p[11]: f_u fringes on
// the centre, one on each endpoint), without c3: the value is V (wedge_lin_v) of the
// wedge, or of the triangle's three rotations, each with its own centre; the per-centre
// terms -b0 (A + B) + (b0 - b2) s come from k_wedge_lin_centres over every pair of N(u).
// The value is taken for every pair, also those the reference's filters drop: such a
// pair has fc3 = 0, so its V cancels its share of the per-centre terms. The filters
// (the same as eval_wedge's) decide the core-match count and which flags are kept.
// A triangle is taken once, at the rotation whose centre u is its smallest vertex.
// the three rotations of a triangle, each vertex the centre once (out of line: the
// triangles are a small share of the wedges, and inline they cost the walk registers)
__device__ __noinline__ u128 wedge_lin_tri(int f, int du, int dv, int dw, int tuv, int tuw, int cvw, uint32_t& fl)
{
  return addx(addx(wedge_lin_v(f, du - 2, dw - 2, dv - 2, tuw - 1, tuv - 1, cvw - 1, fl),
                   wedge_lin_v(f, dv - 2, du - 2, dw - 2, tuv - 1, cvw - 1, tuw - 1, fl), fl),
              wedge_lin_v(f, dw - 2, du - 2, dv - 2, tuw - 1, cvw - 1, tuv - 1, fl), fl);
}

// cb: the centre's binomials C(d_u - 2 - c, f), c = 0..2, when they are inline u64
// values (cb_ok, the same for the whole walk of a centre), for wedges that are not
// triangles
__device__ __forceinline__ bool eval_wedge_lin(const JobDesc* __restrict__ J, int u, int w, int du, int dv, int dw,
                                               int tuv, int tuw, int cvw, bool tri, u128& val, bool& match,
                                               uint32_t& fl, bool cb_ok, ull cb0, ull cb1, ull cb2)
{
  const int* p = J->p;
  if (tri && !(u < w)) return false;
  const int min_u = p[0], lim = p[1], sum1 = p[2], sum2 = p[3], sum3 = p[4], f = p[12];
  bool pass = du >= sum1 && dv >= max(sum1, sum2 - du) &&
              dw >= max(max(sum1, sum3 - du - dv), max(sum2 - du, sum2 - dv));
  if (pass && !tri) pass = du >= min_u && dv >= lim && dw >= max(lim, sum2 - dv);
  uint32_t lf = 0;
  if (!tri && cb_ok) {
    const int a = dw - 1, b = dv - 1, p = tuw, q = tuv, r = cvw - 1;
    const ull t0 = (ull)((long long)(a - p) * (b - q) - r + a + b);
    const ull t1 = (ull)((long long)p * (b - q) + (long long)q * (a - p));
    const ull t2 = (ull)((long long)p * q);
    val = (u128)cb0 * t0 + (u128)cb1 * t1 + (u128)cb2 * t2;  // as wedge_lin_v's inline case
  } else if (!tri) {
    val = wedge_lin_v(f, du - 2, dw - 1, dv - 1, tuw, tuv, cvw - 1, lf);
  } else {
    val = wedge_lin_tri(f, du, dv, dw, tuv, tuw, cvw, lf);
  }
  fl |= pass ? lf : (lf & kOvf128);
  match = pass;
  return true;
}

// The per-centre terms of the jobs evaluated by eval_wedge_lin: for every centre u (in
// the root slice) with b_c = C(d_u - 2 - c, f), (b0 - b2) S_u goes to the job's sum and
// b0 K_u to its negative part neg[j], where S_u = sum_{x in N(u)} C(t_ux, 2) counts the
// pairs of N(u) inside one N(x) (the sum of s over the pairs of N(u)) and
// K_u = (d_u - 1) sum_x d_x - d_u (d_u - 1) - sum_x t_ux is the sum of A + B over them.
// One wave per vertex; 8 bytes per vertex (its index pair) and per position (support,
// neighbour degree) join the launch's algorithmic bytes.
template <int NJ>
__global__ __launch_bounds__(256) void k_wedge_lin_centres(const int* __restrict__ idx, const int* __restrict__ tsup,
                                                           const int* __restrict__ adjdeg, int n,
                                                           const JobDesc* __restrict__ jobs, int njobs,
                                                           Accum* __restrict__ neg)
{
  const int lane = lane_id();
  const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const int nw = (int)(gridDim.x * blockDim.x / kWave);
  const int smod = jobs[0].smod, srem = jobs[0].srem;
  LaneAcc pos[NJ], ng[NJ];
  ull bytes = 0;
  for (int u = gw; u < n; u += nw) {
    if (smod > 1 && u % smod != srem) continue;
    const int ub = idx[u], du = idx[u + 1] - ub;
    if (lane == 0) bytes += 8 + 8 * (ull)du;
    if (du < 2) continue;
    ull s_lo = 0, s_hi = 0, sd = 0, st = 0;
    for (int k = lane; k < du; k += kWave) {
      const ull t = (ull)tsup[ub + k];
      const ull c = (t * (t - 1)) >> 1;  // t < 2^31
      s_lo += c;
      s_hi += s_lo < c;
      sd += (ull)adjdeg[ub + k];
      st += t;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const ull olo = __shfl_xor(s_lo, o, kWave), ohi = __shfl_xor(s_hi, o, kWave);
      const ull nlo = s_lo + olo;
      s_hi += ohi + (nlo < s_lo);
      s_lo = nlo;
      sd += __shfl_xor(sd, o, kWave);
      st += __shfl_xor(st, o, kWave);
    }
    if (lane != 0) continue;
    const u128 S = ((u128)s_hi << 64) | s_lo;
    const u128 K = (u128)(ull)(du - 1) * sd - (u128)((ull)du * (ull)(du - 1)) - st;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      if (j < njobs && jobs[j].p[11]) {
        uint32_t fb = 0;  // the wedges of this centre that pass the filters set the binomials' flags
        const int f = jobs[j].p[12];
        const u128 b0 = binom(du - 2, f, fb), b2 = binom(du - 4, f, fb);
        pos[j].add(mulx(b0 - b2, S, pos[j].fl));
        ng[j].add(mulx(b0, K, ng[j].fl));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    if (j < njobs) {
      flush(pos[j], jobs[j].out);
      flush(ng[j], neg + j);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bytes += __shfl_xor(bytes, o, kWave);
  if (lane == 0 && bytes) atomicAdd((ull*)&jobs[0].out->bytes, bytes);
}

// TL: element type of the common-neighbour lists the kernel reads (positions in N(u):
// uint16_t when every degree is at most 65536, which halves the list traffic of the c3
// scatter, else int)
template <int NJ, bool kFull, int kPivWaves, int kECap, int kHub, typename TL>
__global__ __launch_bounds__(kPivWaves * kWave) void k_wedge_pivot(const int* __restrict__ idx, const int* __restrict__ adj,
                                                           const int* __restrict__ tsup,
                                                           const int* __restrict__ adjdeg,
                                                           const int* __restrict__ adjhub,
                                                           const int* __restrict__ rev,
                                                           const long long* __restrict__ tl_off,
                                                           const TL* __restrict__ tl, const JobDesc* __restrict__ jobs,
                                                           int njobs, int need_c3, const PivotItem* __restrict__ items,
                                                           int nitems, int* __restrict__ head,
                                                           unsigned* __restrict__ dbuf, long long dstride,
                                                           int ecap, int achunk, int wchunk)
{
  // ecap <= kECap: positions per c3 window; achunk, wchunk: positions per shared phase-A
  // chunk and per walk chunk without c3 (the kAChunk / kWalkChunk defaults, or smaller
  // when the caller caps the LDS staging)
  __shared__ unsigned s_e[kPivWaves][kECap];
  __shared__ unsigned s_hub[kHub > 0 ? kHub : 1];
  __shared__ int s_item, s_next, s_nheavy, s_task, s_anext, s_anheavy, s_atask;
  // per wave (launches with c3): the window-cut list of each lane's x in the flattened
  // c3 scatter, its start and its offset in the sequence
  __shared__ long long s_flb[kECap > 1 ? kPivWaves : 1][kWave];
  __shared__ int s_foff[kECap > 1 ? kPivWaves : 1][kWave + 1];
  __shared__ int s_heavy[kHeavyCap], s_hwin[kHeavyCap];
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  unsigned* __restrict__ D = dbuf + (long long)blockIdx.x * dstride;
  const G g{idx, adj};
  LaneAcc acc[NJ];
  // algorithmic global-memory bytes of this schedule: every logical load, store or
  // atomic on a global array counted at its size (atomics twice), no cache credit; a
  // binary search over a list of length L counts bitlen(L + 1) probes. The clearing of
  // D (every 255 items, and at the end) is scratch upkeep that depends on how the
  // dynamic queue spreads items, and is not counted. Per unit:
  //   2-hop path 20 (adj, two atomics counted twice) or 4 (LDS hub counter);
  //   adjacency mark 16, 20 in hub launches, 4 for an LDS hub; neighbour x of the
  //   pivot 12 + 4 bitlen(d_x + 1); centre 16 + 4 bitlen(d_u + 1); walk position 16,
  //   12 for an LDS hub; c3 entry x of T(u,w) per window 16 + s + (8 + s) bitlen(t_ux + 1)
  //   (twice the search in a window that is not the centre's last); list element s
  //   (s = 2 for 16-bit lists, else 4)
  ull bytes = 0;
  // D entries carry the item's epoch in their top 8 bits: an entry written for an
  // earlier item reads as zero, so no pass clears D after an item. Every 255 items,
  // and when the workgroup finishes, D is cleared whole (sequential stores), which
  // also leaves it zero for the next launch.
  unsigned epoch = 0;
  const long long dn = dstride;
  // the hub counters are cleared before every item (the barrier after the item's
  // fetch orders the clearing before its scatter)
  if (kHub > 0)
    for (int i = threadIdx.x; i < kHub; i += kPivWaves * kWave) s_hub[i] = 0u;

  for (;;) {
    if (threadIdx.x == 0) {
      s_item = atomicAdd(head, 1);
      s_next = s_nheavy = s_task = 0;
      s_anext = s_anheavy = s_atask = 0;
    }
    __syncthreads();
    const int it = s_item;
    if (it >= nitems) break;
    const PivotItem I = items[it];
    const int w = I.w, wb = g.beg(w), we = g.end(w), dw = we - wb;
    if (jobs[0].smod > 1) {
      // a root slice counts only centres u with u % smod == srem: a pivot none of whose
      // centres is in the slice contributes nothing, so its 2-hop scatter is skipped
      // (every thread has read s_item above, so the next fetch may overwrite it)
      const int smod = jobs[0].smod, srem = jobs[0].srem;
      int hit = 0;
      for (int p = I.p0 + (int)threadIdx.x; p < I.p1 && !hit; p += kPivWaves * kWave) hit = (adj[wb + p] % smod) == srem;
      if (!__syncthreads_or(hit)) continue;
    }
    if (++epoch == 256) {
      for (long long i = threadIdx.x; i < dn; i += kPivWaves * kWave) D[i] = 0u;
      __syncthreads();
      epoch = 1;
    }
    const unsigned e0 = epoch << 24;

    // Deferred heavy work of phases A and B: entries s_heavy[h] with their task counts,
    // turned by wave 0 into an inclusive prefix in s_hwin (nh <= kHeavyCap) ...
    auto heavy_prefix = [&](int nh) {
      if (wave == 0) {
        int carry = 0;
        for (int b0 = 0; b0 < nh; b0 += kWave) {
          const int h = b0 + lane;
          int v = h < nh ? s_hwin[h] : 0;
#pragma unroll
          for (int o = 1; o < kWave; o <<= 1) {
            const int t = __shfl_up(v, o, kWave);
            if (lane >= o) v += t;
          }
          if (h < nh) s_hwin[h] = v + carry;
          carry += __shfl(v, kWave - 1, kWave);
        }
      }
    };
    // ... and task t mapped to (entry h, task k of that entry)
    auto heavy_task = [&](int t, int nh, int& h, int& k) {
      int lo = 0, hi = nh - 1;  // first h with s_hwin[h] > t
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_hwin[mid] > t) hi = mid; else lo = mid + 1;
      }
      h = lo;
      k = t - (lo > 0 ? s_hwin[lo - 1] : 0);
    };

    // Phase A: D[y] += 2 for every path w - x - y with y > w; D[x] |= 1 for x in N(w).
    // With hub slots, adjhub[e] names the LDS counter of a hub (>= 0) or ~y otherwise.
    // Waves take the x one at a time; an x with more than kAChunk positions past w is
    // deferred (phase A2) and its chunks shared by all waves.
    // Eight list loads per lane are issued before their (fire-and-forget) atomics.
    auto scatter2 = [&](int k0, int k1) {
      for (int b = k0; b < k1; b += 8 * kWave) {
        int hv[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
          const int k = b + r * kWave + lane;
          hv[r] = k < k1 ? (kHub > 0 ? adjhub[k] : ~adj[k]) : 0;
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
          if (b + r * kWave + lane >= k1) continue;
          const int h = hv[r];
          if (kHub > 0 && h >= 0) {
            atomicAdd(&s_hub[h], 2u);
            bytes += 4;
          } else {
            atomicMax(&D[~h], e0);
            atomicAdd(&D[~h], 2u);
            bytes += 20;
          }
        }
      }
    };
    for (;;) {
      int c = 0;
      if (lane == 0) c = atomicAdd(&s_anext, 1);
      c = bcast(c, 0);
      const int i = wb + c;
      if (i >= we) break;
      const int x = adj[i];
      const int xe = g.end(x);
      if (lane == 0 && x > w) {
        const int hx = kHub > 0 ? adjhub[i] : ~x;
        if (kHub > 0 && hx >= 0) {
          atomicOr(&s_hub[hx], 1u);
          bytes += 4;
        } else {
          atomicMax(&D[~hx], e0);
          atomicOr(&D[~hx], 1u);
          bytes += kHub > 0 ? 20 : 16;
        }
      }
      const int s = rev[i] + 1;  // past w in N(x)
      if (lane == 0) bytes += 12;
      const int nch = (xe - s + achunk - 1) / achunk;
      if (nch > 1) {
        int slot = 0;
        if (lane == 0) slot = atomicAdd(&s_anheavy, 1);
        slot = bcast(slot, 0);
        if (slot < kHeavyCap) {
          if (lane == 0) { s_heavy[slot] = i; s_hwin[slot] = nch; }
          continue;
        }
      }
      scatter2(s, xe);
    }
    __syncthreads();
    {
      const int nh = min(s_anheavy, kHeavyCap);
      if (nh > 0) {
        heavy_prefix(nh);
        __syncthreads();
        const int ntask = s_hwin[nh - 1];
        for (;;) {
          int t = 0;
          if (lane == 0) t = atomicAdd(&s_atask, 1);
          t = bcast(t, 0);
          if (t >= ntask) break;
          int h, k;
          heavy_task(t, nh, h, k);
          const int x = adj[s_heavy[h]];
          const int xe = g.end(x);
          const int s = rev[s_heavy[h]] + 1 + k * achunk;
          scatter2(s, min(xe, s + achunk));
        }
        __syncthreads();
      }
    }

    // one window [q_lo, q_hi) of centre u's walk past w: the c3 scatter into this
    // wave's LDS counters (when c3on), then the walk
    // one window [q_lo, q_hi) of centre u's walk past w: the c3 scatter into this
    // wave's LDS counters (when c3on), then the walk
    auto window = [&](int u, int ub, int du, int tuw, bool c3on, long long tb, long long te, int q_lo, int q_hi) {
      unsigned* __restrict__ E = s_e[wave];
      const bool last = q_hi == du;
      // the centre's binomials of the linear jobs (uniform over the wave)
      bool cb_ok[NJ];
      ull cb0[NJ], cb1[NJ], cb2[NJ];
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        cb_ok[j] = false;
        cb0[j] = cb1[j] = cb2[j] = 0;
        if (!(kPivWaves == 16 && kHub == 0) && j < njobs && jobs[j].p[11]) {
          const int f = jobs[j].p[12], nx = du - 2;
          cb_ok[j] = fc3_small_inline_ok(f, nx);
          if (cb_ok[j]) {
            cb0[j] = ubin_small(nx, f);
            cb1[j] = ubin_small(nx - 1, f);
            cb2[j] = ubin_small(nx - 2, f);
          }
        }
      }
      if (c3on) {
        for (int q = lane; q < q_hi - q_lo; q += kWave) E[q] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // scatter: for x in T(u, w), for y in T(u, x) with q_lo <= y < q_hi: E[y - q_lo]++.
        // Each lane cuts the list of its own x to the window; the cut lists of the 64
        // lanes are then streamed as one sequence, eight elements per lane in flight.
        long long* __restrict__ FLB = s_flb[kECap > 1 ? wave : 0];
        int* __restrict__ FOFF = s_foff[kECap > 1 ? wave : 0];
        for (long long t0 = tb; t0 < te; t0 += kWave) {
          const long long t = t0 + lane;
          long long lb = 0, le = 0;
          if (t < te) {
            const int xq = (int)tl[t];
            const long long b0 = tl_off[ub + xq], e0 = tl_off[ub + xq + 1];
            long long lo = b0, hi = e0;  // first entry >= q_lo
            while (lo < hi) { const long long mid = (lo + hi) >> 1; if ((int)tl[mid] < q_lo) lo = mid + 1; else hi = mid; }
            bytes += 16 + sizeof(TL) + (8 + sizeof(TL)) * (64 - __clzll(e0 - b0 + 1));
            lb = lo;
            le = e0;
            if (!last) {  // first entry >= q_hi
              hi = e0;
              while (lo < hi) { const long long mid = (lo + hi) >> 1; if ((int)tl[mid] < q_hi) lo = mid + 1; else hi = mid; }
              le = lo;
              bytes += (8 + sizeof(TL)) * (64 - __clzll(e0 - b0 + 1));
            }
            bytes += (le - lb) * sizeof(TL);
          }
          const int len = (int)(le - lb);
          int incl = len;
#pragma unroll
          for (int o = 1; o < kWave; o <<= 1) {
            const int v = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += v;
          }
          const int tot = __shfl(incl, kWave - 1, kWave);
          __builtin_amdgcn_wave_barrier();  // the previous batch's readers of FLB / FOFF are done
          FLB[lane] = lb;
          FOFF[lane] = incl - len;
          if (lane == kWave - 1) FOFF[kWave] = tot;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          // element g of the sequence is element g - FOFF[k] of lane k's cut list, k the
          // last lane with FOFF[k] <= g (a lane with an empty list shares its offset with
          // the next one, which wins); k only grows with g
          int k = 0;
          {
            int lo = 0, hi = kWave;
            const int g0 = min(lane, max(tot - 1, 0));
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (FOFF[mid] <= g0) lo = mid; else hi = mid;
            }
            k = lo;
          }
          for (int c = 0; c < tot; c += 8 * kWave) {
            int a[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
              const int g = c + r * kWave + lane;
              a[r] = -1;
              if (g < tot) {
                while (FOFF[k + 1] <= g) k++;
                a[r] = (int)tl[FLB[k] + (g - FOFF[k])];
              }
            }
#pragma unroll
            for (int r = 0; r < 8; r++)
              if (a[r] >= 0) atomicAdd(&E[a[r] - q_lo], 1u);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      // walk the window, software-pipelined in steps of P * 64 positions: the list loads
      // of a step run two steps ahead of its evaluation and its 2-hop counter loads one
      // step ahead, so each wait for memory is covered by the evaluation of a step
      auto pos_load = [&](int q, int& h, int& t, int& d) __attribute__((always_inline)) {
        if (q < q_hi) {
          h = kHub > 0 ? adjhub[ub + q] : ~adj[ub + q];
          t = tsup[ub + q];
          d = adjdeg[ub + q];
        }
      };
      auto cnt_load = [&](int q, int h, unsigned& r, bool& gd) __attribute__((always_inline)) {
        if (q < q_hi) {
          gd = !(kHub > 0 && h >= 0);  // its counter is in D, not an LDS hub slot
          r = gd ? D[~h] : s_hub[kHub > 0 ? h : 0];
        }
      };
      auto eval_pos = [&](int q, int tuv, int dv, unsigned r, bool gd) __attribute__((always_inline)) {
        if (q < q_hi) {
          unsigned dv_ = r;
          if (gd) {
            dv_ = (r >> 24) == epoch ? (r & 0xffffffu) : 0u;
            bytes += 4;
          }
          const bool tri = dv_ & 1u;
          const int cvw = (int)(dv_ >> 1);
          const int c3 = c3on ? (int)E[q - q_lo] : 0;
          bytes += 12;
#pragma unroll
          for (int j = 0; j < NJ; j++) {
            if (j < njobs) {
              u128 val;
              bool match = true;
              bool take;
              // the 16-wave c3 shape never runs a linear job (a launch with c3 evaluates
              // them with their c3); the others may run launches without c3
              if (!(kPivWaves == 16 && kHub == 0) && jobs[j].p[11]) {
                take = eval_wedge_lin(jobs + j, u, w, du, dv, dw, tuv, tuw, cvw, tri, val, match, acc[j].fl, cb_ok[j],
                                      cb0[j], cb1[j], cb2[j]);
              } else if constexpr (kHub > 0) {
                // the hub shape runs launches without c3: besides the linear jobs only
                // specialised ones (any other general job makes the launch take c3)
                take = eval_wedge_special(jobs + j, du, dv, dw, cvw, tri, val, acc[j].fl);
              } else {
                take = eval_wedge<kFull>(jobs + j, u, w, du, dv, dw, tuv, tuw, cvw, c3, tri, val, acc[j].fl);
              }
              if (take) {
                acc[j].add(val);
                acc[j].matches += match;
              }
            }
          }
        }
      };
      // P positions per lane per step: two in the 12- and 8-wave hub shapes, one where
      // the registers of a second would spill (16 waves, and the c3 shapes, which
      // measured 3% slower with two)
      constexpr int P = (kHub > 0 && kPivWaves < 16) ? 2 : 1;
      constexpr int kStep = P * kWave;
      const int qa = q_lo + lane;
      int t0[P], d0[P];  // this step: loaded, counters loaded
      unsigned r0[P];
      bool g0[P];
      int h1[P], t1[P], d1[P];  // the next step: loaded
#pragma unroll
      for (int z = 0; z < P; z++) {
        int h0 = 0;
        t0[z] = d0[z] = 0;
        r0[z] = 0u;
        g0[z] = false;
        h1[z] = t1[z] = d1[z] = 0;
        pos_load(qa + z * kWave, h0, t0[z], d0[z]);
        cnt_load(qa + z * kWave, h0, r0[z], g0[z]);
        pos_load(qa + kStep + z * kWave, h1[z], t1[z], d1[z]);
      }
      for (int q0 = q_lo; q0 < q_hi; q0 += kStep) {
        const int q = q0 + lane;
        unsigned r1[P];
        bool g1[P];
        int h2[P], t2[P], d2[P];
#pragma unroll
        for (int z = 0; z < P; z++) {
          r1[z] = 0u;
          g1[z] = false;
          cnt_load(q + kStep + z * kWave, h1[z], r1[z], g1[z]);
        }
#pragma unroll
        for (int z = 0; z < P; z++) {
          h2[z] = t2[z] = d2[z] = 0;
          pos_load(q + 2 * kStep + z * kWave, h2[z], t2[z], d2[z]);
        }
#pragma unroll
        for (int z = 0; z < P; z++) eval_pos(q + z * kWave, t0[z], d0[z], r0[z], g0[z]);
#pragma unroll
        for (int z = 0; z < P; z++) {
          t0[z] = t1[z]; d0[z] = d1[z]; r0[z] = r1[z]; g0[z] = g1[z];
          h1[z] = h2[z]; t1[z] = t2[z]; d1[z] = d2[z];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    };

    // Phase B: centres, one per wave at a time. A centre whose walk past w spans more
    // than one window is deferred to phase B2 (up to kHeavyCap of them per item), where
    // its windows are shared by all waves: a hub centre would otherwise keep one wave
    // busy while the others wait at the item's closing barrier.
    for (;;) {
      int c = 0;
      if (lane == 0) c = atomicAdd(&s_next, 1);
      c = bcast(c, 0);
      const int p = I.p0 + c;
      if (p >= I.p1) break;
      const int u = adj[wb + p];
      const int ub = g.beg(u), ue = g.end(u), du = ue - ub;
      const int tuw = tsup[wb + p];
      const int rq = rev[wb + p] - ub;  // position of w in N(u)
      if (rq + 1 >= du) continue;
      bool in_slice = true;
      if (jobs[0].smod > 1) in_slice = (u % jobs[0].smod) == jobs[0].srem;
      if (!in_slice) continue;
      if (lane == 0) bytes += 20;
      const bool c3on = need_c3 && tuw > 0;
      const int cstep = c3on ? ecap : wchunk;
      const int nwin = (du - (rq + 1) + cstep - 1) / cstep;
      if (nwin > 1) {
        int slot = 0;
        if (lane == 0) slot = atomicAdd(&s_nheavy, 1);
        slot = bcast(slot, 0);
        if (slot < kHeavyCap) {
          if (lane == 0) { s_heavy[slot] = p; s_hwin[slot] = nwin; }
          continue;
        }
      }
      const long long tb = c3on ? tl_off[ub + rq] : 0, te = c3on ? tl_off[ub + rq + 1] : 0;
      // the walk past w in windows of ecap positions, each with its c3 counters in LDS
      // (without c3 the whole walk is one window)
      const int wstep = c3on ? ecap : du;
      for (int q_lo = rq + 1; q_lo < du; q_lo += wstep) window(u, ub, du, tuw, c3on, tb, te, q_lo, min(du, q_lo + wstep));
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Phase B2: the deferred centres' windows as one task list, taken one window per
    // wave at a time
    const int nh = min(s_nheavy, kHeavyCap);
    if (nh > 0) {
      heavy_prefix(nh);
      __syncthreads();
      const int ntask = s_hwin[nh - 1];
      for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(&s_task, 1);
        t = bcast(t, 0);
        if (t >= ntask) break;
        int hh, k;
        heavy_task(t, nh, hh, k);
        const int p = s_heavy[hh];
        const int u = adj[wb + p];
        const int ub = g.beg(u), ue = g.end(u), du = ue - ub;
        const int tuw = tsup[wb + p];
        const int rq = rev[wb + p] - ub;
        const bool c3on = need_c3 && tuw > 0;
        const int cstep = c3on ? ecap : wchunk;
        const long long tb = c3on ? tl_off[ub + rq] : 0, te = c3on ? tl_off[ub + rq + 1] : 0;
        const int q_lo = rq + 1 + k * cstep;
        window(u, ub, du, tuw, c3on, tb, te, q_lo, min(du, q_lo + cstep));
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    if (kHub > 0)
      for (int i = threadIdx.x; i < kHub; i += kPivWaves * kWave) s_hub[i] = 0u;
  }
  // leave D zero for the next launch
  if (epoch > 0) {
    for (long long i = threadIdx.x; i < dn; i += kPivWaves * kWave) D[i] = 0u;
  }
#pragma unroll
  for (int j = 0; j < NJ; j++)
    if (j < njobs) flush(acc[j], jobs[j].out);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bytes += __shfl_xor(bytes, o, kWave);
  if (lane == 0) atomicAdd((ull*)&jobs[0].out->bytes, bytes);
}

// 16-bit copy of the common-neighbour lists (every entry below 65536)
__global__ void k_narrow16(const int* __restrict__ tl, long long total, uint16_t* __restrict__ out)
{
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x)
    out[e] = (uint16_t)tl[e];
}

// degree of the neighbour at every CSR position
__global__ void k_adjdeg(const int* __restrict__ idx, const int* __restrict__ adj, long long m, int* __restrict__ out)
{
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (long long)gridDim.x * blockDim.x) {
    const int x = adj[e];
    out[e] = idx[x + 1] - idx[x];
  }
}

// hub slot (or ~vertex) of the neighbour at every CSR position
__global__ void k_adjhub(const int* __restrict__ adj, long long m, const int* __restrict__ slot, int* __restrict__ out)
{
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (long long)gridDim.x * blockDim.x)
    out[e] = slot[adj[e]];
}

// cost of pivot w: sum of the degrees of its neighbours
__global__ void k_pivot_cost(const int* __restrict__ idx, const int* __restrict__ adj, int n, long long* __restrict__ cost)
{
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  long long s = 0;
  for (int i = idx[w]; i < idx[w + 1]; i++) { const int x = adj[i]; s += idx[x + 1] - idx[x]; }
  cost[w] = s;
}

}  // namespace dev

#define HIP_TRY(expr)                                                    \
  do {                                                                   \
    hipError_t e_ = (expr);                                              \
    if (e_ != hipSuccess) {                                              \
      if (err) *err = std::string(#expr) + ": " + hipGetErrorString(e_); \
      return -5;                                                         \
    }                                                                    \
  } while (0)

void device_derived_free(DeviceGraph* g)
{
  if (g->tsup) (void)hipFree(g->tsup);
  if (g->adjdeg) (void)hipFree(g->adjdeg);
  if (g->adjhub) (void)hipFree(g->adjhub);
  if (g->rev) (void)hipFree(g->rev);
  g->rev = nullptr;
  if (g->tl_off) (void)hipFree(g->tl_off);
  if (g->tl) (void)hipFree(g->tl);
  if (g->tl16) (void)hipFree(g->tl16);
  if (g->piv_cost) (void)hipFree(g->piv_cost);
  if (g->c3tl) (void)hipFree(g->c3tl);
  g->c3tl = nullptr;
  for (auto& kv : g->piv_items) (void)hipFree(kv.second.first);
  g->piv_items.clear();
  for (auto& kv : g->tri_items) (void)hipFree(kv.second.first);
  g->tri_items.clear();
  g->tsup = nullptr;
  g->adjdeg = nullptr;
  g->adjhub = nullptr;
  g->tl_off = nullptr;
  g->tl = nullptr;
  g->tl16 = nullptr;
  g->piv_cost = nullptr;
  g->derived_ready = false;
  g->support_ready = false;
}

// Builds the per-edge supports and pivot costs, and (lists = true) the per-edge
// common-neighbour lists, each once per graph. The lists hold 6 entries per triangle;
// schedules that need only the supports (the edge pass) never allocate them.
int device_derived_build(DeviceGraph* g, void* stream_v, double* ms, std::string* err, bool lists)
{
  // one build per graph, also when several host threads count on it at once
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (ms) *ms = 0;
  if (g->derived_ready || (!lists && g->support_ready)) return 0;
  hipStream_t stream = (hipStream_t)stream_v;
  const int n = g->n;
  const long long m = g->m;
  EventPair ev;
  HIP_TRY(ev.create());
  HIP_TRY(hipEventRecord(ev.a, stream));
  DevScratch sc(stream);
  int* d_head = nullptr;
  HIP_TRY(sc.get(&d_head, sizeof(int) * 2));
  HIP_TRY(hipMemsetAsync(d_head, 0, sizeof(int) * 2, stream));
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // work items of the intersection passes: each vertex u with its positions cut into
  // ranges (256 positions above 1024), highest degree first so the longest items start
  // early; a counting sort over the degrees
  dev::ComItem* d_citems = nullptr;
  int ncitems = 0;
  auto common_items = [&]() -> int {
    if (d_citems) return 0;
    const int md = g->maxdeg;
    std::vector<long long> start((size_t)md + 2, 0);
    for (int v = 0; v < n; v++) {
      const int d = g->h_idx[v + 1] - g->h_idx[v];
      if (d > 0) start[(size_t)(md - d) + 1] += d <= 1024 ? 1 : (d + 255) / 256;
    }
    for (size_t i = 1; i < start.size(); i++) start[i] += start[i - 1];
    std::vector<dev::ComItem> items((size_t)start[(size_t)md + 1]);
    for (int v = 0; v < n; v++) {
      const int d = g->h_idx[v + 1] - g->h_idx[v];
      if (d == 0) continue;
      long long& at = start[(size_t)(md - d)];
      const int step = d <= 1024 ? d : 256;
      for (int p = 0; p < d; p += step) items[(size_t)at++] = dev::ComItem{v, p, std::min(d, p + step), 0};
    }
    ncitems = (int)items.size();
    if (sc.get(&d_citems, sizeof(dev::ComItem) * std::max<size_t>(1, items.size())) != hipSuccess) return -5;
    if (!items.empty() && hipMemcpyAsync(d_citems, items.data(), sizeof(dev::ComItem) * items.size(),
                                         hipMemcpyHostToDevice, stream) != hipSuccess)
      return -5;
    return hipStreamSynchronize(stream) == hipSuccess ? 0 : -5;  // the host items go out of scope
  };
  const int cgrid = cus * 8;
  if (!g->support_ready) {
    HIP_TRY(hipMalloc((void**)&g->tsup, sizeof(int) * std::max<long long>(1, m)));
    HIP_TRY(hipMalloc((void**)&g->piv_cost, sizeof(long long) * std::max(1, n)));
    HIP_TRY(hipMalloc((void**)&g->adjdeg, sizeof(int) * std::max<long long>(1, m)));
    hipLaunchKernelGGL(dev::k_adjdeg, dim3(cus * 8), dim3(256), 0, stream, g->idx, g->adj, m, g->adjdeg);
    HIP_TRY(hipGetLastError());
    // hub slots: the kHubSlots vertices of highest degree (ties by id) get LDS slots
    // 0, 1, ... in that order; every other vertex v is stored as ~v
    {
      std::vector<int> start((size_t)g->maxdeg + 2, 0), slot((size_t)std::max(1, n));
      for (int v = 0; v < n; v++) start[(size_t)(g->maxdeg - (g->h_idx[v + 1] - g->h_idx[v])) + 1]++;
      for (size_t i = 1; i < start.size(); i++) start[i] += start[i - 1];
      for (int v = 0; v < n; v++) {
        const int r = start[(size_t)(g->maxdeg - (g->h_idx[v + 1] - g->h_idx[v]))]++;
        slot[v] = r < dev::kHubSlots ? r : ~v;
      }
      int* d_slot = nullptr;
      HIP_TRY(sc.get(&d_slot, sizeof(int) * (size_t)std::max(1, n)));
      HIP_TRY(hipMalloc((void**)&g->adjhub, sizeof(int) * std::max<long long>(1, m)));
      HIP_TRY(hipMemcpyAsync(d_slot, slot.data(), sizeof(int) * (size_t)std::max(1, n), hipMemcpyHostToDevice, stream));
      hipLaunchKernelGGL(dev::k_adjhub, dim3(cus * 8), dim3(256), 0, stream, g->adj, m, d_slot, g->adjhub);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(stream));  // the host slot array goes out of scope
    }
    if (common_items()) { if (err) *err = "upload of the intersection work items failed"; return -5; }
    hipLaunchKernelGGL(dev::k_common<false>, dim3(cgrid), dim3(dev::kComBlock), 0, stream, g->idx, g->adj,
                       g->adjdeg, d_citems, ncitems, d_head, g->tsup, (const long long*)nullptr, (int*)nullptr,
                       (int*)nullptr);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(dev::k_pivot_cost, dim3((n + 255) / 256), dim3(256), 0, stream, g->idx, g->adj, n, g->piv_cost);
    HIP_TRY(hipGetLastError());
    g->h_piv_cost.resize(n);
    HIP_TRY(hipMemcpyAsync(g->h_piv_cost.data(), g->piv_cost, sizeof(long long) * n, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    g->derived_bytes = 3 * sizeof(int) * (size_t)m + sizeof(long long) * (size_t)n;
    g->support_ready = true;
  }
  if (lists) {
    // tl_off = exclusive prefix sum of tsup (64-bit)
    HIP_TRY(hipMalloc((void**)&g->tl_off, sizeof(long long) * (m + 1)));
    HIP_TRY(hipMemsetAsync(g->tl_off, 0, sizeof(long long), stream));
    size_t tmp_bytes = 0;
    hipcub::TransformInputIterator<long long, dev::Widen, const int*> in(g->tsup, dev::Widen());
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, in, g->tl_off + 1, (int)m, stream));
    void* tmp = nullptr;
    HIP_TRY(sc.get(&tmp, tmp_bytes + 16));
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, in, g->tl_off + 1, (int)m, stream));
    long long total = 0;
    HIP_TRY(hipMemcpyAsync(&total, g->tl_off + m, sizeof(long long), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    g->tl_total = total;
    HIP_TRY(hipMalloc((void**)&g->tl, sizeof(int) * std::max<long long>(1, total)));
    HIP_TRY(hipMalloc((void**)&g->rev, sizeof(int) * std::max<long long>(1, m)));
    HIP_TRY(hipMemsetAsync(d_head + 1, 0, sizeof(int), stream));
    if (common_items()) { if (err) *err = "upload of the intersection work items failed"; return -5; }
    hipLaunchKernelGGL(dev::k_common<true>, dim3(cgrid), dim3(dev::kComBlock), 0, stream, g->idx, g->adj,
                       g->adjdeg, d_citems, ncitems, d_head + 1, g->tsup, (const long long*)g->tl_off, g->tl,
                       g->rev);
    HIP_TRY(hipGetLastError());
    g->derived_bytes += sizeof(long long) * ((size_t)m + 1) + sizeof(int) * (size_t)total + sizeof(int) * (size_t)m;
    // 16-bit copy of the lists for the pivot kernel: a position in N(u) is below
    // d_u <= maxdeg, so it fits when maxdeg <= 65536
    if (g->maxdeg <= 65536 && total > 0) {
      HIP_TRY(hipMalloc((void**)&g->tl16, sizeof(uint16_t) * (size_t)total));
      hipLaunchKernelGGL(dev::k_narrow16, dim3(cus * 8), dim3(256), 0, stream, (const int*)g->tl, total, g->tl16);
      HIP_TRY(hipGetLastError());
      g->derived_bytes += sizeof(uint16_t) * (size_t)total;
    }
    g->derived_ready = true;
  }
  HIP_TRY(hipEventRecord(ev.b, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, ev.a, ev.b));
  if (ms) *ms = t;
  return 0;
}

// waves per workgroup of the pivot launch: Tune::block of 256, 512, 768 or 1024 threads
// selects the 4-, 8-, 12- and 16-wave shapes. Launches with c3 default to 16 waves (four
// per SIMD against two for the 64 KB c3 shapes: 6.4 s against 8.0 s on R-MAT s20), and
// 12 waves take that shape too. Launches without c3 come in 16-, 12- and 8-wave hub
// shapes (at most 128, 168 and 256 VGPRs per lane) and default to 12 (C2: 690 ms
// against 751 ms for 16 waves, whose walk spills, and 1066 ms for 8 before the walk
// was pipelined).
static int pivot_waves(const Tune& t, bool c3)
{
  if (t.block <= 0) return c3 ? 16 : 12;
  if (t.block >= 1024) return 16;
  if (t.block >= 768) return 12;
  return t.block >= 512 ? 8 : 4;
}

// Grid of the pivot launch (before clamping to the number of items); its device
// scratch is one dense D array of n counters per workgroup.
static int pivot_grid(const Tune& t)
{
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // as many workgroups as fit at once: one per CU for the 128 KB shapes, two with 64 KB
  // (the queue is dynamic; 256, 512 and 768 workgroups of 16 waves time the same on
  // R-MAT s20, and each holds a D array of n counters)
  return t.grid > 0 ? t.grid : (pivot_waves(t, true) > 4 ? cus : cus * 2);
}

size_t pivot_scratch_bytes(const DeviceGraph& g, const Tune& t)
{
  const size_t grid = (size_t)pivot_grid(t);
  const size_t dstride = ((size_t)g.n + 63) / 64 * 64;
  return sizeof(unsigned) * grid * dstride;  // the work items stay with the graph
}

// The cost above which a pivot's centres are split over several work items: 1/4096 of
// the launch's total (16 items per CU of an MI355X), at least 2^16 and chunk * 4096.
// It does not depend on the device, so the items, and the algorithmic bytes, are a
// function of the graph alone.
long long pivot_split_budget(long long total_cost, int chunk)
{
  return std::max<long long>(std::max<long long>(1 << 16, (long long)chunk << 12), total_cost / 4096);
}

// Prepares one pivot-schedule launch of up to kMaxPivotJobs wedge-class jobs.
int plan_wedge_pivot(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream_v, LaunchPlan* lp,
                     std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int nj = (int)jobs.size();
  if (nj > dev::kMaxPivotJobs) { if (err) *err = "too many jobs for one pivot launch"; return -7; }
  // 2 * c(w, y) + 1 must fit the 24 count bits of a D entry (c(w, y) <= maxdeg); the
  // class dispatch sends such graphs to the per-root kernel, this is the last guard
  if (g->maxdeg >= (1 << 23)) {
    if (err) *err = "maximum degree 2^23 or more: the pivot kernel's 24-bit 2-hop counters could overflow";
    return -75;
  }
  int rc = device_derived_build(g, stream_v, &lp->index_ms, err, true);
  if (rc) return rc;
  // FSGC_PIVOT_C3_PER_WEDGE=1 in the environment (read per launch plan) evaluates the
  // jobs that are linear in the triple region with their per-wedge c3 like the others
  // (timing comparisons and tests; the counts are the same)
  // A root slice evaluates them per wedge too: the reference gives all three rotations
  // of a triangle to its smallest vertex, while the per-centre terms give each rotation
  // to its own centre, so only whole counts agree.
  const char* pw = getenv("FSGC_PIVOT_C3_PER_WEDGE");
  const bool per_wedge_c3 = (pw && *pw && *pw != '0') || (nj > 0 && jobs[0].smod > 1);
  bool need_c3 = false, lin = false;
  for (auto& j : jobs) {
    j.idx = g->idx; j.adj = g->adj; j.n = g->n; j.m = (int32_t)g->m;
    if (j.kind != kWedgeGeneral || per_wedge_c3) j.p[11] = 0;
    need_c3 = need_c3 || (j.kind == kWedgeGeneral && !j.p[11]);
  }
  // a launch that computes c3 anyway evaluates every job with it
  for (auto& j : jobs) {
    if (need_c3) j.p[11] = 0;
    lin = lin || j.p[11];
  }
  // work items: pivots split into centre ranges of bounded walk cost. Every item of a
  // split pivot repeats its 2-hop scatter (phase A), so a pivot is split only when its
  // cost exceeds an eighth of one workgroup's share of the whole launch; heaviest-first
  // order then balances the rest.
  // The items depend on the graph, the chunk and the sample only: they are built once per
  // graph and key and kept on the device until the graph is freed (a launch still running
  // on another stream may read them).
  const dev::PivotItem* d_items = nullptr;
  int nitems = 0;
  {
    static std::mutex items_mu;
    std::lock_guard<std::mutex> lk(items_mu);
    const long long key = ((long long)t.chunk << 32) | (unsigned)std::max(1, t.sample);
    auto it = g->piv_items.find(key);
    if (it == g->piv_items.end()) {
      long long total_cost = 0;
      for (int w = 0; w < g->n; w++) total_cost += g->h_piv_cost[w];
      const long long budget = pivot_split_budget(total_cost, t.chunk);
      std::vector<dev::PivotItem> items;
      std::vector<long long> icost;
      for (int w = 0; w < g->n; w++) {
        const int dw = g->h_idx[w + 1] - g->h_idx[w];
        if (dw == 0) continue;
        const long long c = g->h_piv_cost[w];
        const long long k = std::max<long long>(1, std::min<long long>(dw, c / std::max(budget, c / 8 + 1)));
        for (long long s = 0; s < k; s++) {
          const int p0 = (int)(dw * s / k), p1 = (int)(dw * (s + 1) / k);
          if (p1 > p0) { items.push_back(dev::PivotItem{w, p0, p1, 0}); icost.push_back(c / k + c / 2); }
        }
      }
      // heaviest first
      std::vector<int> order(items.size());
      for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return icost[a] > icost[b]; });
      std::vector<dev::PivotItem> sorted(items.size());
      for (size_t i = 0; i < order.size(); i++) sorted[i] = items[order[i]];
      sample_items(&sorted, t.sample);
      void* d = nullptr;
      HIP_TRY(hipMalloc(&d, sizeof(dev::PivotItem) * std::max<size_t>(1, sorted.size())));
      if (!sorted.empty())
        HIP_TRY(hipMemcpy(d, sorted.data(), sizeof(dev::PivotItem) * sorted.size(), hipMemcpyHostToDevice));
      it = g->piv_items.emplace(key, std::make_pair(d, (int)sorted.size())).first;
    }
    d_items = (const dev::PivotItem*)it->second.first;
    nitems = it->second.second;
  }

  const int grid = std::max(1, std::min<int>(pivot_grid(t), nitems));
  const long long dstride = ((long long)g->n + 63) / 64 * 64;
  // jobs linear in the triple region take a negative part (nj more accumulators)
  rc = plan_accumulators(lp, nj, stream, err, lin);
  if (rc) return rc;
  DevScratch& sc = *lp->scratch;
  unsigned* dbuf = nullptr;
  JobDesc* d_jobs = nullptr;
  HIP_TRY(sc.get(&dbuf, sizeof(unsigned) * dstride * grid));
  const int waves = pivot_waves(t, need_c3);
  for (int k = 0; k < nj; k++) jobs[k].out = lp->d_acc + k;
  HIP_TRY(sc.get(&d_jobs, sizeof(JobDesc) * std::max(1, nj)));
  // D starts at zero; every launch leaves it zero again (phase C of the kernel)
  HIP_TRY(hipMemsetAsync(dbuf, 0, sizeof(unsigned) * dstride * grid, stream));
  if (nj) HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));  // the host job array goes out of scope
  const int c3 = need_c3 ? 1 : 0;
  bool full = false;  // any fringe vector with a multi-vertex anchor set
  for (auto& j : jobs)
    if (j.kind == kWedgeGeneral)
      for (int r = 0; r < j.rot; r++)
        full = full || (j.fr[r][3] | j.fr[r][5] | j.fr[r][6]) != 0;  // pair anchor sets
  if (nitems == 0) return 0;
  const int* idx = g->idx;
  const int* adj = g->adj;
  const int* tsup = g->tsup;
  const int* adjdeg = g->adjdeg;
  const int* adjhub = g->adjhub;
  const int* rev = g->rev;
  const long long* tl_off = g->tl_off;
  const int* tl = g->tl;
  const uint16_t* tl16 = g->tl16;
  // LDS staging: Tune::lds_cap > 0 caps the positions of a c3 window and of the shared
  // phase-A and walk chunks (the launch clamps the window to its shape's array); the
  // results do not depend on it
  const int ecap = t.lds_cap > 0 ? t.lds_cap : 1 << 30;
  const int achunk = t.lds_cap > 0 ? std::min(t.lds_cap, dev::kAChunk) : dev::kAChunk;
  const int wchunk = t.lds_cap > 0 ? std::min(t.lds_cap, dev::kWalkChunk) : dev::kWalkChunk;
  // FSGC_PIVOT_LISTS32=1 in the environment (read per launch plan) makes the kernel
  // read the 32-bit common-neighbour lists even where the 16-bit copy exists: the path
  // graphs of maximum degree above 65536 take, exercised on small graphs by the tests
  const char* l32 = getenv("FSGC_PIVOT_LISTS32");
  const bool lists32 = l32 && *l32 && *l32 != '0';
  lp->nitems = nitems;
  if (lin) {
    const int n = g->n;
    Accum* neg = lp->d_acc + nj;
    int dev = 0, cus = 256;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int cgrid = std::max(1, std::min(cus * 8, (n + 3) / 4));
    lp->post = [=](void* sv) {
      hipStream_t s = (hipStream_t)sv;
      if (nj == 1)
        hipLaunchKernelGGL(dev::k_wedge_lin_centres<1>, dim3(cgrid), dim3(256), 0, s, idx, tsup, adjdeg, n, d_jobs, nj, neg);
      else if (nj == 2)
        hipLaunchKernelGGL(dev::k_wedge_lin_centres<2>, dim3(cgrid), dim3(256), 0, s, idx, tsup, adjdeg, n, d_jobs, nj, neg);
      else
        hipLaunchKernelGGL(dev::k_wedge_lin_centres<dev::kMaxPivotJobs>, dim3(cgrid), dim3(256), 0, s, idx, tsup, adjdeg,
                           n, d_jobs, nj, neg);
    };
  }
  lp->kernel = [=](void* sv, int first, int count, int* head) {
    hipStream_t s = (hipStream_t)sv;
#define FSGC_PIVOT_LAUNCH(NJ, FULL, W, CAP, HUB, TL, TLP)                                                          \
  hipLaunchKernelGGL((dev::k_wedge_pivot<NJ, FULL, W, CAP, HUB, TL>), dim3(grid), dim3(W * dev::kWave), 0, s, idx, \
                     adj, tsup, adjdeg, adjhub, rev, tl_off, TLP, d_jobs, nj, c3, d_items + first, count, head,   \
                     dbuf,                                                                                        \
                     dstride, std::min(CAP, ecap), achunk, wchunk)
#define FSGC_PIVOT_C3(NJ, FULL, W, CAP)                                   \
  do {                                                                    \
    if (tl16 && !lists32) FSGC_PIVOT_LAUNCH(NJ, FULL, W, CAP, 0, uint16_t, tl16); \
    else FSGC_PIVOT_LAUNCH(NJ, FULL, W, CAP, 0, int, tl);                 \
  } while (0)
// hub shapes evaluate only the specialised and linear jobs, whatever kFull says
#define FSGC_PIVOT_SHAPE(NJ, FULL)                                                      \
  do {                                                                                  \
    if (waves >= 12 && c3) FSGC_PIVOT_C3(NJ, FULL, 16, 2048);                            \
    else if (waves == 16) FSGC_PIVOT_LAUNCH(NJ, false, 16, 1, dev::kHubSlots, int, tl); \
    else if (waves == 12) FSGC_PIVOT_LAUNCH(NJ, false, 12, 1, dev::kHubSlots, int, tl); \
    else if (waves == 8 && !c3) FSGC_PIVOT_LAUNCH(NJ, false, 8, 1, dev::kHubSlots, int, tl); \
    else if (waves == 8) FSGC_PIVOT_C3(NJ, FULL, 8, 4096);                               \
    else FSGC_PIVOT_C3(NJ, FULL, 4, 4096);                                               \
  } while (0)
    if (nj == 1) {
      if (full) FSGC_PIVOT_SHAPE(1, true); else FSGC_PIVOT_SHAPE(1, false);
    } else if (nj == 2) {
      if (full) FSGC_PIVOT_SHAPE(2, true); else FSGC_PIVOT_SHAPE(2, false);
    } else {
      if (full) FSGC_PIVOT_SHAPE(dev::kMaxPivotJobs, true); else FSGC_PIVOT_SHAPE(dev::kMaxPivotJobs, false);
    }
#undef FSGC_PIVOT_SHAPE
#undef FSGC_PIVOT_C3
#undef FSGC_PIVOT_LAUNCH
  };
  return 0;
}

// Runs up to kMaxPivotJobs wedge-class jobs on one graph in one pivot-schedule launch.
int device_run_wedge_pivot(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream,
                           std::vector<JobResult>* out, double* kernel_ms, std::string* err)
{
  out->assign(jobs.size(), JobResult());
  if (kernel_ms) *kernel_ms = 0;
  if (jobs.empty()) return 0;
  LaunchPlan lp;
  int rc = plan_wedge_pivot(jobs, g, t, stream, &lp, err);
  return rc ? rc : launch_run(lp, stream, out, kernel_ms, err);
}

}  // namespace fsgc
