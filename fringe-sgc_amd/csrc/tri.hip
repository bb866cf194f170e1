// Triangle-based counting on the rank schedule (gfx950).
//
// Every triangle {x, y, z} is visited once, from its vertex of lowest rank
// (rank = (degree, id)): for x, each higher-ranked neighbour y, and each common
// neighbour z of x and y ranked above y. The per-edge common-neighbour lists T(x, .)
// (positions in N(x), built once per graph, see pivot.hip) give
//   t_xy, t_xz from the support array, t_yz by one search in N(y), and
//   c3 = |N(x) & N(y) & N(z)| = |T(x,y) & T(x,z)|: T(x,y) is marked in LDS (indexed by
//   position in N(x)) and each lane sums the marks over T(x,z) for its own z.
// Since x has the lowest degree of the three, the lists walked are the shortest ones.
// The triangle is then relabelled to the reference's roles (u > v > w by id) and the
// reference's filters and formulas are applied: triangle cores FC:967-1062
// (general), FC:774-931 (4-clique and tailed 4-clique), and the triangle kernels of
// the edge core FC:1864-1966.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "devmath.h"
#include "devscratch.h"
#include "fsgc_plan.h"

namespace fsgc {
namespace dev {

constexpr int kTriBlock = 256;
constexpr int kTriWaves = kTriBlock / kWave;
constexpr int kMarkCap = 8192;  // LDS marks per wave (positions of N(x))
constexpr int kMaxTriJobs = 8;

struct TriItem {
  int32_t x;
  int32_t p0;  // positions [p0, p1) of N(x) taken as y
  int32_t p1;
  int32_t pad;
};

__device__ __forceinline__ bool rank_less(int da, int a, int db, int b) { return da < db || (da == db && a < b); }

__device__ __forceinline__ int lbound(const int* __restrict__ a, int l, int r, int x)
{
  while (l < r) {
    const int mid = (l + r) >> 1;
    if (a[mid] < x) l = mid + 1; else r = mid;
  }
  return l;
}

template <bool kFull>
__device__ __forceinline__ u128 fc3_tri(const int* fr, const int* venn, uint32_t& fl)
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

// One triangle in reference roles (u > v > w by id) for one job. `c3lt` is the number
// of common neighbours q < w, `wq` the sum over them of C(d_q - 3, tails).
template <bool kFull>
__device__ __forceinline__ bool eval_triangle(const JobDesc* __restrict__ J, int du, int dv, int dw, int tuv, int tuw,
                                              int tvw, int c3, int c3lt, u128 wq, u128& val, uint32_t& fl)
{
  const int* p = J->p;
  if (J->cls == kEdge) {  // FC:2054-2103 filters, then FC:1864-1966
    if (!(du > 1 && dv > 1)) return false;
    if (J->kind == kEdgeTriangle) { val = 1; return true; }
    const int t = p[2];
    val = addx(addx(ncr(du - 2, t, fl), ncr(dv - 2, t, fl), fl), ncr(dw - 2, t, fl), fl);
    return true;
  }
  const int sum1 = p[0], sum2 = p[1], sum3 = p[2], tails = p[3];
  if (du < sum1 || dv < max(sum1, sum2 - du)) return false;  // FC:1167-1218
  if (J->kind != kTriGeneral) {
    if (dw < 3) return false;
    if (J->kind == kTriClique) { val = (u128)(unsigned)c3lt; return true; }
    const u128 c123 = addx(addx(ncr(du - 3, tails, fl), ncr(dv - 3, tails, fl), fl), ncr(dw - 3, tails, fl), fl);
    val = addx(mulx(c123, (u128)(unsigned)c3lt, fl), wq, fl);
    return true;
  }
  const int thres2 = max(max(sum1, sum3 - du - dv), max(sum2 - du, sum2 - dv));
  if (dw < thres2) return false;
  int venn[8];  // FC:948-962
  const int cc12 = tuv - (c3 + 1), c13 = tuw - (c3 + 1), c23 = tvw - (c3 + 1);
  venn[0] = 0;
  venn[iuvw] = c3;
  venn[iuv] = cc12;
  venn[iuw] = c13;
  venn[ivw] = c23;
  venn[iu] = du - (cc12 + c13 + (c3 + 2));
  venn[iv] = dv - (cc12 + c23 + (c3 + 2));
  venn[iw] = dw - (c13 + c23 + (c3 + 2));
  u128 s = 0;
  for (int q = 0; q < J->rot; q++) s = addx(s, fc3_tri<kFull>(J->fr[q], venn, fl), fl);
  val = s;
  return true;
}

// T(x,z) lists longer than this are counted by the whole wave (see k_triangle_rank)
constexpr int kCoopLen = 64;

// Four waves per SIMD: the full-fc3 routine (fc3_peeled) would otherwise take 248
// VGPRs and hold the kernel to two; at four it spills (C4: 1.43 s against 1.55 s at
// three waves and 2.06 s at two)
template <int NJ, bool kFull>
__global__ __launch_bounds__(kTriBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_triangle_rank(const int* __restrict__ idx, const int* __restrict__ adj,
                                                             const int* __restrict__ tsup,
                                                             const int* __restrict__ adjdeg,
                                                             const int* __restrict__ rev,
                                                             const long long* __restrict__ tl_off,
                                                             const int* __restrict__ tl,
                                                             const JobDesc* __restrict__ jobs, int njobs, int need_c3,
                                                             int need_lt, int tails, const TriItem* __restrict__ items,
                                                             int nitems,
                                                             int* __restrict__ head, unsigned char* __restrict__ gmark,
                                                             long long mstride, int markcap, int* __restrict__ c3out)
{
  __shared__ unsigned char s_mark[kTriWaves][kMarkCap];
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  unsigned char* __restrict__ Mg = gmark + ((long long)blockIdx.x * kTriWaves + wave) * mstride;
  const G g{idx, adj};
  LaneAcc acc[NJ];
  const int smod = jobs[0].smod, srem = jobs[0].srem;
  // algorithmic global-memory bytes: every logical load or store on a global array at
  // its size, no cache credit (LDS traffic not counted)
  ull bytes = 0;
  // LDS holds whatever the previous kernel left: clear this wave's marks once
  for (int i = lane; i < kMarkCap; i += kWave) s_mark[wave][i] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();

  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(head, 1);
    it = bcast(it, 0);
    if (it >= nitems) break;
    const TriItem I = items[it];
    const int x = I.x, xb = g.beg(x), dx = g.deg(x);
    const bool lds = dx <= markcap;  // markcap <= kMarkCap
    unsigned char* __restrict__ M = lds ? s_mark[wave] : Mg;
    if (lane == 0) bytes += 16 + 8;
    for (int py = I.p0; py < I.p1; py++) {
      const int y = adj[xb + py];
      const int dy = adjdeg[xb + py];
      if (lane == 0) bytes += 8;
      if (!rank_less(dx, x, dy, y)) continue;
      const long long yb = tl_off[xb + py], ye = tl_off[xb + py + 1];
      if (lane == 0) bytes += 16;
      if (ye - yb < 1) continue;
      const int txy = tsup[xb + py];
      // T(y, x) lists the same common neighbours as T(x, y), in the same (id) order,
      // as positions in N(y): entry k of one gives entry k of the other
      const int yb2 = g.beg(y);
      const long long yxb = tl_off[rev[xb + py]];  // the reverse edge (y, x)
      if (lane == 0) bytes += 4 + 4 + 8 + (need_c3 ? (ye - yb) * (lds ? 8 : 10) : 0);
      // mark T(x, y)
      if (need_c3) {
        for (long long k = yb + lane; k < ye; k += kWave) M[tl[k]] = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      // lanes over z in T(x, y) ranked above y. Each lane counts c3 = |T(x,y) & T(x,z)|
      // over its own list T(x,z), except lists longer than kCoopLen in launches that need
      // no c3lt: those are counted by the whole wave, one list at a time, so that one
      // hub list does not keep the other lanes waiting.
      for (long long k0 = yb; k0 < ye; k0 += kWave) {
        const long long k = k0 + lane;
        bool act = false;
        int da = 0, db = 0, dc = 0, tab = 0, tac = 0, tbc = 0, wpos = 0;
        long long zb = 0, ze = 0;
        if (k < ye) {
          const int pz = tl[k];
          const int z = adj[xb + pz];
          const int dz = adjdeg[xb + pz];
          bytes += 12;
          if (rank_less(dy, y, dz, z)) {
            const int txz = tsup[xb + pz];
            const int tyz = tsup[yb2 + tl[yxb + (k - yb)]];
            bytes += 12;
            // reference roles by id: u > v > w
            int a = x, b = y, c = z;
            da = dx; db = dy; dc = dz;
            tab = txy; tac = txz; tbc = tyz;
            // sort (a, b, c) descending by id, carrying degrees and supports
            if (a < b) { int t = a; a = b; b = t; t = da; da = db; db = t; t = tac; tac = tbc; tbc = t; }
            if (b < c) { int t = b; b = c; c = t; t = db; db = dc; dc = t; t = tab; tab = tac; tac = t; }
            if (a < b) { int t = a; a = b; b = t; t = da; da = db; db = t; t = tac; tac = tbc; tbc = t; }
            const int u = a;
            act = !(smod > 1 && (u % smod) != srem);
            if (act && need_c3) {
              zb = tl_off[xb + pz];
              ze = tl_off[xb + pz + 1];
              wpos = need_lt ? lbound(adj, xb, xb + dx, c) - xb : 0;  // positions below w = c
              bytes += 16 + (need_lt ? 4 * (32 - __clz(dx + 1)) : 0) + (ze - zb) * (lds ? 4 : 5);
            }
          }
        }
        int c3 = 0, c3lt = 0;
        u128 wq = 0;
        uint32_t fl0 = 0;
        if (need_c3) {
          const bool big = kFull && act && !need_lt && (ze - zb) > kCoopLen;
          if (act && !big) {
            // four list loads per lane in flight, then their marks
            auto mark_hit = [&](int pq) __attribute__((always_inline)) {
              if (M[pq]) {
                c3++;
                if (need_lt && pq < wpos) {
                  c3lt++;
                  if (tails) wq = addx(wq, ncr(adjdeg[xb + pq] - 3, tails, fl0), fl0);
                }
              }
            };
            long long q = zb;
            for (; q + 4 <= ze; q += 4) {
              int pq[4];
#pragma unroll
              for (int r = 0; r < 4; r++) pq[r] = tl[q + r];
#pragma unroll
              for (int r = 0; r < 4; r++) mark_hit(pq[r]);
            }
            for (; q < ze; q++) mark_hit(tl[q]);
          }
          ull bm = __ballot(big);
          while (bm) {
            const int who = __builtin_ctzll(bm);
            bm &= bm - 1;
            const long long B = bcast(zb, who), Ee = bcast(ze, who);
            int cnt = 0;
            for (long long q = B + lane; q < Ee; q += kWave) cnt += M[tl[q]] ? 1 : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, kWave);
            if (lane == who) c3 = cnt;
          }
        }
        // the triple count of this triangle at z's entry of T(x, y) (x < y < z by rank)
        if (c3out && act) c3out[k] = c3;
        if (act) {
#pragma unroll
          for (int j = 0; j < NJ; j++) {
            if (j < njobs) {
              u128 val;
              if (eval_triangle<kFull>(jobs + j, da, db, dc, tab, tac, tbc, c3, c3lt, wq, val, acc[j].fl)) {
                acc[j].add(val);
                acc[j].matches++;
                acc[j].fl |= fl0;
              }
            }
          }
        }
      }
      // unmark
      if (need_c3) {
        __builtin_amdgcn_wave_barrier();
        for (long long k = yb + lane; k < ye; k += kWave) M[tl[k]] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; j++)
    if (j < njobs) flush(acc[j], jobs[j].out);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bytes += __shfl_xor(bytes, o, kWave);
  if (lane == 0) atomicAdd((ull*)&jobs[0].out->bytes, bytes);
}

// One sweep over the directed CSR entries evaluating every edge-core motif of the
// closed-form kind (FC:1999-2051: c_uv is the support t_uv, read from the index) and
// every star motif (FC:2152-2169, evaluated at each vertex's first entry). Lanes take
// consecutive entries; the owner of an entry comes from a search over the item's
// vertex range.
template <int NJ>
__global__ __launch_bounds__(256) void k_edge_pass(const int* __restrict__ idx, const int* __restrict__ adj,
                                                  const int* __restrict__ tsup, const int* __restrict__ adjdeg,
                                                  const JobDesc* __restrict__ jobs,
                                                  int njobs, const WorkItem* __restrict__ items, int nitems,
                                                  int* __restrict__ head)
{
  const int lane = lane_id();
  const G g{idx, adj};
  LaneAcc acc[NJ];
  const int smod = jobs[0].smod, srem = jobs[0].srem;
  ull bytes = 0;  // algorithmic global-memory bytes, as in the rank kernel
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(head, 1);
    it = bcast(it, 0);
    if (it >= nitems) break;
    const WorkItem I = items[it];
    if (lane == 0) bytes += sizeof(WorkItem);
    for (long long j = I.begin + lane; j < I.end; j += kWave) {
      const int u = owner_of(idx, I.u0, I.u1, j);
      bytes += 4 * (32 - __clz(I.u1 - I.u0 + 1)) + 28;
      if (smod > 1 && (u % smod) != srem) continue;
      const int v = adj[j];
      const int du = g.deg(u), dv = adjdeg[j];
      const bool first = (j == idx[u]);
      const int c = tsup[j];
#pragma unroll
      for (int k = 0; k < NJ; k++) {
        if (k >= njobs) continue;
        const JobDesc* J = jobs + k;
        if (J->cls == kVertex) {
          if (first && du >= J->p[0]) { acc[k].add(ncr(du, J->p[0], acc[k].fl)); acc[k].matches++; }
        } else {
          const int* p = J->p;  // f_u, f_uv, f_v, lim1, lim2 (FC:2054-2103 filters)
          if (v < u && du > p[3] && dv > p[3] && du + dv > p[4]) {
            acc[k].add(pair_formula(du, dv, c, 1, p[0], p[1], p[2], acc[k].fl));
            acc[k].matches++;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NJ; k++)
    if (k < njobs) flush(acc[k], jobs[k].out);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bytes += __shfl_xor(bytes, o, kWave);
  if (lane == 0) atomicAdd((ull*)&jobs[0].out->bytes, bytes);
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

bool triangle_schedule_ok(const JobDesc& j)
{
  return j.cls == kTriangle || (j.cls == kEdge && (j.kind == kEdgeTriangle || j.kind == kEdgeTailed));
}

bool edgepass_schedule_ok(const JobDesc& j)
{
  return (j.cls == kEdge && j.kind == kEdgeGeneral) || (j.cls == kVertex && j.kind == kVertexStars);
}

// Grid of the rank launch (before clamping to the items) and its device scratch: one
// byte-mark array of maxdeg entries per wave for vertices whose list exceeds LDS.
static int triangle_grid(const Tune& t)
{
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return t.grid > 0 ? t.grid : cus * 4;
}

size_t triangle_scratch_bytes(const DeviceGraph& g, const Tune& t)
{
  const size_t mstride = ((size_t)g.maxdeg + 63) / 64 * 64;
  return mstride * (size_t)triangle_grid(t) * dev::kTriWaves;  // the work items stay with the graph
}

size_t edgepass_scratch_bytes(const DeviceGraph& g, const Tune& t)
{
  const size_t step = (size_t)std::max<long long>(256, (long long)t.chunk * 32);
  return sizeof(WorkItem) * ((size_t)g.m / step + 2) + 4096;
}

// Prepares one sweep over the edges for up to kMaxTriJobs closed-form edge and star jobs.
int plan_edgepass(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream_v, LaunchPlan* lp,
                  std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int nj = (int)jobs.size();
  if (nj > dev::kMaxTriJobs) { if (err) *err = "too many jobs for one edge pass"; return -7; }
  int rc = device_derived_build(g, stream_v, &lp->index_ms, err, false);
  if (rc) return rc;
  for (auto& j : jobs) { j.idx = g->idx; j.adj = g->adj; j.n = g->n; j.m = (int32_t)g->m; }
  std::vector<WorkItem> items;
  const long long step = std::max<long long>(256, (long long)t.chunk * 32);
  const std::vector<int32_t>& hi = g->h_idx;
  for (long long b = 0; b < g->m; b += step) {
    const long long e = std::min<long long>(g->m, b + step);
    const int u0 = (int)(std::upper_bound(hi.begin(), hi.end(), (int32_t)b) - hi.begin()) - 1;
    const int u1 = (int)(std::upper_bound(hi.begin(), hi.end(), (int32_t)(e - 1)) - hi.begin()) - 1;
    items.push_back(WorkItem{0, u0, u1, 0, b, e});
  }
  sample_items(&items, t.sample);
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int grid = t.grid > 0 ? t.grid : cus * 8;
  grid = std::max(1, std::min<int>(grid, ((int)items.size() + 3) / 4));
  rc = plan_accumulators(lp, nj, stream, err);
  if (rc) return rc;
  JobDesc* d_jobs = nullptr;
  WorkItem* d_items = nullptr;
  for (int k = 0; k < nj; k++) jobs[k].out = lp->d_acc + k;
  HIP_TRY(lp->scratch->get(&d_jobs, sizeof(JobDesc) * std::max(1, nj)));
  HIP_TRY(lp->scratch->get(&d_items, sizeof(WorkItem) * std::max<size_t>(1, items.size())));
  if (nj) HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
  if (!items.empty())
    HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));  // the host arrays go out of scope
  const int nitems = (int)items.size();
  if (nitems == 0) return 0;
  const int* idx = g->idx;
  const int* adj = g->adj;
  const int* tsup = g->tsup;
  const int* adjdeg = g->adjdeg;
  lp->nitems = nitems;
  lp->kernel = [=](void* sv, int first, int count, int* head) {
    hipStream_t s = (hipStream_t)sv;
    if (nj == 1)
      hipLaunchKernelGGL((dev::k_edge_pass<1>), dim3(grid), dim3(256), 0, s, idx, adj, tsup, adjdeg, d_jobs, nj,
                         d_items + first, count, head);
    else
      hipLaunchKernelGGL((dev::k_edge_pass<dev::kMaxTriJobs>), dim3(grid), dim3(256), 0, s, idx, adj, tsup, adjdeg,
                         d_jobs, nj, d_items + first, count, head);
  };
  return 0;
}

int device_run_edgepass(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream,
                        std::vector<JobResult>* out, double* kernel_ms, std::string* err)
{
  out->assign(jobs.size(), JobResult());
  if (kernel_ms) *kernel_ms = 0;
  if (jobs.empty()) return 0;
  LaunchPlan lp;
  int rc = plan_edgepass(jobs, g, t, stream, &lp, err);
  return rc ? rc : launch_run(lp, stream, out, kernel_ms, err);
}

// ---------------------------------------------------------------- no-allocation count
//
// One edge pass for one motif with every piece of device scratch it needs (work-queue
// head, accumulator, job descriptor, work items) in a fixed array of this file, so the
// call allocates nothing. The array's device address comes from hipGetSymbolAddress,
// once per device; kernels and copies only ever see that address. Calls on one device
// share the array: a mutex serialises them within the process.
constexpr size_t kNoAllocScratch = 4u << 20;  // bytes
constexpr size_t kNoAllocJobOff = 512;         // [0, 512): queue head (0) and the accumulator (256)
constexpr size_t kNoAllocItemsOff = kNoAllocJobOff + (sizeof(JobDesc) + 255) / 256 * 256;
constexpr long long kNoAllocStep = 8192;       // CSR positions per work item, as the allocating edge pass
constexpr int kNoAllocMaxItems = (int)((kNoAllocScratch - kNoAllocItemsOff) / sizeof(WorkItem));
constexpr int kNoAllocMaxDevices = 64;

namespace dev {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

// clears the head and the accumulator, and writes the work items of the edge pass:
// item i covers CSR positions [i * step, min(m, (i + 1) * step)) with the owners of
// its first and last position (the last vertex u with idx[u] <= position)
__global__ void k_noalloc_setup(const int* __restrict__ idx, int n, long long m, long long step,
                                unsigned char* __restrict__ base, int nitems)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < (int)(kNoAllocJobOff / 8)) reinterpret_cast<ull*>(base)[t] = 0ull;
  if (t >= nitems) return;
  WorkItem* items = reinterpret_cast<WorkItem*>(base + kNoAllocItemsOff);
  const long long b = (long long)t * step, e = min(m, b + step);
  auto owner = [&](long long pos) {
    int lo = 0, hi = n;  // idx[lo] <= pos < idx[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((long long)idx[mid] <= pos) lo = mid; else hi = mid;
    }
    return lo;
  };
  items[t] = WorkItem{0, owner(b), owner(e - 1), 0, b, e};
}

}  // namespace dev

int edgepass_count_noalloc(const DeviceGraph* g, JobDesc job, JobResult* out, double* kernel_ms, std::string* err)
{
  static std::mutex mu;
  static unsigned char* ptr[kNoAllocMaxDevices] = {};
  std::lock_guard<std::mutex> lk(mu);
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) { if (err) *err = "device index out of range"; return -19; }
  if (!ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(dev::scratch)));
    ptr[dev] = (unsigned char*)p;
  }
  unsigned char* base = ptr[dev];
  const long long m = g->m;
  const long long nitems = (m + kNoAllocStep - 1) / kNoAllocStep;
  if (nitems > kNoAllocMaxItems) {
    if (err) *err = "graph too large for the no-allocation scratch: " + std::to_string(m) + " directed edges, at most " +
                    std::to_string((long long)kNoAllocMaxItems * kNoAllocStep);
    return -27;
  }
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  job.idx = g->idx; job.adj = g->adj; job.n = g->n; job.m = (int32_t)m;
  job.out = reinterpret_cast<Accum*>(base + 256);
  const auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipMemcpy(base + kNoAllocJobOff, &job, sizeof(JobDesc), hipMemcpyHostToDevice));
  const int setup = std::max<int>((int)nitems, (int)(kNoAllocJobOff / 8));
  hipLaunchKernelGGL(dev::k_noalloc_setup, dim3((setup + 255) / 256), dim3(256), 0, 0, (const int*)g->idx, g->n, m,
                     kNoAllocStep, base, (int)nitems);
  HIP_TRY(hipGetLastError());
  if (nitems > 0) {
    const int grid = std::max(1, std::min<int>(cus * 8, (int)((nitems + 3) / 4)));
    hipLaunchKernelGGL((dev::k_edge_pass<1>), dim3(grid), dim3(256), 0, 0, (const int*)g->idx, (const int*)g->adj,
                       (const int*)g->tsup, (const int*)g->adjdeg, reinterpret_cast<const JobDesc*>(base + kNoAllocJobOff),
                       1, reinterpret_cast<const WorkItem*>(base + kNoAllocItemsOff), (int)nitems,
                       reinterpret_cast<int*>(base));
    HIP_TRY(hipGetLastError());
  }
  Accum a;
  HIP_TRY(hipMemcpy(&a, base + 256, sizeof(Accum), hipMemcpyDeviceToHost));  // waits for the kernels
  if (kernel_ms) *kernel_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  out->sum = ((unsigned __int128)a.hi << 64) | a.lo;
  out->matches = a.matches;
  out->flags = a.flags;
  out->bytes = a.bytes;
  out->index_ms = 0;
  return 0;
}

// Prepares one rank-schedule launch of up to kMaxTriJobs triangle-based jobs.
int plan_triangles(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream_v, LaunchPlan* lp,
                   std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int nj = (int)jobs.size();
  if (nj > dev::kMaxTriJobs) { if (err) *err = "too many jobs for one triangle launch"; return -7; }
  int tails = -1;
  for (auto& j : jobs) {
    if (j.cls == kTriangle && j.kind == kTriTailedClique) {
      if (tails >= 0 && tails != j.p[3]) { if (err) *err = "tailed 4-clique jobs with different tails"; return -22; }
      tails = j.p[3];
    }
  }
  int rc = device_derived_build(g, stream_v, &lp->index_ms, err, true);
  if (rc) return rc;
  bool need_c3 = false, need_lt = false, full = false;
  for (auto& j : jobs) {
    j.idx = g->idx; j.adj = g->adj; j.n = g->n; j.m = (int32_t)g->m;
    if (j.cls == kTriangle) {
      need_c3 = true;
      if (j.kind != kTriGeneral) need_lt = true;
      // fringes anchored on a pair of core vertices need the full fc3; single-vertex and
      // all-three anchor sets have the inline form
      for (int r = 0; r < j.rot; r++) full = full || (j.fr[r][3] | j.fr[r][5] | j.fr[r][6]) != 0;
    }
  }
  // items: vertices x with their positions split into bounded pieces; they depend on the
  // graph, the chunk and the sample only, so they are built once per graph and key and
  // kept on the device until the graph is freed
  // default item size: 64 positions when the launch counts triple intersections (C4:
  // 935 -> 831 ms, C3's triangle cores 1-6% faster), else 256 (the triangle kinds of
  // the edge core: 74 ms against 122 ms with 64)
  const int chunk = t.chunk_set ? t.chunk : (need_c3 ? 64 : 256);
  const dev::TriItem* d_items = nullptr;
  int nitems = 0;
  {
    static std::mutex items_mu;
    std::lock_guard<std::mutex> lk(items_mu);
    const long long key = ((long long)chunk << 32) | (unsigned)std::max(1, t.sample);
    auto it = g->tri_items.find(key);
    if (it == g->tri_items.end()) {
      std::vector<dev::TriItem> items;
      const long long budget = std::max<long long>(64, chunk);
      for (int x = 0; x < g->n; x++) {
        const int dx = g->h_idx[x + 1] - g->h_idx[x];
        if (dx < 2) continue;
        const int per = (int)std::max<long long>(1, budget * 64 / std::max(64, dx));
        for (int p = 0; p < dx; p += per) items.push_back(dev::TriItem{x, p, std::min(dx, p + per), 0});
      }
      sample_items(&items, t.sample);
      void* d = nullptr;
      HIP_TRY(hipMalloc(&d, sizeof(dev::TriItem) * std::max<size_t>(1, items.size())));
      if (!items.empty())
        HIP_TRY(hipMemcpy(d, items.data(), sizeof(dev::TriItem) * items.size(), hipMemcpyHostToDevice));
      it = g->tri_items.emplace(key, std::make_pair(d, (int)items.size())).first;
    }
    d_items = (const dev::TriItem*)it->second.first;
    nitems = it->second.second;
  }
  int grid = triangle_grid(t);
  grid = std::max(1, std::min<int>(grid, (nitems + dev::kTriWaves - 1) / dev::kTriWaves));
  const long long mstride = ((long long)g->maxdeg + 63) / 64 * 64;
  rc = plan_accumulators(lp, nj, stream, err);
  if (rc) return rc;
  unsigned char* gmark = nullptr;
  JobDesc* d_jobs = nullptr;
  HIP_TRY(lp->scratch->get(&gmark, (size_t)mstride * grid * dev::kTriWaves));
  for (int k = 0; k < nj; k++) jobs[k].out = lp->d_acc + k;
  HIP_TRY(lp->scratch->get(&d_jobs, sizeof(JobDesc) * std::max(1, nj)));
  // marks start at zero; every launch unmarks what it marked
  HIP_TRY(hipMemsetAsync(gmark, 0, (size_t)mstride * grid * dev::kTriWaves, stream));
  if (nj) HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));  // the host job array goes out of scope
  if (nitems == 0) return 0;
  const int* idx = g->idx;
  const int* adj = g->adj;
  const int* tsup = g->tsup;
  const int* adjdeg = g->adjdeg;
  const int* rev = g->rev;
  const long long* tl_off = g->tl_off;
  const int* tl = g->tl;
  const int c3 = need_c3 ? 1 : 0, lt = need_lt ? 1 : 0, tl_arg = std::max(0, tails);
  // Tune::lds_cap > 0 caps the degree up to which a vertex's marks sit in LDS (larger
  // ones use the global mark array); the results do not depend on it
  const int markcap = t.lds_cap > 0 ? std::min(t.lds_cap, dev::kMarkCap) : dev::kMarkCap;
  lp->nitems = nitems;
  lp->kernel = [=](void* sv, int first, int count, int* head) {
    hipStream_t s = (hipStream_t)sv;
#define FSGC_TRI_LAUNCH(NJ, FULL)                                                                               \
  hipLaunchKernelGGL((dev::k_triangle_rank<NJ, FULL>), dim3(grid), dim3(dev::kTriBlock), 0, s, idx, adj, tsup,  \
                     adjdeg, rev, tl_off, tl, d_jobs, nj, c3, lt, tl_arg, d_items + first, count, head, gmark, mstride, \
                     markcap, nullptr)
    if (nj == 1) {
      if (full) FSGC_TRI_LAUNCH(1, true); else FSGC_TRI_LAUNCH(1, false);
    } else if (nj == 2) {
      if (full) FSGC_TRI_LAUNCH(2, true); else FSGC_TRI_LAUNCH(2, false);
    } else {
      if (full) FSGC_TRI_LAUNCH(dev::kMaxTriJobs, true); else FSGC_TRI_LAUNCH(dev::kMaxTriJobs, false);
    }
#undef FSGC_TRI_LAUNCH
  };
  return 0;
}

// The triple count of every triangle {a, b, c} (a < b < c by rank) at c's entry of
// T(a, b): one rank-schedule pass with no jobs that writes the c3 it computes. Built
// once per graph, for the clique schedule (4 bytes per list entry).
int device_c3tl_build(DeviceGraph* g, void* stream_v, double* ms, std::string* err)
{
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  if (ms) *ms = 0;
  if (g->c3tl) return 0;
  hipStream_t stream = (hipStream_t)stream_v;
  int rc = device_derived_build(g, stream_v, nullptr, err, true);
  if (rc) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  int32_t* c3 = nullptr;
  HIP_TRY(hipMalloc((void**)&c3, sizeof(int32_t) * (size_t)std::max<long long>(1, g->tl_total)));
  {
    DevScratch sc(stream);
    std::vector<dev::TriItem> items;
    for (int x = 0; x < g->n; x++) {
      const int dx = g->h_idx[x + 1] - g->h_idx[x];
      if (dx < 2) continue;
      const int per = std::max(1, 256 * 64 / std::max(64, dx));
      for (int p = 0; p < dx; p += per) items.push_back(dev::TriItem{x, p, std::min(dx, p + per), 0});
    }
    int dev = 0, cus = 256;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::max(1, std::min<int>(cus * 4, ((int)items.size() + dev::kTriWaves - 1) / dev::kTriWaves));
    const long long mstride = ((long long)g->maxdeg + 63) / 64 * 64;
    JobDesc job;
    std::memset(&job, 0, sizeof(job));
    job.smod = 1;
    unsigned char* gmark = nullptr;
    JobDesc* d_job = nullptr;
    Accum* d_acc = nullptr;
    dev::TriItem* d_items = nullptr;
    int* head = nullptr;
    HIP_TRY(sc.get(&gmark, (size_t)mstride * grid * dev::kTriWaves));
    HIP_TRY(sc.get(&d_job, sizeof(JobDesc)));
    HIP_TRY(sc.get(&d_acc, sizeof(Accum)));
    HIP_TRY(sc.get(&head, sizeof(int)));
    HIP_TRY(sc.get(&d_items, sizeof(dev::TriItem) * std::max<size_t>(1, items.size())));
    job.out = d_acc;
    HIP_TRY(hipMemsetAsync(c3, 0, sizeof(int32_t) * (size_t)std::max<long long>(1, g->tl_total), stream));
    HIP_TRY(hipMemsetAsync(gmark, 0, (size_t)mstride * grid * dev::kTriWaves, stream));
    HIP_TRY(hipMemsetAsync(d_acc, 0, sizeof(Accum), stream));
    HIP_TRY(hipMemsetAsync(head, 0, sizeof(int), stream));
    HIP_TRY(hipMemcpyAsync(d_job, &job, sizeof(JobDesc), hipMemcpyHostToDevice, stream));
    if (!items.empty())
      HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(dev::TriItem) * items.size(), hipMemcpyHostToDevice,
                             stream));
    if (!items.empty()) {
      hipLaunchKernelGGL((dev::k_triangle_rank<1, true>), dim3(grid), dim3(dev::kTriBlock), 0, stream,
                         (const int*)g->idx, (const int*)g->adj, (const int*)g->tsup, (const int*)g->adjdeg,
                         (const int*)g->rev, (const long long*)g->tl_off, (const int*)g->tl, (const JobDesc*)d_job, 0, 1, 0, 0,
                         (const dev::TriItem*)d_items, (int)items.size(), head, gmark, mstride, dev::kMarkCap, c3);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(stream));
  }
  g->c3tl = c3;
  g->derived_bytes += sizeof(int32_t) * (size_t)std::max<long long>(1, g->tl_total);
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Runs up to kMaxTriJobs triangle-based jobs (same graph) in one rank-schedule launch.
int device_run_triangles(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream,
                         std::vector<JobResult>* out, double* kernel_ms, std::string* err)
{
  out->assign(jobs.size(), JobResult());
  if (kernel_ms) *kernel_ms = 0;
  if (jobs.empty()) return 0;
  LaunchPlan lp;
  int rc = plan_triangles(jobs, g, t, stream, &lp, err);
  return rc ? rc : launch_run(lp, stream, out, kernel_ms, err);
}

}  // namespace fsgc
