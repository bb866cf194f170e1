// 4-clique cores on the rank schedule (gfx950).
//
// A general core whose four vertices are pairwise adjacent (FC:489-614 with
// list = {1, 3, 7}) is counted by the reference once per ordered embedding: every
// permutation of every 4-clique that passes the per-position degree filters, with
// commonX (FC:345-405) walking the anchors' lists and functionX (FC:408-486)
// enumerating fringe placements. Here each 4-clique {x, y, z, w} (x < y < z < w by
// rank = (degree, id)) is visited once, from x:
//   * for each higher-ranked neighbour y of x, the list T(x, y) (common neighbours,
//     positions in N(x)) is marked in LDS with each entry's index;
//   * a lane takes z in T(x, y) ranked above y and scans T(x, z): marked entries are
//     the common neighbours of x, y, z; those ranked above z close a clique with w;
//   * the intersection sizes of the clique's neighbourhoods come from the per-edge
//     index: degrees, the six supports (the pairs' entries found through the twin
//     lists T(y, x), T(z, x), which list the same vertices in the same order), the four
//     triple counts (c3tl, written by a rank pass for every triangle at its highest
//     vertex's entry of the list of its two lower ones; the triangle {y, z, w} by one
//     search in T(y, z)), and the quadruple count as the popcount of two bit rows:
//     row z holds, over the indices of T(x, y), the common neighbours of x, y, z
//     (written by z's lane in a first pass);
//   * N[U] = |intersection of N(v), v in U| - (4 - |U|) non-core vertices are adjacent to
//     every clique vertex of U; the fringe count of an assignment of the clique's
//     vertices to the core positions is an inclusion-exclusion over coinciding fringes
//     (tables built on the host, one per class of assignments that give equal counts),
//     evaluated modulo 2^128 and divided exactly (the host admits a motif only where
//     the true value is below 2^128).
// The sum over cliques and assignments is the reference's sum over ordered embeddings,
// with the same degree filters and the same number of counted matches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "devmath.h"
#include "devscratch.h"
#include "fsgc_plan.h"

namespace fsgc {
namespace dev {

constexpr int kCqBlock = 256;
constexpr int kCqWaves = kCqBlock / kWave;
constexpr int kCqMarkCap = 4096;  // LDS marks per wave (positions of N(x))

struct CqItem {
  int32_t x;
  int32_t p0;  // positions [p0, p1) of N(x) taken as y
  int32_t p1;
  int32_t pad;
};

__device__ __forceinline__ bool rank_lt(int da, int a, int db, int b) { return da < db || (da == db && a < b); }

__device__ __forceinline__ int lbound_i(const int* __restrict__ a, int l, int r, int x)
{
  while (l < r) {
    const int mid = (l + r) >> 1;
    if (a[mid] < x) l = mid + 1; else r = mid;
  }
  return l;
}

__device__ __forceinline__ long long lbound_ll(const int* __restrict__ a, long long l, long long r, int x)
{
  while (l < r) {
    const long long mid = (l + r) >> 1;
    if (a[mid] < x) l = mid + 1; else r = mid;
  }
  return l;
}

// F of assignment class c: (sum of the class's terms modulo 2^128) / den, exact
__device__ __forceinline__ u128 clique_class_value(const CliqueTab* __restrict__ T, int c,
                                                   const unsigned (*__restrict__ sN)[kCqBlock], int tid)
{
  u128 v = 0;
  const CliqueTerm* __restrict__ terms = T->terms;
  const int t1 = T->term_off[c + 1];
  for (int t = T->term_off[c]; t < t1; t++) {
    const CliqueTerm term = terms[t];
    const uint32_t pk = term.packed;
    const int nf = (int)(pk >> 28);
    ull a = 1;
    int f = 0;
    // N < 2^21: three factors fit 64 bits, six fit 128
    for (; f < nf && f < 3; f++) a *= (ull)sN[(pk >> (4 * f)) & 15][tid];
    u128 prod = a;
    if (f < nf) {
      ull b = 1;
      for (; f < nf && f < 6; f++) b *= (ull)sN[(pk >> (4 * f)) & 15][tid];
      prod = (u128)a * b;
      for (; f < nf; f++) prod = mul_lo128(prod, (u128)sN[(pk >> (4 * f)) & 15][tid]);
    }
    v += mul_lo128(prod, (u128)(__int128)term.coef);
  }
  return mul_lo128(v >> T->den_shift, ((u128)T->den_inv_hi << 64) | T->den_inv_lo);
}

// Every job on one clique whose vertices' degrees and ids are d[0..3], K[0..3] (rank
// order) and whose N values this lane holds in sN[.][tid].
template <int NJ>
__device__ __forceinline__ void clique_jobs(const CliqueTab* __restrict__ tabs, int njobs, int smod, int srem,
                                            const int d[4], const int K[4], const unsigned (*__restrict__ sN)[kCqBlock],
                                            int tid, LaneAcc* acc)
{
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    if (j >= njobs) continue;
    const CliqueTab* __restrict__ T = tabs + j;
    // the assignments that pass the degree filters (FC:617-663, FC:489-564) and the slice
    uint32_t pass = 0;
    for (int i = 0; i < 24; i++) {
      const int a0 = T->perm[i][0], a1 = T->perm[i][1], a2 = T->perm[i][2], a3 = T->perm[i][3];
      bool ok = d[a0] >= T->thr[0] && d[a1] >= T->thr[1] && d[a2] >= T->thr[2] && d[a3] >= T->thr[3];
      if (ok && smod > 1) ok = (K[a0] % smod) == srem;
      pass |= (uint32_t)ok << i;
    }
    if (!pass) continue;
    acc[j].matches += __popc(pass);
    for (int c = 0; c < T->ncls; c++) {
      const int cnt = __popc(pass & T->cls_perms[c]);
      if (!cnt) continue;
      const u128 val = clique_class_value(T, c, sN, tid);
      acc[j].add(mulx(val, (u128)(unsigned)cnt, acc[j].fl));
      for (int s = 0; s < T->nck_n; s++)
        if ((int)sN[T->nck_mask[c][s]][tid] >= T->nck_thr[s]) acc[j].fl |= kRefNckInexact;
    }
  }
}

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int NJ>
__global__ __launch_bounds__(kCqBlock) void k_clique_rank(const int* __restrict__ idx, const int* __restrict__ adj,
                                                          const int* __restrict__ tsup, const int* __restrict__ adjdeg,
                                                          const long long* __restrict__ tl_off,
                                                          const int* __restrict__ tl, const int* __restrict__ c3tl,
                                                          const JobDesc* __restrict__ jobs,
                                                          const CliqueTab* __restrict__ tabs, int njobs,
                                                          const CqItem* __restrict__ items, int nitems,
                                                          int* __restrict__ head, int* __restrict__ gmark,
                                                          int* __restrict__ gzmap, ull* __restrict__ grows,
                                                          long long mstride, long long rowcap, int markcap)
{
  __shared__ int s_mark[kCqWaves][kCqMarkCap];
  __shared__ unsigned s_N[16][kCqBlock];
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const int tid = threadIdx.x;
  const long long wid = (long long)blockIdx.x * kCqWaves + wave;
  int* __restrict__ Mg = gmark + wid * mstride;
  int* __restrict__ Z = gzmap + wid * mstride;
  ull* __restrict__ rows = grows + wid * rowcap;
  LaneAcc acc[NJ];
  const int smod = jobs[0].smod, srem = jobs[0].srem;
  for (int i = lane; i < kCqMarkCap; i += kWave) s_mark[wave][i] = 0;
  s_N[0][tid] = 0;
  wave_sync();

  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(head, 1);
    it = bcast(it, 0);
    if (it >= nitems) break;
    const CqItem I = items[it];
    const int x = I.x, xb = idx[x], dx = idx[x + 1] - xb;
    int* __restrict__ M = dx <= markcap ? s_mark[wave] : Mg;
    for (int py = I.p0; py < I.p1; py++) {
      const int y = adj[xb + py];
      const int dy = adjdeg[xb + py];
      if (!rank_lt(dx, x, dy, y)) continue;
      const long long yb = tl_off[xb + py];
      const int t = (int)(tl_off[xb + py + 1] - yb);
      if (t < 2) continue;
      const int txy = tsup[xb + py];
      const int yb2 = idx[y];
      // T(y, x): the same common neighbours as T(x, y), as positions in N(y)
      const long long yxb = tl_off[lbound_i(adj, yb2, idx[y + 1], x)];
      // mark T(x, y) with each entry's index + 1; Z[j] = rank among the entries ranked
      // above y (the row of z = entry j), or -1
      int nz = 0;
      for (int j0 = 0; j0 < t; j0 += kWave) {
        const int j = j0 + lane;
        bool inz = false;
        if (j < t) {
          const int p = tl[yb + j];
          M[p] = j + 1;
          inz = rank_lt(dy, y, adjdeg[xb + p], adj[xb + p]);
        }
        const ull bm = __ballot(inz);
        if (j < t) Z[j] = inz ? nz + __popcll(bm & ((1ull << lane) - 1)) : -1;
        nz += __popcll(bm);
      }
      wave_sync();
      if (nz >= 2) {
        const int R = (t + 63) >> 6;
        const bool use_rows = (long long)nz * R <= rowcap;
        // pass 1: row of z = the common neighbours of x, y, z as bits over T(x, y)
        if (use_rows) {
          for (int j0 = 0; j0 < t; j0 += kWave) {
            const int j = j0 + lane;
            if (j >= t) continue;
            const int zi = Z[j];
            if (zi < 0) continue;
            ull* __restrict__ row = rows + (long long)zi * R;
            for (int r = 0; r < R; r++) row[r] = 0;
            const int pz = tl[yb + j];
            const long long zb = tl_off[xb + pz], ze = tl_off[xb + pz + 1];
            for (long long q = zb; q < ze; q++) {
              const int mk = M[tl[q]];
              if (mk) row[(mk - 1) >> 6] |= 1ull << ((mk - 1) & 63);
            }
          }
          wave_sync();
        }
        // pass 2: cliques (x, y, z, w), one lane per z
        for (int j0 = 0; j0 < t; j0 += kWave) {
          const int j = j0 + lane;
          if (j >= t) continue;
          const int zi = Z[j];
          if (zi < 0) continue;
          const int pz = tl[yb + j];
          const int z = adj[xb + pz], dz = adjdeg[xb + pz];
          const int txz = tsup[xb + pz];
          const long long zb = tl_off[xb + pz], ze = tl_off[xb + pz + 1];
          const int zb2 = idx[z];
          const long long zxb = tl_off[lbound_i(adj, zb2, idx[z + 1], x)];
          const int pyz = tl[yxb + j];
          const int tyz = tsup[yb2 + pyz];
          const long long yzb = tl_off[yb2 + pyz], yze = tl_off[yb2 + pyz + 1];
          const int c3xyz = c3tl[yb + j];
          const ull* __restrict__ rz = rows + (long long)zi * R;
          for (long long q = zb; q < ze; q++) {
            const int pw = tl[q];
            const int mk = M[pw];
            if (!mk) continue;
            const int jw = mk - 1;
            if (Z[jw] < 0) continue;  // w ranked below y
            const int w = adj[xb + pw], dw = adjdeg[xb + pw];
            if (!rank_lt(dz, z, dw, w)) continue;
            // the clique (x, y, z, w)
            const int txw = tsup[xb + pw];
            const int tzw = tsup[zb2 + tl[zxb + (q - zb)]];
            const int pyw = tl[yxb + jw];
            const int tyw = tsup[yb2 + pyw];
            const int c3xyw = c3tl[yb + jw];
            const int c3xzw = c3tl[q];
            const int c3yzw = c3tl[lbound_ll(tl, yzb, yze, pyw)];
            int c4 = 0;
            if (use_rows) {
              const ull* __restrict__ rw = rows + (long long)Z[jw] * R;
              for (int r = 0; r < R; r++) c4 += __popcll(rz[r] & rw[r]);
            } else {
              const long long wb = tl_off[xb + pw], we = tl_off[xb + pw + 1];
              for (long long q2 = zb; q2 < ze; q2++) {
                const int p2 = tl[q2];
                if (!M[p2]) continue;
                const long long h = lbound_ll(tl, wb, we, p2);
                c4 += (h < we && tl[h] == p2) ? 1 : 0;
              }
            }
            // N[U], U a mask over (x, y, z, w) = bits 0..3
            s_N[1][tid] = dx - 3;
            s_N[2][tid] = dy - 3;
            s_N[4][tid] = dz - 3;
            s_N[8][tid] = dw - 3;
            s_N[3][tid] = txy - 2;
            s_N[5][tid] = txz - 2;
            s_N[9][tid] = txw - 2;
            s_N[6][tid] = tyz - 2;
            s_N[10][tid] = tyw - 2;
            s_N[12][tid] = tzw - 2;
            s_N[7][tid] = c3xyz - 1;
            s_N[11][tid] = c3xyw - 1;
            s_N[13][tid] = c3xzw - 1;
            s_N[14][tid] = c3yzw - 1;
            s_N[15][tid] = c4;
            const int d[4] = {dx, dy, dz, dw};
            const int K[4] = {x, y, z, w};
            clique_jobs<NJ>(tabs, njobs, smod, srem, d, K, s_N, tid, acc);
          }
        }
      }
      wave_sync();
      for (int j = lane; j < t; j += kWave) M[tl[yb + j]] = 0;
      wave_sync();
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; j++)
    if (j < njobs) flush(acc[j], jobs[j].out);
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

// bit rows per wave (words): nz rows of ceil(|T(x, y)| / 64) words; a pair (x, y) whose
// rows do not fit counts its quadruple intersections by searches instead
constexpr long long kCliqueRowCap = 1ll << 17;

static int clique_grid(const Tune& t)
{
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return t.grid > 0 ? t.grid : cus * 2;
}

size_t clique_scratch_bytes(const DeviceGraph& g, const Tune& t)
{
  const size_t mstride = ((size_t)g.maxdeg + 63) / 64 * 64;
  const size_t waves = (size_t)clique_grid(t) * dev::kCqWaves;
  return waves * (2 * sizeof(int) * mstride + sizeof(dev::ull) * (size_t)kCliqueRowCap) +
         sizeof(dev::CqItem) * ((size_t)g.n + 1024) + kMaxCliqueJobs * (sizeof(CliqueTab) + 16 * kCliqueMaxTerms);
}

// Prepares one launch of up to kMaxCliqueJobs 4-clique jobs (same graph).
int plan_cliques(std::vector<JobDesc>& jobs, DeviceGraph* g, const Tune& t, void* stream_v, LaunchPlan* lp,
                 std::string* err)
{
  hipStream_t stream = (hipStream_t)stream_v;
  const int nj = (int)jobs.size();
  if (nj > kMaxCliqueJobs) { if (err) *err = "too many jobs for one clique launch"; return -7; }
  int rc = device_derived_build(g, stream_v, &lp->index_ms, err, true);
  if (rc) return rc;
  double c3ms = 0;
  rc = device_c3tl_build(g, stream_v, &c3ms, err);
  if (rc) return rc;
  lp->index_ms += c3ms;
  std::vector<CliqueTab> tabs(nj);
  std::vector<std::vector<CliqueTerm>> terms(nj);
  for (int k = 0; k < nj; k++) {
    if (!clique_fits(jobs[k], g->maxdeg)) { if (err) *err = "motif not countable on the clique schedule"; return -22; }
    rc = build_clique_tab(jobs[k], &tabs[k], &terms[k], err);
    if (rc) return rc;
    if (jobs[k].smod != jobs[0].smod || jobs[k].srem != jobs[0].srem) {
      if (err) *err = "clique jobs of one launch take the same root slice";
      return -22;
    }
    jobs[k].idx = g->idx; jobs[k].adj = g->adj; jobs[k].n = g->n; jobs[k].m = (int32_t)g->m;
  }
  std::vector<dev::CqItem> items;
  const long long budget = std::max<long long>(64, t.chunk);
  for (int x = 0; x < g->n; x++) {
    const int dx = g->h_idx[x + 1] - g->h_idx[x];
    if (dx < 3) continue;
    const int per = (int)std::max<long long>(1, budget * 64 / std::max(64, dx));
    for (int p = 0; p < dx; p += per) items.push_back(dev::CqItem{x, p, std::min(dx, p + per), 0});
  }
  sample_items(&items, t.sample);
  int grid = clique_grid(t);
  grid = std::max(1, std::min<int>(grid, ((int)items.size() + dev::kCqWaves - 1) / dev::kCqWaves));
  const long long mstride = ((long long)g->maxdeg + 63) / 64 * 64;
  // FSGC_CLIQUE_ROWCAP=k (tests) lowers the bit-row budget per wave, sending more
  // pairs to the search path; the counts do not depend on it
  long long rowcap = kCliqueRowCap;
  if (const char* e = getenv("FSGC_CLIQUE_ROWCAP")) rowcap = std::max(1ll, std::min(kCliqueRowCap, atoll(e)));
  rc = plan_accumulators(lp, nj, stream, err);
  if (rc) return rc;
  int* gmark = nullptr;
  int* gzmap = nullptr;
  dev::ull* grows = nullptr;
  JobDesc* d_jobs = nullptr;
  CliqueTab* d_tabs = nullptr;
  CliqueTerm* d_terms = nullptr;
  dev::CqItem* d_items = nullptr;
  const size_t waves = (size_t)grid * dev::kCqWaves;
  HIP_TRY(lp->scratch->get(&gmark, sizeof(int) * (size_t)mstride * waves));
  HIP_TRY(lp->scratch->get(&gzmap, sizeof(int) * (size_t)mstride * waves));
  HIP_TRY(lp->scratch->get(&grows, sizeof(dev::ull) * (size_t)rowcap * waves));
  size_t nterm = 0;
  for (auto& v : terms) nterm += v.size();
  HIP_TRY(lp->scratch->get(&d_terms, sizeof(CliqueTerm) * std::max<size_t>(1, nterm)));
  std::vector<CliqueTerm> all;
  for (int k = 0; k < nj; k++) {
    tabs[k].terms = d_terms + all.size();
    all.insert(all.end(), terms[k].begin(), terms[k].end());
  }
  for (int k = 0; k < nj; k++) jobs[k].out = lp->d_acc + k;
  HIP_TRY(lp->scratch->get(&d_jobs, sizeof(JobDesc) * std::max(1, nj)));
  HIP_TRY(lp->scratch->get(&d_tabs, sizeof(CliqueTab) * std::max(1, nj)));
  HIP_TRY(lp->scratch->get(&d_items, sizeof(dev::CqItem) * std::max<size_t>(1, items.size())));
  // marks start at zero; every launch unmarks what it marked
  HIP_TRY(hipMemsetAsync(gmark, 0, sizeof(int) * (size_t)mstride * waves, stream));
  if (nj) {
    HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(d_tabs, tabs.data(), sizeof(CliqueTab) * nj, hipMemcpyHostToDevice, stream));
  }
  if (!all.empty())
    HIP_TRY(hipMemcpyAsync(d_terms, all.data(), sizeof(CliqueTerm) * all.size(), hipMemcpyHostToDevice, stream));
  if (!items.empty())
    HIP_TRY(hipMemcpyAsync(d_items, items.data(), sizeof(dev::CqItem) * items.size(), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipStreamSynchronize(stream));  // the host arrays go out of scope
  const int nitems = (int)items.size();
  if (nitems == 0) return 0;
  const int* idx = g->idx;
  const int* adj = g->adj;
  const int* tsup = g->tsup;
  const int* adjdeg = g->adjdeg;
  const long long* tl_off = g->tl_off;
  const int* tl = g->tl;
  const int* c3tl = g->c3tl;
  // Tune::lds_cap > 0 caps the degree up to which x's marks sit in LDS
  const int markcap = t.lds_cap > 0 ? std::min(t.lds_cap, dev::kCqMarkCap) : dev::kCqMarkCap;
  lp->nitems = nitems;
  lp->kernel = [=](void* sv, int first, int count, int* head) {
    hipStream_t s = (hipStream_t)sv;
#define FSGC_CQ_LAUNCH(NJ)                                                                                       \
  hipLaunchKernelGGL((dev::k_clique_rank<NJ>), dim3(grid), dim3(dev::kCqBlock), 0, s, idx, adj, tsup, adjdeg,    \
                     tl_off, tl, c3tl, d_jobs, d_tabs, nj, d_items + first, count, head, gmark, gzmap, grows,    \
                     mstride, rowcap, markcap)
    if (nj == 1) FSGC_CQ_LAUNCH(1);
    else if (nj == 2) FSGC_CQ_LAUNCH(2);
    else FSGC_CQ_LAUNCH(kMaxCliqueJobs);
#undef FSGC_CQ_LAUNCH
  };
  return 0;
}

}  // namespace fsgc
