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

constexpr int kCqBlock = 128;
constexpr int kCqWaves = kCqBlock / kWave;
constexpr int kCqMarkCap = 2048;  // LDS marks per wave (positions of N(x))
constexpr int kCqBuf = 2 * kWave;  // clique records per wave awaiting evaluation
// a record: N[U] for the masks U = 1..15 (rows 0..14), the degrees (15..18) and the ids
// (19..22) of the clique's vertices in rank order
constexpr int kRecRows = 23;

struct CqItem {
  int32_t x;
  int32_t p0;  // positions [p0, p1) of N(x) taken as y
  int32_t p1;
  int32_t pad;
};

// per wave: one batch of up to 64 entries z of T(x, y) ranked above y, whose lists
// T(x, z) are streamed as one sequence
struct CqBatch {
  long long zb[kWave];   // T(x, z) in tl
  long long zxb[kWave];  // T(z, x) in tl
  long long yzb[kWave];  // T(y, z) in tl
  long long yze[kWave];
  int off[kWave + 1];    // offsets of the lists in the sequence
  int z[kWave], dz[kWave], zb2[kWave], txz[kWave], tyz[kWave], c3[kWave], zi[kWave];
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

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// F of assignment class c for the record in column `slot`: (the class's terms summed
// modulo 2^128) / den, exact
__device__ __forceinline__ u128 clique_class_value(const CliqueTab* __restrict__ T, int c,
                                                   const unsigned (*__restrict__ rec)[kCqBuf], int slot)
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
    // N < 2^21: three factors fit 64 bits
    for (; f < nf && f < 3; f++) a *= (ull)rec[((pk >> (4 * f)) & 15) - 1][slot];
    u128 prod = a;
    if (f < nf) {
      ull b = 1;
      for (; f < nf && f < 6; f++) b *= (ull)rec[((pk >> (4 * f)) & 15) - 1][slot];
      prod = (u128)a * b;
      for (; f < nf; f++) prod = mul_lo128(prod, (u128)rec[((pk >> (4 * f)) & 15) - 1][slot]);
    }
    v += mul_lo128(prod, (u128)(__int128)term.coef);
  }
  return mul_lo128(v >> T->den_shift, ((u128)T->den_inv_hi << 64) | T->den_inv_lo);
}

// Every job on the records [0, n) of the buffer, one per lane: the assignments that
// pass the degree filters (FC:617-663, FC:489-564) and the slice, then per class of
// assignments its count times the number of them.
template <int NJ>
__device__ __forceinline__ void clique_eval(const CliqueTab* __restrict__ tabs, int njobs, int smod, int srem,
                                            const unsigned (*__restrict__ rec)[kCqBuf], int n, LaneAcc* acc)
{
  const int lane = lane_id();
  const bool on = lane < n;
  int d0 = 0, d1 = 0, d2 = 0, d3 = 0, k0 = 0, k1 = 0, k2 = 0, k3 = 0;
  if (on) {
    d0 = (int)rec[15][lane]; d1 = (int)rec[16][lane]; d2 = (int)rec[17][lane]; d3 = (int)rec[18][lane];
    k0 = (int)rec[19][lane]; k1 = (int)rec[20][lane]; k2 = (int)rec[21][lane]; k3 = (int)rec[22][lane];
  }
  auto deg = [&](int a) { return a == 0 ? d0 : a == 1 ? d1 : a == 2 ? d2 : d3; };
  auto vid = [&](int a) { return a == 0 ? k0 : a == 1 ? k1 : a == 2 ? k2 : k3; };
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    if (j >= njobs) continue;
    const CliqueTab* __restrict__ T = tabs + j;
    constexpr uint32_t kAll = (1u << 24) - 1;
    uint32_t pass = 0;
    const int dmin = min(min(d0, d1), min(d2, d3));
    if (on && dmin >= T->max_thr) {
      // every vertex meets every position's threshold; a slice keeps the assignments
      // whose root is in it
      if (smod <= 1) {
        pass = kAll;
      } else {
        if (k0 % smod == srem) pass |= T->root_perms[0];
        if (k1 % smod == srem) pass |= T->root_perms[1];
        if (k2 % smod == srem) pass |= T->root_perms[2];
        if (k3 % smod == srem) pass |= T->root_perms[3];
      }
    } else if (on) {
      for (int i = 0; i < 24; i++) {
        const int a0 = T->perm[i][0], a1 = T->perm[i][1], a2 = T->perm[i][2], a3 = T->perm[i][3];
        bool ok = deg(a0) >= T->thr[0] && deg(a1) >= T->thr[1] && deg(a2) >= T->thr[2] && deg(a3) >= T->thr[3];
        if (ok && smod > 1) ok = (vid(a0) % smod) == srem;
        pass |= (uint32_t)ok << i;
      }
    }
    acc[j].matches += __popc(pass);
    // a clique that passes every assignment takes the merged table of all of them
    const bool all = T->use_all && pass == kAll;
    if (__ballot(all)) {
      if (all) acc[j].add(clique_class_value(T, T->ncls, rec, lane));
    }
    for (int c = 0; c < T->ncls; c++) {
      const int cnt = __popc(pass & T->cls_perms[c]);
      if (cnt > 0)
        for (int s = 0; s < T->nck_n; s++)
          if ((int)rec[T->nck_mask[c][s] - 1][lane] >= T->nck_thr[s]) acc[j].fl |= kRefNckInexact;
      const bool need = cnt > 0 && !all;
      if (!__ballot(need)) continue;
      if (need) {
        const u128 val = clique_class_value(T, c, rec, lane);
        acc[j].add(mulx(val, (u128)(unsigned)cnt, acc[j].fl));
      }
    }
  }
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
  __shared__ CqBatch s_bat[kCqWaves];
  __shared__ unsigned s_rec[kCqWaves][kRecRows][kCqBuf];
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const long long wid = (long long)blockIdx.x * kCqWaves + wave;
  int* __restrict__ Mg = gmark + wid * mstride;
  int* __restrict__ Z = gzmap + wid * mstride;
  ull* __restrict__ rows = grows + wid * rowcap;
  CqBatch& B = s_bat[wave];
  unsigned (*__restrict__ rec)[kCqBuf] = s_rec[wave];
  LaneAcc acc[NJ];
  const int smod = jobs[0].smod, srem = jobs[0].srem;
  int nrec = 0;  // records in the buffer (wave-uniform)
  for (int i = lane; i < kCqMarkCap; i += kWave) s_mark[wave][i] = 0;
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
      // mark T(x, y) with each entry's index + 1; Z[j] = the row of z = entry j (its
      // rank among the entries ranked above y), or -1
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
        if (use_rows) {
          for (long long i = lane; i < (long long)nz * R; i += kWave) rows[i] = 0;
          wave_sync();
        }
        // sweep 0 writes the bit rows (row z: the common neighbours of x, y, z over the
        // indices of T(x, y)); sweep 1 takes the cliques, whose quadruple counts read
        // the rows of both z and w
        for (int sweep = use_rows ? 0 : 1; sweep < 2; sweep++) {
          for (int j0 = 0; j0 < t; j0 += kWave) {
            // the batch: entries j0 .. j0 + 63 of T(x, y) ranked above y
            const int j = j0 + lane;
            const bool inz = j < t && Z[j] >= 0;
            int len = 0;
            if (inz) {
              const int pz = tl[yb + j];
              const long long zb = tl_off[xb + pz];
              len = (int)(tl_off[xb + pz + 1] - zb);
              B.zb[lane] = zb;
              B.zi[lane] = Z[j];
              if (sweep == 1) {
                const int z = adj[xb + pz];
                const int zb2 = idx[z];
                const int pyz = tl[yxb + j];
                B.z[lane] = z;
                B.dz[lane] = adjdeg[xb + pz];
                B.zb2[lane] = zb2;
                B.txz[lane] = tsup[xb + pz];
                B.zxb[lane] = tl_off[lbound_i(adj, zb2, idx[z + 1], x)];
                B.tyz[lane] = tsup[yb2 + pyz];
                B.yzb[lane] = tl_off[yb2 + pyz];
                B.yze[lane] = tl_off[yb2 + pyz + 1];
                B.c3[lane] = c3tl[yb + j];
              }
            }
            int incl = len;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
              const int v = __shfl_up(incl, o, kWave);
              if (lane >= o) incl += v;
            }
            const int tot = __shfl(incl, kWave - 1, kWave);
            B.off[lane] = incl - len;
            if (lane == kWave - 1) B.off[kWave] = tot;
            wave_sync();
            // element g of the sequence is element g - off[k] of lane k's list, k the last
            // lane with off[k] <= g (empty lists share the next one's offset); k only grows
            int k = 0;
            {
              int lo = 0, hi = kWave;
              const int g0 = min(lane, max(tot - 1, 0));
              while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (B.off[mid] <= g0) lo = mid; else hi = mid;
              }
              k = lo;
            }
            for (int c0 = 0; c0 < tot; c0 += kWave) {
              const int g = c0 + lane;
              bool found = false;
              long long q = 0;
              int jw = 0, p = 0;
              if (g < tot) {
                while (B.off[k + 1] <= g) k++;
                q = B.zb[k] + (g - B.off[k]);
                p = tl[q];
                const int mk = M[p];
                if (mk) {
                  jw = mk - 1;
                  if (sweep == 0) {
                    atomicOr(&rows[(long long)B.zi[k] * R + (jw >> 6)], 1ull << (jw & 63));
                  } else if (Z[jw] >= 0) {  // w ranked above y
                    found = rank_lt(B.dz[k], B.z[k], adjdeg[xb + p], adj[xb + p]);
                  }
                }
              }
              if (sweep == 0) continue;
              const ull fm = __ballot(found);
              if (!fm) continue;
              if (found) {
                // the clique (x, y, z, w) into the next free record
                const int slot = nrec + __popcll(fm & ((1ull << lane) - 1));
                const int w = adj[xb + p], dw = adjdeg[xb + p];
                const int pyw = tl[yxb + jw];
                int c4 = 0;
                if (use_rows) {
                  const ull* __restrict__ rz = rows + (long long)B.zi[k] * R;
                  const ull* __restrict__ rw = rows + (long long)Z[jw] * R;
                  for (int r = 0; r < R; r++) c4 += __popcll(rz[r] & rw[r]);
                } else {
                  const long long wb = tl_off[xb + p], we = tl_off[xb + p + 1];
                  const long long zb = B.zb[k], ze = B.zb[k] + (B.off[k + 1] - B.off[k]);
                  for (long long q2 = zb; q2 < ze; q2++) {
                    const int p2 = tl[q2];
                    if (!M[p2]) continue;
                    const long long h = lbound_ll(tl, wb, we, p2);
                    c4 += (h < we && tl[h] == p2) ? 1 : 0;
                  }
                }
                // N[U] = |intersection of N(v), v in U| - (4 - |U|), U over (x, y, z, w)
                rec[0][slot] = dx - 3;
                rec[1][slot] = dy - 3;
                rec[2][slot] = txy - 2;
                rec[3][slot] = B.dz[k] - 3;
                rec[4][slot] = B.txz[k] - 2;
                rec[5][slot] = B.tyz[k] - 2;
                rec[6][slot] = B.c3[k] - 1;
                rec[7][slot] = dw - 3;
                rec[8][slot] = tsup[xb + p] - 2;
                rec[9][slot] = tsup[yb2 + pyw] - 2;
                rec[10][slot] = c3tl[yb + jw] - 1;
                rec[11][slot] = tsup[B.zb2[k] + tl[B.zxb[k] + (q - B.zb[k])]] - 2;
                rec[12][slot] = c3tl[q] - 1;
                rec[13][slot] = c3tl[lbound_ll(tl, B.yzb[k], B.yze[k], pyw)] - 1;
                rec[14][slot] = c4;
                rec[15][slot] = dx;
                rec[16][slot] = dy;
                rec[17][slot] = B.dz[k];
                rec[18][slot] = dw;
                rec[19][slot] = x;
                rec[20][slot] = y;
                rec[21][slot] = B.z[k];
                rec[22][slot] = w;
              }
              nrec += __popcll(fm);
              wave_sync();
              if (nrec >= kWave) {
                clique_eval<NJ>(tabs, njobs, smod, srem, rec, kWave, acc);
                wave_sync();
                // keep the records past the first 64
                for (int r = 0; r < kRecRows; r++)
                  if (lane + kWave < nrec) rec[r][lane] = rec[r][lane + kWave];
                nrec -= kWave;
                wave_sync();
              }
            }
            wave_sync();  // the batch arrays are rewritten by the next batch
          }
          // the row bits were set by atomics in L2: the rows' lines in this CU's L1
          // (read by an earlier pair) are invalidated before sweep 1 reads them
          if (sweep == 0) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
          wave_sync();
        }
      }
      for (int j = lane; j < t; j += kWave) M[tl[yb + j]] = 0;
      wave_sync();
    }
  }
  if (nrec > 0) clique_eval<NJ>(tabs, njobs, smod, srem, rec, nrec, acc);
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
  return t.grid > 0 ? t.grid : cus * 3;
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
    // the merged table sums 24 counts, each below maxdeg^fringes / den
    {
      unsigned __int128 p = 24;
      bool ok = true;
      for (int f = job_fringes(jobs[k]); f > 0 && ok; f--) {
        if (p > (~(unsigned __int128)0) / (unsigned __int128)std::max(1, g->maxdeg)) ok = false;
        else p *= (unsigned __int128)std::max(1, g->maxdeg);
      }
      tabs[k].use_all = ok ? 1 : 0;
    }
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
