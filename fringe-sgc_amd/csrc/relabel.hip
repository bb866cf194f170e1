// Degree relabelling of a resident graph (gfx950).
//
// The reference renumbers the data graph in place before a triangle-core count
// (relabelByDegree, FC:1221-1288; init_sort, create_map, create_edge_list,
// create_nindex and create_nlist, FC:1065-1164): vertices ordered by degree, highest
// first, ties by the larger id first (its descending sort of (degree << idbits) + id),
// and every list rebuilt in the new ids, sorted. Counts do not depend on the labels,
// so the relabelled graph counts the same as the original; what changes is locality
// (the hubs share a few cache lines of every per-vertex array).
//
// Here the caller's graph is left as it is and the relabelled graph is a new
// resident graph. The steps are HBM-bound sorts and scatters:
//   1. key[u] = (d_u << idbits) | u, radix-sorted (descending, or ascending for the
//      opposite order) over idbits + degbits bits; position i of the sorted keys is
//      the new id of vertex key & mask;
//   2. map[old] = new, and the new degree sequence, whose exclusive scan is the new
//      index array;
//   3. one wave per old vertex writes its list, mapped, at the vertex's new offset as
//      keys (new source << idbits) | new neighbour: the keys are grouped by source, so
//      one radix sort over 2 * idbits bits sorts every list in place;
//   4. the new neighbour ids are the low bits of the sorted keys.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "devscratch.h"
#include "fsgc_plan.h"

namespace fsgc {
namespace dev {

typedef unsigned long long ull;

__global__ void k_rl_keys(const int* __restrict__ idx, int n, int idbits, ull* __restrict__ keys)
{
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    keys[u] = ((ull)(idx[u + 1] - idx[u]) << idbits) | (ull)u;
}

// map[old] = new; deg[new] = degree, deg[n] = 0 (the scan's last term)
__global__ void k_rl_map(const ull* __restrict__ keys, int n, int idbits, int* __restrict__ map,
                         int* __restrict__ deg)
{
  const ull mask = (1ull << idbits) - 1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
    if (i == n) {
      deg[n] = 0;
    } else {
      const ull k = keys[i];
      map[(int)(k & mask)] = i;
      deg[i] = (int)(k >> idbits);
    }
  }
}

// one wave per old vertex: its list, mapped, as sort keys at the vertex's new offset
__global__ void k_rl_edges(const int* __restrict__ idx, const int* __restrict__ adj, int n,
                           const int* __restrict__ map, const int* __restrict__ nidx, int idbits,
                           ull* __restrict__ ekeys)
{
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int u = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); u < n; u += waves) {
    const int beg = idx[u], end = idx[u + 1];
    const int nu = map[u];
    const ull hi = (ull)nu << idbits;
    const long long base = nidx[nu];
    for (int e = beg + lane; e < end; e += 64) ekeys[base + (e - beg)] = hi | (ull)map[adj[e]];
  }
}

__global__ void k_rl_nlist(const ull* __restrict__ ekeys, long long m, int idbits, int* __restrict__ adj)
{
  const ull mask = (1ull << idbits) - 1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < m; e += (long long)gridDim.x * blockDim.x)
    adj[e] = (int)(ekeys[e] & mask);
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

static int bits_for(unsigned long long v)  // bits to hold every value in [0, v]
{
  int b = 1;
  while (b < 63 && (v >> b) != 0) b++;
  return b;
}

using dev::ull;

int device_graph_relabel(const DeviceGraph& src, int order, DeviceGraph* dst, std::vector<int32_t>* h_map,
                         double* ms, std::string* err)
{
  const int n = src.n;
  const long long m = src.m;
  if (n < 1) {
    if (err) *err = "empty graph";
    return -22;
  }
  hipStream_t stream = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{stream};
  EventPair ev;
  HIP_TRY(ev.create());
  int dev = 0, cus = 256;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int idbits = bits_for((unsigned long long)(n - 1));
  const int degbits = bits_for((unsigned long long)std::max(0, src.maxdeg));
  const size_t nk = (size_t)std::max<long long>((long long)n, m) + 1;

  dst->n = n;
  dst->m = m;
  dst->maxdeg = src.maxdeg;
  HIP_TRY(hipMalloc((void**)&dst->idx, sizeof(int32_t) * ((size_t)n + 1)));
  HIP_TRY(hipMalloc((void**)&dst->adj, sizeof(int32_t) * std::max<size_t>(1, (size_t)m)));

  HIP_TRY(hipEventRecord(ev.a, stream));
  {
    DevScratch sc(stream);
    ull* k1 = nullptr;
    ull* k2 = nullptr;
    int* map = nullptr;
    int* deg = nullptr;
    HIP_TRY(sc.get(&k1, sizeof(ull) * nk));
    HIP_TRY(sc.get(&k2, sizeof(ull) * nk));
    HIP_TRY(sc.get(&map, sizeof(int) * (size_t)n));
    HIP_TRY(sc.get(&deg, sizeof(int) * ((size_t)n + 1)));
    const int grid = cus * 8;

    // 1. vertex keys, sorted
    hipLaunchKernelGGL(dev::k_rl_keys, dim3(grid), dim3(256), 0, stream, src.idx, n, idbits, k1);
    HIP_TRY(hipGetLastError());
    hipcub::DoubleBuffer<ull> vk(k1, k2);
    size_t tmp_bytes = 0;
    const int vbits = std::min(64, idbits + degbits);
    if (order == 0)
      HIP_TRY(hipcub::DeviceRadixSort::SortKeysDescending(nullptr, tmp_bytes, vk, n, 0, vbits, stream));
    else
      HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, vk, n, 0, vbits, stream));
    size_t tmp2 = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, deg, dst->idx, n + 1, stream));
    hipcub::DoubleBuffer<ull> ek(k1, k2);
    size_t tmp3 = 0;
    if (m > 0) HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp3, ek, (int)m, 0, 2 * idbits, stream));
    void* tmp = nullptr;
    HIP_TRY(sc.get(&tmp, std::max(tmp_bytes, std::max(tmp2, tmp3)) + 16));
    if (order == 0)
      HIP_TRY(hipcub::DeviceRadixSort::SortKeysDescending(tmp, tmp_bytes, vk, n, 0, vbits, stream));
    else
      HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, vk, n, 0, vbits, stream));

    // 2. old -> new and the new index
    hipLaunchKernelGGL(dev::k_rl_map, dim3(grid), dim3(256), 0, stream, (const ull*)vk.Current(), n, idbits, map, deg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp2, deg, dst->idx, n + 1, stream));

    if (m > 0) {
      // 3. mapped lists as keys grouped by new source, then one sort over both ids
      ull* ebuf = vk.Current() == k1 ? k2 : k1;  // the buffer not holding the vertex keys
      hipLaunchKernelGGL(dev::k_rl_edges, dim3(grid), dim3(256), 0, stream, src.idx, src.adj, n, (const int*)map,
                         (const int*)dst->idx, idbits, ebuf);
      HIP_TRY(hipGetLastError());
      hipcub::DoubleBuffer<ull> ek2(ebuf, ebuf == k1 ? k2 : k1);
      HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, tmp3, ek2, (int)m, 0, 2 * idbits, stream));
      // 4. new neighbour ids
      hipLaunchKernelGGL(dev::k_rl_nlist, dim3(grid), dim3(256), 0, stream, (const ull*)ek2.Current(), m, idbits,
                         dst->adj);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ev.b, stream));
    if (h_map) {
      h_map->resize((size_t)n);
      HIP_TRY(hipMemcpyAsync(h_map->data(), map, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, stream));
    }
    dst->h_idx.resize((size_t)n + 1);
    HIP_TRY(hipMemcpyAsync(dst->h_idx.data(), dst->idx, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost,
                           stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }  // scratch handed back on the stream
  HIP_TRY(hipStreamSynchronize(stream));
  float t = 0;
  HIP_TRY(hipEventElapsedTime(&t, ev.a, ev.b));
  if (ms) *ms = t;
  return 0;
}

}  // namespace fsgc
