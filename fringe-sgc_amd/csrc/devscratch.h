// Stream-ordered device temporaries of one counting call (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace fsgc {

// Every temporary of a call is taken with hipMallocAsync on the call's stream and
// handed back with hipFreeAsync on that stream when the call returns (error paths
// included). Calls on different streams therefore never wait on each other through
// an allocation or a free, which a device-wide hipFree would force.
struct DevScratch {
  hipStream_t stream;
  std::vector<void*> ptrs;
  size_t bytes = 0;

  explicit DevScratch(hipStream_t s) : stream(s) {}
  DevScratch(const DevScratch&) = delete;
  DevScratch& operator=(const DevScratch&) = delete;
  ~DevScratch()
  {
    for (void* p : ptrs) (void)hipFreeAsync(p, stream);
  }

  template <class T>
  hipError_t get(T** out, size_t nbytes)
  {
    void* p = nullptr;
    nbytes = std::max<size_t>(nbytes, 256);
    const hipError_t e = hipMallocAsync(&p, nbytes, stream);
    if (e != hipSuccess) return e;
    ptrs.push_back(p);
    bytes += nbytes;
    *out = (T*)p;
    return hipSuccess;
  }
};

}  // namespace fsgc

namespace fsgc {

// A pair of timing events destroyed with the call.
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t create()
  {
    hipError_t e = hipEventCreate(&a);
    return e != hipSuccess ? e : hipEventCreate(&b);
  }
  ~EventPair()
  {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace fsgc
