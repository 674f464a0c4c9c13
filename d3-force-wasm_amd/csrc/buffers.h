// Host-side buffers of the d3fgpu runtime: pinned host arrays (what the C
// ABI hands to its callers) and device arrays, both with a capacity that
// only grows.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "common.h"

namespace d3f {

// A pinned host array with capacity; realloc keeps the contents.
template <class T>
struct Pinned {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t n, size_t keep) {
    if (n <= cap) return hipSuccess;
    T* q = nullptr;
    HIPCHECK(hipHostMalloc((void**)&q, n * sizeof(T), hipHostMallocDefault));
    if (p) {
      if (keep) std::memcpy(q, p, std::min(keep, cap) * sizeof(T));
      HIPCHECK(hipHostFree(p));
    }
    p = q;
    cap = n;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

template <class T>
struct Device {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t n, size_t keep = 0, hipStream_t stream = nullptr) {
    if (n <= cap) return hipSuccess;
    T* q = nullptr;
    HIPCHECK(hipMalloc((void**)&q, n * sizeof(T)));
    if (p) {
      if (keep) HIPCHECK(hipMemcpyAsync(q, p, std::min(keep, cap) * sizeof(T), hipMemcpyDeviceToDevice, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      HIPCHECK(hipFree(p));
    }
    p = q;
    cap = n;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace d3f
