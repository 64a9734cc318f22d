// Strided reductions over device arrays, p[start], p[start + stride], ...
// below n, with the loads issued kLoadBatch at a time: a plain strided loop
// waits for each load before it issues the next, so a few thousand partial
// values cost microseconds of serial latency.
#pragma once
#include <hip/hip_runtime.h>

namespace ausm {

constexpr int kLoadBatch = 8;

// NaN-propagating max; exact, so the grouping does not change the result.
// Starts from 0 (the values reduced are non-negative wave speeds).
template <typename R>
__device__ __forceinline__ R strided_max(const R* p, int n, int start, int stride) {
  R m = R(0);
  for (int base = start; base < n; base += kLoadBatch * stride) {
    R v[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int i = base + u * stride;
      v[u] = i < n ? p[i] : R(0);
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) m = (v[u] > m || v[u] != v[u]) ? v[u] : m;
  }
  return m;
}

// Sum in index order (p[start] first), the order of a plain strided loop, so
// the result is bitwise that loop's.
template <typename R>
__device__ __forceinline__ R strided_sum(const R* p, int n, int start, int stride) {
  R s = R(0);
  for (int base = start; base < n; base += kLoadBatch * stride) {
    R v[kLoadBatch];
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u) {
      const int i = base + u * stride;
      v[u] = i < n ? p[i] : R(0);
    }
#pragma unroll
    for (int u = 0; u < kLoadBatch; ++u)
      if (base + u * stride < n) s += v[u];
  }
  return s;
}

}  // namespace ausm
