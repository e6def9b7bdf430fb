// ct_primitives.h -- wavefront / workgroup building blocks (64-lane wavefronts, 256-lane
// workgroups).  No hipMemsetAsync anywhere: buffers are cleared by kernels so that every
// entry point stays safe to capture in a HIP graph.
#pragma once

#include "ct_internal.h"

namespace ct {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Exclusive scan over a BLOCK-lane workgroup (BLOCK a multiple of 64, at most 1024).  lds
// must hold BLOCK/64 elements.  Every lane must call it (it contains barriers).  *total
// receives the workgroup sum.
template <typename T, int BLOCK = kBlock>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds, T* total) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  T before = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < BLOCK / 64; ++k) {
    const T s = lds[k];
    if (k < w) before += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return before + x - v;
}

// Segmented (by equal consecutive key) min/max reduction across a wavefront of
// order-preserving FP64 keys; the last lane of each run applies one pair of atomics.
// Runs must be contiguous in lane order; key < 0 marks an inactive lane.
__device__ __forceinline__ void wave_run_minmax(int key, unsigned long long lo,
                                                unsigned long long hi,
                                                unsigned long long* __restrict__ out_lo,
                                                unsigned long long* __restrict__ out_hi) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int k2 = __shfl_up(key, d, 64);
    const unsigned long long lo2 = __shfl_up(lo, d, 64);
    const unsigned long long hi2 = __shfl_up(hi, d, 64);
    if (lane >= d && k2 == key) {
      lo = lo2 < lo ? lo2 : lo;
      hi = hi2 > hi ? hi2 : hi;
    }
  }
  const int knext = __shfl_down(key, 1, 64);
  const bool tail = (lane == 63) || (knext != key);
  if (key >= 0 && tail) {
    atomicMin(out_lo + key, lo);
    atomicMax(out_hi + key, hi);
  }
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

template <typename T>
inline int fill(T* p, int64_t n, T value, hipStream_t st) {
  if (n <= 0) return CT_OK;
  k_fill<T><<<grid_for(n), kBlock, 0, st>>>(p, n, value);
  CT_LAUNCH_CHECK();
  return CT_OK;
}

int exclusive_scan(const int64_t* in, int64_t n, int64_t* out, void* ws, size_t ws_bytes,
                   hipStream_t st);

}  // namespace ct
