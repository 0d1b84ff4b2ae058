// Shared device helpers for the gfx950 dictionary-indexing kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <limits.h>

#include "../../include/ebsd_amd.h"

#define EBSD_WAVE 64

// Convert a hipError_t into the ABI's positive error range.
#define EBSD_HIP_RET(expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return (int)_e;                    \
  } while (0)

#define EBSD_LAUNCH_CHECK() EBSD_HIP_RET(hipGetLastError())

// Sum over one 64-lane wavefront.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, EBSD_WAVE);
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, EBSD_WAVE));
  return v;
}

// Block-wide reductions for a 256-thread workgroup (4 waves). `scratch` needs
// 4 floats of LDS; every thread receives the result.
__device__ __forceinline__ float block_sum_256(float v, float* scratch) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[w] = v;
  __syncthreads();
  return (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
}

// Sum / maximum over a workgroup of BLOCK threads (BLOCK / 64 slots of
// scratch); BLOCK = 256 is block_sum_256 / block_max_256. The wave partials
// are combined in a fixed pairwise order.
template <int BLOCK>
__device__ __forceinline__ float block_sum(float v, float* scratch) {
  if constexpr (BLOCK == 256) {
    return block_sum_256(v, scratch);
  } else {
    constexpr int W = BLOCK / 64;
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    float p[W];
#pragma unroll
    for (int i = 0; i < W; ++i) p[i] = scratch[i];
    // pairwise levels; an odd element is carried to the next level
#pragma unroll
    for (int len = W; len > 1; len = (len + 1) / 2) {
#pragma unroll
      for (int i = 0; i < len / 2; ++i) p[i] = p[2 * i] + p[2 * i + 1];
      if (len & 1) p[len / 2] = p[len - 1];
    }
    return p[0];
  }
}

__device__ __forceinline__ float block_max_256(float v, float* scratch);

template <int BLOCK>
__device__ __forceinline__ float block_max(float v, float* scratch) {
  if constexpr (BLOCK == 256) {
    return block_max_256(v, scratch);
  } else {
    constexpr int W = BLOCK / 64;
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    float m = scratch[0];
#pragma unroll
    for (int i = 1; i < W; ++i) m = fmaxf(m, scratch[i]);
    return m;
  }
}

__device__ __forceinline__ float block_max_256(float v, float* scratch) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[w] = v;
  __syncthreads();
  return fmaxf(fmaxf(scratch[0], scratch[1]), fmaxf(scratch[2], scratch[3]));
}

// Total order used by every top-k in this library: a higher score wins, an
// equal score is won by the lower dictionary index. The result of a top-k is
// then independent of how the dictionary is tiled or split.
__device__ __forceinline__ bool topk_better(float va, int ia, float vb, int ib) {
  return (va > vb) || (va == vb && ia < ib);
}

// Insert (v, id) into a descending register list of length K, given that it
// beats the current last entry. Fully unrolled so the list stays in VGPRs.
template <int K>
__device__ __forceinline__ void topk_insert(float (&tv)[K], int (&ti)[K], float v, int id) {
  tv[K - 1] = v;
  ti[K - 1] = id;
#pragma unroll
  for (int j = K - 1; j > 0; --j) {
    const bool b = topk_better(tv[j], ti[j], tv[j - 1], ti[j - 1]);
    const float va = b ? tv[j] : tv[j - 1];
    const float vb = b ? tv[j - 1] : tv[j];
    const int ia = b ? ti[j] : ti[j - 1];
    const int ib = b ? ti[j - 1] : ti[j];
    tv[j - 1] = va;
    tv[j] = vb;
    ti[j - 1] = ia;
    ti[j] = ib;
  }
}

template <int K>
__device__ __forceinline__ void topk_init(float (&tv)[K], int (&ti)[K]) {
#pragma unroll
  for (int j = 0; j < K; ++j) {
    tv[j] = -INFINITY;
    ti[j] = INT_MAX;
  }
}

template <int K>
__device__ __forceinline__ void topk_offer(float (&tv)[K], int (&ti)[K], float v, int id) {
  if (topk_better(v, id, tv[K - 1], ti[K - 1])) topk_insert<K>(tv, ti, v, id);
}
