// Device helpers: deterministic block reductions and XCD-aware block order.
#pragma once

#include <hip/hip_runtime.h>

namespace fs {

// Sum over a block of kBlock threads with a fixed combination order
// (wavefront shuffles, then the per-wave partials in wave order).  The
// result is valid in thread 0.
template <int BS, class T> __device__ __forceinline__ T block_sum(T v) {
  static_assert(BS % 64 == 0, "block must be whole wavefronts");
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __shared__ T s_w[BS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) s_w[wid] = v;
  __syncthreads();
  T out = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) out += s_w[i];
  }
  __syncthreads();
  return out;
}

template <int BS, class T> __device__ __forceinline__ T block_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  __shared__ T s_w[BS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) s_w[wid] = v;
  __syncthreads();
  T out = v;
  if (threadIdx.x == 0) {
    out = s_w[0];
#pragma unroll
    for (int i = 1; i < BS / 64; ++i) out = fmax(out, s_w[i]);
  }
  __syncthreads();
  return out;
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remapping the block index gives every XCD one
// contiguous range of cells, so the plane neighbours a gather touches are
// mostly resident in the same L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

} // namespace fs
