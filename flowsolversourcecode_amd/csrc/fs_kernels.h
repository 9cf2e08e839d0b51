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

// Fold of the occupied ELL slots of row i (column-major nbr/anb, neighbour
// k >= 0) in slot order: acc = op(acc, anb_j, x_k) for j = 0, 1, ..., w-1.
// The slots go in batches of B: the batch's neighbour indices are loaded
// first, then its coefficients and gathered values, then the terms are
// combined in slot order.  A row thus keeps B independent loads in flight
// instead of waiting on each slot's index before fetching its value; slots
// past w re-read slot w-1 and empty slots read x_i, both discarded, so the
// result is exactly that of the plain slot loop.
// Neighbours of the ELL rows: the 0-based column-major indices (-1 = none)
// and, when every offset k - i of the mesh fits 16 bits, the same neighbours
// as int16 offsets from the row (0 = none: a cell is never its own
// neighbour) packed two slots to a 32-bit word (slot 2p in the low half,
// 2p+1 in the high half, word p of cell i at d16[p n + i]): half the index
// bytes in full-width loads.
struct EllNbr {
  const int *i32 = nullptr;
  const int *d16 = nullptr;
  EllNbr() = default;
  __host__ __device__ EllNbr(const int *p, const int *d = nullptr) : i32(p), d16(d) {}
};

// A load of matrix data (coefficients, neighbour indices) that a solve
// streams once per use: with NT it is marked non-temporal, so that it does
// not displace the gathered vectors from L2 and the Infinity Cache.
template <bool NT, class T> __device__ __forceinline__ T ld_mat(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// the neighbours of slots j0 .. j0+B-1 of row i (B even; slots past w
// repeat the last slot), -1 for none
template <int B, bool NT = false>
__device__ __forceinline__ void ell_batch(const EllNbr &nb, int i, int n, int w, int j0, int *k) {
  if (nb.d16) {
    const int wp = (w + 1) >> 1;
#pragma unroll
    for (int u = 0; u < B / 2; ++u) {
      const int v = ld_mat<NT>(&nb.d16[(size_t)min((j0 >> 1) + u, wp - 1) * n + i]);
      const int d0 = (short)(v & 0xffff), d1 = v >> 16;
      k[2 * u] = d0 ? i + d0 : -1;
      k[2 * u + 1] = d1 ? i + d1 : -1;
    }
  } else {
#pragma unroll
    for (int u = 0; u < B; ++u) k[u] = ld_mat<NT>(&nb.i32[(size_t)min(j0 + u, w - 1) * n + i]);
  }
}

template <int B, bool NT = false, class T, class Op>
__device__ __forceinline__ T ell_fold(int i, int n, int w, EllNbr nb, const T *__restrict__ anb, const T *x, T acc,
                                      Op op) {
  for (int j0 = 0; j0 < w; j0 += B) {
    int k[B];
    T a[B], xv[B];
    ell_batch<B, NT>(nb, i, n, w, j0, k);
#pragma unroll
    for (int u = 0; u < B; ++u) {
      a[u] = ld_mat<NT>(&anb[(size_t)min(j0 + u, w - 1) * n + i]);
      xv[u] = x[k[u] >= 0 ? k[u] : i];
    }
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (j0 + u < w && k[u] >= 0) acc = op(acc, a[u], xv[u]);
  }
  return acc;
}

// batch width for a row of w slots: all of a hexahedron's or tetrahedron's
// slots at once
constexpr int kEllBatch = 8;

} // namespace fs
