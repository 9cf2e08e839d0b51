// Learned warm start for the p' solve: a message-passing network over the
// cell graph, trained and run with hand-written kernels.  The dense layers
// run on the matrix cores (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation); the message aggregation is a gather over the ELL slots,
// bound by memory like the solver's own kernels.  No counterpart in the
// reference; the definition is stated in include/fsamd_learn.h.
//
// Storage: weights FP32 master + BF16 copies in the layouts the matrix cores
// read (K contiguous per lane: forward Wt[out][in], backward Wb[in][out]);
// activations BF16 [cell][64] (128-byte rows); gradients of activations FP32.
// Every reduction runs over fixed per-block partials in a fixed order, so a
// step is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "fs_common.h"
#include "fs_kernels.h"
#include "fs_mpnn.h"
#include "fs_solve.h"

namespace fs {
namespace {

constexpr int kH = 64;     // hidden width
constexpr int kF = 8;      // input features (4 used, padded)
constexpr int kIn = 2 * kH; // input width of a message-passing layer
constexpr int kMaxLayers = 8;
constexpr int kStatBlocks = 64;
constexpr int kRowsFwd = 64;   // cells per block of the dense kernels
constexpr int kChunkW = 256;   // cells per block of the weight-gradient kernel

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

inline int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }

// float -> bfloat16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x8 load_bf8(const unsigned short *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// ---------------------------------------------------------------- features
// per (sample, block) partial sums of b^2, ac, vol, x
constexpr int kStats = 5; // per sample: s_b, s_a, mean vol, mean x, mean xp

template <class T>
__global__ void k_stats(int N, const int *__restrict__ ids, const T *__restrict__ b, const T *__restrict__ ac,
                        const T *__restrict__ vol, const T *__restrict__ x, const T *__restrict__ xp,
                        double *__restrict__ part) {
  const int s = blockIdx.y, id = ids[s];
  const size_t base = (size_t)id * N;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
    const double bi = (double)b[base + i];
    a0 += bi * bi;
    a1 += (double)ac[base + i];
    a2 += (double)vol[base + i];
    if (x) a3 += (double)x[base + i];
    if (xp) a4 += (double)xp[base + i];
  }
  a0 = block_sum<kBlock>(a0);
  a1 = block_sum<kBlock>(a1);
  a2 = block_sum<kBlock>(a2);
  a3 = block_sum<kBlock>(a3);
  a4 = block_sum<kBlock>(a4);
  if (threadIdx.x == 0) {
    double *p = part + ((size_t)s * gridDim.x + blockIdx.x) * kStats;
    p[0] = a0;
    p[1] = a1;
    p[2] = a2;
    p[3] = a3;
    p[4] = a4;
  }
}

// scal[s] = {s_b, s_a, mean vol, mean x, mean xp}
__global__ void k_stats_fin(int N, int nb, const double *__restrict__ part, double *__restrict__ scal) {
  const int s = blockIdx.x;
  if (threadIdx.x != 0) return;
  double a[kStats] = {0, 0, 0, 0, 0};
  for (int k = 0; k < nb; ++k)
    for (int q = 0; q < kStats; ++q) a[q] += part[((size_t)s * nb + k) * kStats + q];
  scal[s * kStats + 0] = sqrt(a[0] / N) + 1e-300;
  for (int q = 1; q < kStats; ++q) scal[s * kStats + q] = a[q] / N;
}

// X0, edge weights E[j][c] and the normalised target of the batch cells
template <class T>
__global__ void k_feat(int N, int W, int Bn, const int *__restrict__ ids, const int *__restrict__ nbr0,
                       const T *__restrict__ b, const T *__restrict__ ac, const T *__restrict__ anb,
                       const T *__restrict__ vol, const T *__restrict__ x, const T *__restrict__ xp,
                       const double *__restrict__ scal, float *__restrict__ X0, float *__restrict__ E,
                       float *__restrict__ Tg) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= Bn) return;
  const int s = c / N, i = c - s * N, id = ids[s];
  const size_t base = (size_t)id * N;
  const double *sc = scal + (size_t)s * kStats;
  const double sb = sc[0], sa = sc[1], vm = sc[2], xm = sc[3], xpm = sc[4];
  const double aci = (double)ac[base + i];
  int nn = 0;
  for (int j = 0; j < W; ++j) {
    const int k = nbr0[(size_t)j * N + i];
    float e = 0.f;
    if (k >= 0) {
      ++nn;
      e = (float)(-(double)anb[((size_t)id * W + j) * N + i] / aci);
    }
    E[(size_t)j * Bn + c] = e;
  }
  float *xr = X0 + (size_t)c * kF;
  xr[0] = (float)((double)b[base + i] / sb);
  xr[1] = (float)(aci / sa);
  xr[2] = (float)((double)nn / W);
  xr[3] = (float)((double)vol[base + i] / vm);
  xr[4] = xp ? (float)(((double)xp[base + i] - xpm) * sa / sb) : 0.f;
  xr[5] = xr[6] = xr[7] = 0.f;
  if (Tg) Tg[c] = x ? (float)(((double)x[base + i] - xm) * sa / sb) : 0.f;
}

// ---------------------------------------------------------------- forward
// H0 = bf16(relu(X0 We + be)); thread per (cell, 8 outputs)
__global__ void k_enc(int Bn, const float *__restrict__ X0, const float *__restrict__ weq,
                      const float *__restrict__ be, unsigned short *__restrict__ H0) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(t >> 3), g = (int)(t & 7);
  if (c >= Bn) return;
  const float *xr = X0 + (size_t)c * kF;
  u16x8 out;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int o = g * 8 + q;
    float z = 0.f;
#pragma unroll
    for (int f = 0; f < kF; ++f) z += xr[f] * weq[f * kH + o];
    z += be[o];
    out[q] = f2bf(fmaxf(z, 0.f));
  }
  *reinterpret_cast<u16x8 *>(H0 + (size_t)c * kH + g * 8) = out;
}

// M[c] = bf16(sum_j E[j][c] H[nbr_j(c)]); thread per (cell, 8 features)
__global__ void k_agg(int N, int W, int Bn, const int *__restrict__ nbr0, const float *__restrict__ E,
                      const unsigned short *__restrict__ H, unsigned short *__restrict__ M) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(t >> 3), g = (int)(t & 7);
  if (c >= Bn) return;
  const int s = c / N, i = c - s * N;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < W; ++j) {
    const int k = nbr0[(size_t)j * N + i];
    if (k < 0) continue;
    const float e = E[(size_t)j * Bn + c];
    const u16x8 h = *reinterpret_cast<const u16x8 *>(H + ((size_t)s * N + k) * kH + g * 8);
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += e * bf2f(h[q]);
  }
  u16x8 out;
#pragma unroll
  for (int q = 0; q < 8; ++q) out[q] = f2bf(acc[q]);
  *reinterpret_cast<u16x8 *>(M + (size_t)c * kH + g * 8) = out;
}

// z = [H M] W + b on the matrix cores; Hn = bf16(H + relu(z)); mask = z > 0.
// Block: 4 waves, 64 cells; wave w: cells 16w..16w+15 x all 64 outputs
// (4 tiles of 16x16), K = 128 in 4 steps of 32.
__global__ void __launch_bounds__(256)
k_dense_fwd(int Bn, const unsigned short *__restrict__ H, const unsigned short *__restrict__ M,
            const unsigned short *__restrict__ wt, const float *__restrict__ bias, unsigned short *__restrict__ Hn,
            unsigned char *__restrict__ mask) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * kRowsFwd + wave * 16;
  const int row = r0 + (lane & 15), kg = lane >> 4;
  floatx4 acc[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    bf16x8 a;
    if (row < Bn) {
      const unsigned short *src = ks < 2 ? H + (size_t)row * kH + ks * 32 + kg * 8
                                         : M + (size_t)row * kH + (ks - 2) * 32 + kg * 8;
      a = load_bf8(src);
    } else {
      a = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int col = nt * 16 + (lane & 15);
      const bf16x8 bv = load_bf8(wt + (size_t)col * kIn + ks * 32 + kg * 8);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int col = nt * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = r0 + kg * 4 + r;
      if (rr >= Bn) continue;
      const float z = acc[nt][r] + bias[col];
      const size_t at = (size_t)rr * kH + col;
      Hn[at] = f2bf(bf2f(H[at]) + fmaxf(z, 0.f));
      mask[at] = z > 0.f;
    }
  }
}

// y = H wd + bd per cell, partial sums of y per (sample, block)
__global__ void k_decode(int N, int Bn, const unsigned short *__restrict__ H, const float *__restrict__ wdq,
                         float bd, float *__restrict__ Y, double *__restrict__ part) {
  const int s = blockIdx.y;
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
    const int c = s * N + i;
    const unsigned short *h = H + (size_t)c * kH;
    float y = 0.f;
#pragma unroll 8
    for (int o = 0; o < kH; ++o) y += bf2f(h[o]) * wdq[o];
    y += bd;
    Y[c] = y;
    acc += y;
  }
  acc = block_sum<kBlock>(acc);
  if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = acc;
  (void)Bn;
}

// ymean[s] from the partials
__global__ void k_mean_fin(int N, int nb, const double *__restrict__ part, double *__restrict__ ymean) {
  if (threadIdx.x != 0) return;
  const int s = blockIdx.x;
  double a = 0.0;
  for (int k = 0; k < nb; ++k) a += part[(size_t)s * nb + k];
  ymean[s] = a / N;
}

// yc = y - mean; loss partials and dy = 2 (yc - t)/total (training), or the
// prediction x = yc s_b/s_a (inference, out != null)
__global__ void k_loss(int N, const double *__restrict__ ymean, const double *__restrict__ scal,
                       float *__restrict__ Y, const float *__restrict__ Tg, float *__restrict__ DY, double inv_total,
                       double *__restrict__ part, double *__restrict__ out) {
  const int s = blockIdx.y;
  const double ym = ymean[s];
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
    const int c = s * N + i;
    const float yc = (float)((double)Y[c] - ym);
    Y[c] = yc;
    if (out) {
      out[(size_t)s * N + i] = (double)yc * scal[s * kStats] / scal[s * kStats + 1];
    } else {
      const double d = (double)yc - (double)Tg[c];
      acc += d * d;
      DY[c] = (float)(2.0 * d * inv_total);
    }
  }
  if (!out) {
    acc = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = acc;
  }
}

// ---------------------------------------------------------------- backward
// dH[c][o] = dy[c] wd[o]
__global__ void k_dec_bwd(int Bn, const float *__restrict__ DY, const float *__restrict__ wdq, float *__restrict__ dH) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)Bn * kH) return;
  const int c = (int)(t >> 6), o = (int)(t & 63);
  dH[t] = DY[c] * wdq[o];
}

// Per-block column sums: out[blk][o] = sum over the block's cells of
// a[c] * V[c][o] (a = null: 1), V as BF16 (Vb) or FP32 (Vf); out[blk][64] =
// sum of a[c] when a is given.  Block: 256 threads = 4 cell lanes x 64 columns.
__global__ void k_colsum(int Bn, int chunk, const float *__restrict__ a, const unsigned short *__restrict__ Vb,
                         const float *__restrict__ Vf, float *__restrict__ out) {
  const int o = threadIdx.x & 63, r = threadIdx.x >> 6;
  const int c0 = blockIdx.x * chunk, c1 = min(Bn, c0 + chunk);
  float s = 0.f, sa = 0.f;
  for (int c = c0 + r; c < c1; c += 4) {
    const float v = Vb ? bf2f(Vb[(size_t)c * kH + o]) : Vf[(size_t)c * kH + o];
    const float w = a ? a[c] : 1.f;
    s += w * v;
    if (o == 0 && a) sa += w;
  }
  __shared__ float red[4][kH + 1];
  red[r][o] = s;
  if (o == 0) red[r][kH] = sa;
  __syncthreads();
  if (r == 0) {
    out[(size_t)blockIdx.x * (kH + 1) + o] = red[0][o] + red[1][o] + red[2][o] + red[3][o];
    if (o == 0) out[(size_t)blockIdx.x * (kH + 1) + kH] = red[0][kH] + red[1][kH] + red[2][kH] + red[3][kH];
  }
}

// dz = bf16(mask ? dH : 0)
__global__ void k_relu_bwd(int Bn, const float *__restrict__ dH, const unsigned char *__restrict__ mask,
                           unsigned short *__restrict__ DZ) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long long)Bn * kH) return;
  DZ[t] = f2bf(mask[t] ? dH[t] : 0.f);
}

// D = dz W^T on the matrix cores: [Bn x 64] x [64 x 128] -> FP32 [Bn x 128].
// Wave: 16 cells x 128 columns (8 tiles), K = 64 in 2 steps.
__global__ void __launch_bounds__(256)
k_dense_bwd_x(int Bn, const unsigned short *__restrict__ DZ, const unsigned short *__restrict__ wb,
              float *__restrict__ D) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * kRowsFwd + wave * 16;
  const int row = r0 + (lane & 15), kg = lane >> 4;
  floatx4 acc[8];
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const bf16x8 a = row < Bn ? load_bf8(DZ + (size_t)row * kH + ks * 32 + kg * 8)
                              : __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
      const int col = nt * 16 + (lane & 15); // input feature
      const bf16x8 bv = load_bf8(wb + (size_t)col * kH + ks * 32 + kg * 8);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) {
    const int col = nt * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = r0 + kg * 4 + r;
      if (rr < Bn) D[(size_t)rr * kIn + col] = acc[nt][r];
    }
  }
}

// dH_prev[c] = dH[c] + D[c][0:64] + sum over neighbours k of c of
// e_{k,c} D[k][64:128] (the transpose of the aggregation; e_{k,c} is the
// weight of c in k's row, found through the mirror slot)
__global__ void k_agg_t(int N, int W, int Bn, const int *__restrict__ nbr0, const int *__restrict__ mir,
                        const float *__restrict__ E, const float *__restrict__ dH, const float *__restrict__ D,
                        float *__restrict__ dHp) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int c = (int)(t >> 3), g = (int)(t & 7);
  if (c >= Bn) return;
  const int s = c / N, i = c - s * N;
  float acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = dH[(size_t)c * kH + g * 8 + q] + D[(size_t)c * kIn + g * 8 + q];
  for (int j = 0; j < W; ++j) {
    const int k = nbr0[(size_t)j * N + i];
    if (k < 0) continue;
    const int jm = mir[(size_t)j * N + i];
    const size_t ck = (size_t)s * N + k;
    const float e = E[(size_t)jm * Bn + ck];
    const float *dk = D + ck * kIn + kH + g * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += e * dk[q];
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) dHp[(size_t)c * kH + g * 8 + q] = acc[q];
}

// Partial dW = [H M]^T dz over the block's cells on the matrix cores.  The
// block stages 32 cells at a time transposed in LDS (feature-major, K =
// cells contiguous); wave w owns the 2 x 4 output tiles of input features
// 32w..32w+31.  out[blk][in][o], FP32.
constexpr int kStageCells = 32;
constexpr int kStageStride = kStageCells + 8; // bf16 elements per staged row
__global__ void __launch_bounds__(256)
k_dense_bwd_w(int Bn, const unsigned short *__restrict__ H, const unsigned short *__restrict__ M,
              const unsigned short *__restrict__ DZ, float *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) unsigned short xs[kIn * kStageStride];
  __shared__ __attribute__((aligned(16))) unsigned short zs[kH * kStageStride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, kg = lane >> 4;
  const int c0 = blockIdx.x * kChunkW;
  floatx4 acc[2][4];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int cs = c0; cs < c0 + kChunkW; cs += kStageCells) {
    __syncthreads();
    // stage: xs[f][cell] for f < 128, zs[o][cell]; zeros past Bn
    for (int e = threadIdx.x; e < kStageCells * kIn; e += blockDim.x) {
      const int cl = e / kIn, f = e - cl * kIn, c = cs + cl;
      unsigned short v = 0;
      if (c < Bn) v = f < kH ? H[(size_t)c * kH + f] : M[(size_t)c * kH + f - kH];
      xs[f * kStageStride + cl] = v;
    }
    for (int e = threadIdx.x; e < kStageCells * kH; e += blockDim.x) {
      const int cl = e / kH, o = e - cl * kH, c = cs + cl;
      zs[o * kStageStride + cl] = c < Bn ? DZ[(size_t)c * kH + o] : (unsigned short)0;
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int f = wave * 32 + mt * 16 + (lane & 15);
      const bf16x8 a = load_bf8(xs + f * kStageStride + kg * 8);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const int o = nt * 16 + (lane & 15);
        const bf16x8 bv = load_bf8(zs + o * kStageStride + kg * 8);
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[mt][nt], 0, 0, 0);
      }
    }
  }
  float *dst = out + (size_t)blockIdx.x * kIn * kH;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int o = nt * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = wave * 32 + mt * 16 + kg * 4 + r;
        dst[f * kH + o] = acc[mt][nt][r];
      }
    }
}

// encoder gradient partials: out[blk][f][o] = sum X0[c][f] dz0[c][o] (f < 8),
// out[blk][8][o] = sum dz0[c][o], dz0 = (H0 > 0) ? dH : 0
__global__ void k_enc_bwd(int Bn, int chunk, const float *__restrict__ X0, const unsigned short *__restrict__ H0,
                          const float *__restrict__ dH, float *__restrict__ out) {
  const int o = threadIdx.x & 63, r = threadIdx.x >> 6;
  const int c0 = blockIdx.x * chunk, c1 = min(Bn, c0 + chunk);
  float s[kF + 1];
#pragma unroll
  for (int f = 0; f <= kF; ++f) s[f] = 0.f;
  for (int c = c0 + r; c < c1; c += 4) {
    const float dz = bf2f(H0[(size_t)c * kH + o]) > 0.f ? dH[(size_t)c * kH + o] : 0.f;
#pragma unroll
    for (int f = 0; f < kF; ++f) s[f] += X0[(size_t)c * kF + f] * dz;
    s[kF] += dz;
  }
  __shared__ float red[4][kF + 1][kH];
#pragma unroll
  for (int f = 0; f <= kF; ++f) red[r][f][o] = s[f];
  __syncthreads();
  if (r == 0)
    for (int f = 0; f <= kF; ++f)
      out[((size_t)blockIdx.x * (kF + 1) + f) * kH + o] = red[0][f][o] + red[1][f][o] + red[2][f][o] + red[3][f][o];
}

// dst[k] = sum over blocks b (in order) of part[b][k], k < K
__global__ void k_reduce(int nb, int K, const float *__restrict__ part, float *__restrict__ dst) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= K) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * K + k];
  dst[k] = s;
}

// the bias and wd gradients from k_colsum partials (nb blocks of 65)
__global__ void k_reduce_cols(int nb, const float *__restrict__ part, float *__restrict__ dcol, float *__restrict__ dscalar) {
  const int o = threadIdx.x;
  if (o > kH) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * (kH + 1) + o];
  if (o < kH) dcol[o] = s;
  else if (dscalar) *dscalar = s;
}

__global__ void k_adam(int n, float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float c1, float c2) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// BF16 copies of the weights in the layouts the kernels read
__global__ void k_quant(int L, const float *__restrict__ th, float *__restrict__ weq, unsigned short *__restrict__ wt,
                        unsigned short *__restrict__ wb, float *__restrict__ wdq) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int per = kIn * kH;
  if (i < kF * kH) weq[i] = bf2f(f2bf(th[i]));
  if (i < kH) wdq[i] = bf2f(f2bf(th[kF * kH + kH + L * (per + kH) + i]));
  if (i >= L * per) return;
  const int l = i / per, r = i - l * per, k = r / kH, o = r - k * kH;
  const unsigned short q = f2bf(th[kF * kH + kH + l * (per + kH) + r]);
  wt[(size_t)l * per + o * kIn + k] = q;
  wb[(size_t)l * per + k * kH + o] = q;
}

__global__ void k_to_f32(long long n, const double *__restrict__ a, float *__restrict__ b) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

} // namespace
} // namespace fs

// ====================================================================== host
struct fs_mpnn {
  int N = 0, W = 0, L = 3, P = 0, cap = 0, count = 0, steps = 0;
  bool use_prev = false;  // fifth input: the previous p' solution
  hipStream_t s = nullptr;
  std::vector<int> h_nbr0;
  fs::DevBuf<int> nbr0, mir, ids;
  fs::DevBuf<float> theta, grad, am, av, weq, wdq;
  fs::DevBuf<unsigned short> wt, wb;
  fs::DevBuf<float> ds_b, ds_x, ds_ac, ds_anb, ds_vol, ds_xp;
  // batch work buffers (grown on demand)
  long long bn_cap = 0;
  fs::DevBuf<double> part, scal, ymean;
  fs::DevBuf<float> X0, E, Tg, Y, DY, dH, dH2, D, gpart;
  std::vector<fs::DevBuf<unsigned short>> Hs, Ms;
  std::vector<fs::DevBuf<unsigned char>> masks;
  fs::DevBuf<unsigned short> DZ;
  std::vector<double> h_part;
};

namespace fs {

int mpnn_offset_W(int l) { return kF * kH + kH + l * (kIn * kH + kH); }

static void mpnn_grow(fs_mpnn *n, int B) {
  const long long Bn = (long long)B * n->N;
  n->ids.alloc(B);
  n->part.alloc((size_t)B * kStatBlocks * kStats);
  n->scal.alloc((size_t)B * kStats);
  n->ymean.alloc(B);
  if (Bn <= n->bn_cap) return;
  n->X0.alloc(Bn * kF);
  n->E.alloc(Bn * n->W);
  n->Tg.alloc(Bn);
  n->Y.alloc(Bn);
  n->DY.alloc(Bn);
  n->dH.alloc(Bn * kH);
  n->dH2.alloc(Bn * kH);
  n->D.alloc(Bn * kIn);
  n->DZ.alloc(Bn * kH);
  n->Hs.resize(n->L + 1);
  n->Ms.resize(n->L);
  n->masks.resize(n->L);
  for (auto &h : n->Hs) h.alloc(Bn * kH);
  for (auto &m : n->Ms) m.alloc(Bn * kH);
  for (auto &m : n->masks) m.alloc(Bn * kH);
  const size_t nbw = cdiv(Bn, kChunkW);
  n->gpart.alloc(std::max(nbw * kIn * kH, nbw * (kF + 1) * kH));
  n->bn_cap = Bn;
}

static void quantize(fs_mpnn *n) {
  hipLaunchKernelGGL(k_quant, dim3(cdiv((long long)n->L * kIn * kH, kBlock)), dim3(kBlock), 0, n->s, n->L,
                     n->theta.p, n->weq.p, n->wt.p, n->wb.p, n->wdq.p);
}

// Forward over the batch already described by ids / features.  Returns the
// number of stat blocks per sample.
template <class T>
static void mpnn_forward(fs_mpnn *n, int B, const T *b, const T *ac, const T *anb, const T *vol, const T *x,
                         const T *xp, float bd) {
  hipStream_t s = n->s;
  const int N = n->N, W = n->W;
  const int Bn = B * N;
  const int nb = std::min(kStatBlocks, cdiv(N, kBlock));
  if (!n->use_prev) xp = nullptr;
  hipLaunchKernelGGL(k_stats<T>, dim3(nb, B), dim3(kBlock), 0, s, N, n->ids.p, b, ac, vol, x, xp, n->part.p);
  hipLaunchKernelGGL(k_stats_fin, dim3(B), dim3(64), 0, s, N, nb, n->part.p, n->scal.p);
  hipLaunchKernelGGL(k_feat<T>, dim3(cdiv(Bn, kBlock)), dim3(kBlock), 0, s, N, W, Bn, n->ids.p, n->nbr0.p, b, ac,
                     anb, vol, x, xp, n->scal.p, n->X0.p, n->E.p, x ? n->Tg.p : nullptr);
  hipLaunchKernelGGL(k_enc, dim3(cdiv((long long)Bn * 8, kBlock)), dim3(kBlock), 0, s, Bn, n->X0.p, n->weq.p,
                     n->theta.p + kF * kH, n->Hs[0].p);
  for (int l = 0; l < n->L; ++l) {
    hipLaunchKernelGGL(k_agg, dim3(cdiv((long long)Bn * 8, kBlock)), dim3(kBlock), 0, s, N, W, Bn, n->nbr0.p, n->E.p,
                       n->Hs[l].p, n->Ms[l].p);
    hipLaunchKernelGGL(k_dense_fwd, dim3(cdiv(Bn, kRowsFwd)), dim3(256), 0, s, Bn, n->Hs[l].p, n->Ms[l].p,
                       n->wt.p + (size_t)l * kIn * kH, n->theta.p + mpnn_offset_W(l) + kIn * kH, n->Hs[l + 1].p,
                       n->masks[l].p);
  }
  hipLaunchKernelGGL(k_decode, dim3(nb, B), dim3(kBlock), 0, s, N, Bn, n->Hs[n->L].p, n->wdq.p, bd, n->Y.p,
                     n->part.p);
  hipLaunchKernelGGL(k_mean_fin, dim3(B), dim3(64), 0, s, N, nb, n->part.p, n->ymean.p);
}

static float host_bd(fs_mpnn *n) {
  float bd = 0.f;
  FS_HIP_THROW(hipMemcpyAsync(&bd, n->theta.p + n->P - 1, sizeof(float), hipMemcpyDeviceToHost, n->s));
  FS_HIP_THROW(hipStreamSynchronize(n->s));
  return bd;
}

// loss (and gradient into n->grad when want_grad)
static double mpnn_loss_grad(fs_mpnn *n, const int *ids, int B, bool want_grad) {
  hipStream_t s = n->s;
  for (int k = 0; k < B; ++k)
    if (ids[k] < 0 || ids[k] >= n->count) throw std::runtime_error("sample id out of range");
  mpnn_grow(n, B);
  FS_HIP_THROW(hipMemcpyAsync(n->ids.p, ids, sizeof(int) * B, hipMemcpyHostToDevice, s));
  const int N = n->N, Bn = B * N;
  const float bd = host_bd(n);
  mpnn_forward<float>(n, B, n->ds_b.p, n->ds_ac.p, n->ds_anb.p, n->ds_vol.p, n->ds_x.p, n->ds_xp.p, bd);
  const int nb = std::min(kStatBlocks, cdiv(N, kBlock));
  hipLaunchKernelGGL(k_loss, dim3(nb, B), dim3(kBlock), 0, s, N, n->ymean.p, n->scal.p, n->Y.p, n->Tg.p, n->DY.p,
                     1.0 / Bn, n->part.p, (double *)nullptr);
  n->h_part.resize((size_t)B * nb);
  FS_HIP_THROW(hipMemcpyAsync(n->h_part.data(), n->part.p, sizeof(double) * B * nb, hipMemcpyDeviceToHost, s));
  if (want_grad) {
    const int L = n->L;
    float *g = n->grad.p;
    // decoder
    const int nbc = cdiv(Bn, kChunkW);
    hipLaunchKernelGGL(k_colsum, dim3(nbc), dim3(256), 0, s, Bn, kChunkW, n->DY.p, n->Hs[L].p, (const float *)nullptr,
                       n->gpart.p);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(128), 0, s, nbc, n->gpart.p, g + mpnn_offset_W(L),
                       g + n->P - 1);
    hipLaunchKernelGGL(k_dec_bwd, dim3(cdiv((long long)Bn * kH, kBlock)), dim3(kBlock), 0, s, Bn, n->DY.p, n->wdq.p,
                       n->dH.p);
    float *dH = n->dH.p, *dH2 = n->dH2.p;
    for (int l = L - 1; l >= 0; --l) {
      hipLaunchKernelGGL(k_relu_bwd, dim3(cdiv((long long)Bn * kH, kBlock)), dim3(kBlock), 0, s, Bn, dH,
                         n->masks[l].p, n->DZ.p);
      hipLaunchKernelGGL(k_colsum, dim3(nbc), dim3(256), 0, s, Bn, kChunkW, (const float *)nullptr, n->DZ.p,
                         (const float *)nullptr, n->gpart.p);
      hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(128), 0, s, nbc, n->gpart.p,
                         g + mpnn_offset_W(l) + kIn * kH, (float *)nullptr);
      hipLaunchKernelGGL(k_dense_bwd_w, dim3(nbc), dim3(256), 0, s, Bn, n->Hs[l].p, n->Ms[l].p, n->DZ.p,
                         n->gpart.p);
      hipLaunchKernelGGL(k_reduce, dim3(cdiv(kIn * kH, kBlock)), dim3(kBlock), 0, s, nbc, kIn * kH, n->gpart.p,
                         g + mpnn_offset_W(l));
      hipLaunchKernelGGL(k_dense_bwd_x, dim3(cdiv(Bn, kRowsFwd)), dim3(256), 0, s, Bn, n->DZ.p,
                         n->wb.p + (size_t)l * kIn * kH, n->D.p);
      hipLaunchKernelGGL(k_agg_t, dim3(cdiv((long long)Bn * 8, kBlock)), dim3(kBlock), 0, s, N, n->W, Bn, n->nbr0.p,
                         n->mir.p, n->E.p, dH, n->D.p, dH2);
      std::swap(dH, dH2);
    }
    hipLaunchKernelGGL(k_enc_bwd, dim3(nbc), dim3(256), 0, s, Bn, kChunkW, n->X0.p, n->Hs[0].p, dH, n->gpart.p);
    hipLaunchKernelGGL(k_reduce, dim3(cdiv((kF + 1) * kH, kBlock)), dim3(kBlock), 0, s, nbc, (kF + 1) * kH,
                       n->gpart.p, g);
  }
  FS_HIP_THROW(hipStreamSynchronize(s));
  double loss = 0.0;
  for (double v : n->h_part) loss += v;
  return loss / Bn;
}

} // namespace fs

using fs::set_error;

#define MPNN_GUARD_BEGIN try {
#define MPNN_GUARD_END                                                        \
  }                                                                           \
  catch (const std::exception &e) {                                           \
    set_error(e.what());                                                      \
    return -2;                                                                \
  }

extern "C" {

int fs_mpnn_create(int n_cells, int width, const int *elem_nbr, int layers, int use_prev, unsigned long long seed,
                   int capacity, fs_mpnn **out) {
  MPNN_GUARD_BEGIN
  if (!out || !elem_nbr || n_cells <= 0 || width <= 0 || layers < 1 || layers > fs::kMaxLayers || capacity < 0 ||
      (use_prev != 0 && use_prev != 1)) {
    set_error("fs_mpnn_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  for (long long t = 0; t < (long long)n_cells * width; ++t)
    if (elem_nbr[t] < 0 || elem_nbr[t] > n_cells) {
      set_error("fs_mpnn_create: ElementNeighbor entry out of range at linear index " + std::to_string(t));
      return -1;
    }
  auto n = std::make_unique<fs_mpnn>();
  n->N = n_cells;
  n->W = width;
  n->L = layers;
  n->use_prev = use_prev != 0;
  n->P = fs::mpnn_offset_W(layers) + fs::kH + 1;
  FS_HIP(hipStreamCreate(&n->s));
  // 0-based neighbours and the mirror slots (slot of i in the list of nbr_j(i))
  n->h_nbr0.resize((size_t)n_cells * width);
  for (size_t t = 0; t < n->h_nbr0.size(); ++t) n->h_nbr0[t] = elem_nbr[t] - 1;
  std::vector<int> mir(n->h_nbr0.size(), -1);
  for (int i = 0; i < n_cells; ++i)
    for (int j = 0; j < width; ++j) {
      const int k = n->h_nbr0[(size_t)j * n_cells + i];
      if (k < 0) continue;
      for (int q = 0; q < width; ++q)
        if (n->h_nbr0[(size_t)q * n_cells + k] == i) {
          mir[(size_t)j * n_cells + i] = q;
          break;
        }
      if (mir[(size_t)j * n_cells + i] < 0) {
        set_error("fs_mpnn_create: neighbour lists are not symmetric (cell " + std::to_string(i + 1) + ")");
        (void)hipStreamDestroy(n->s);
        return -1;
      }
    }
  n->nbr0.upload(n->h_nbr0, n->s);
  n->mir.upload(mir, n->s);
  // He initialisation
  std::vector<float> th(n->P, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int i = 0; i < fs::kF * fs::kH; ++i) th[i] = (float)(nd(rng) * std::sqrt(2.0 / fs::kF));
  for (int l = 0; l < layers; ++l)
    for (int i = 0; i < fs::kIn * fs::kH; ++i) th[fs::mpnn_offset_W(l) + i] = (float)(nd(rng) * std::sqrt(2.0 / fs::kIn));
  for (int i = 0; i < fs::kH; ++i) th[fs::mpnn_offset_W(layers) + i] = (float)(nd(rng) * std::sqrt(1.0 / fs::kH));
  n->theta.upload(th, n->s);
  n->grad.alloc(n->P);
  n->am.alloc(n->P);
  n->av.alloc(n->P);
  FS_HIP(hipMemsetAsync(n->grad.p, 0, sizeof(float) * n->P, n->s));
  FS_HIP(hipMemsetAsync(n->am.p, 0, sizeof(float) * n->P, n->s));
  FS_HIP(hipMemsetAsync(n->av.p, 0, sizeof(float) * n->P, n->s));
  n->weq.alloc(fs::kF * fs::kH);
  n->wdq.alloc(fs::kH);
  n->wt.alloc((size_t)layers * fs::kIn * fs::kH);
  n->wb.alloc((size_t)layers * fs::kIn * fs::kH);
  fs::quantize(n.get());
  n->cap = capacity;
  const size_t cn = (size_t)std::max(capacity, 1) * n_cells;
  n->ds_b.alloc(cn);
  n->ds_x.alloc(cn);
  n->ds_ac.alloc(cn);
  n->ds_vol.alloc(cn);
  n->ds_xp.alloc(cn);
  n->ds_anb.alloc(cn * width);
  FS_HIP(hipStreamSynchronize(n->s));
  *out = n.release();
  return 0;
  MPNN_GUARD_END
}

int fs_mpnn_destroy(fs_mpnn *n) {
  if (!n) return 0;
  (void)hipStreamSynchronize(n->s);
  hipStream_t s = n->s;
  delete n;
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int fs_mpnn_param_count(fs_mpnn *n) { return n ? n->P : -1; }

int fs_mpnn_get_params(fs_mpnn *n, double *theta, int count) {
  MPNN_GUARD_BEGIN
  if (!n || !theta || count != n->P) {
    set_error("fs_mpnn_get_params: NULL argument or count != fs_mpnn_param_count");
    return -1;
  }
  std::vector<float> h(n->P);
  FS_HIP(hipMemcpyAsync(h.data(), n->theta.p, sizeof(float) * n->P, hipMemcpyDeviceToHost, n->s));
  FS_HIP(hipStreamSynchronize(n->s));
  for (int i = 0; i < n->P; ++i) theta[i] = h[i];
  return 0;
  MPNN_GUARD_END
}

int fs_mpnn_set_params(fs_mpnn *n, const double *theta, int count) {
  MPNN_GUARD_BEGIN
  if (!n || !theta || count != n->P) {
    set_error("fs_mpnn_set_params: NULL argument or count != fs_mpnn_param_count");
    return -1;
  }
  std::vector<float> h(n->P);
  for (int i = 0; i < n->P; ++i) h[i] = (float)theta[i];
  FS_HIP(hipMemcpyAsync(n->theta.p, h.data(), sizeof(float) * n->P, hipMemcpyHostToDevice, n->s));
  FS_HIP(hipMemsetAsync(n->am.p, 0, sizeof(float) * n->P, n->s));
  FS_HIP(hipMemsetAsync(n->av.p, 0, sizeof(float) * n->P, n->s));
  n->steps = 0;
  fs::quantize(n);
  FS_HIP(hipStreamSynchronize(n->s));
  return 0;
  MPNN_GUARD_END
}

int fs_mpnn_add_samples(fs_mpnn *n, int ns, const double *b, const double *x, const double *ac, const double *anb,
                        const double *vol, const double *x_prev) {
  MPNN_GUARD_BEGIN
  if (!n || ns < 0 || !b || !x || !ac || !anb || !vol) {
    set_error("fs_mpnn_add_samples: invalid arguments");
    return -1;
  }
  if (n->count + ns > n->cap) {
    set_error("fs_mpnn_add_samples: dataset capacity exceeded");
    return -1;
  }
  const size_t N = n->N, off = (size_t)n->count * N, cnt = (size_t)ns * N;
  fs::DevBuf<double> tmp;
  tmp.alloc(cnt * n->W);
  auto put = [&](const double *src, float *dst, size_t m) {
    FS_HIP_THROW(hipMemcpyAsync(tmp.p, src, sizeof(double) * m, hipMemcpyHostToDevice, n->s));
    hipLaunchKernelGGL(fs::k_to_f32, dim3(fs::cdiv((long long)m, fs::kBlock)), dim3(fs::kBlock), 0, n->s, (long long)m,
                       tmp.p, dst);
    FS_HIP_THROW(hipStreamSynchronize(n->s));
  };
  put(b, n->ds_b.p + off, cnt);
  put(x, n->ds_x.p + off, cnt);
  put(ac, n->ds_ac.p + off, cnt);
  put(vol, n->ds_vol.p + off, cnt);
  put(anb, n->ds_anb.p + off * n->W, cnt * n->W);
  if (x_prev) put(x_prev, n->ds_xp.p + off, cnt);
  else FS_HIP(hipMemsetAsync(n->ds_xp.p + off, 0, sizeof(float) * cnt, n->s));
  FS_HIP(hipStreamSynchronize(n->s));
  n->count += ns;
  return n->count;
  MPNN_GUARD_END
}

int fs_mpnn_sample_count(fs_mpnn *n) { return n ? n->count : -1; }

int fs_mpnn_get_sample(fs_mpnn *n, int k, double *b, double *x, double *ac, double *anb, double *vol,
                       double *x_prev) {
  MPNN_GUARD_BEGIN
  if (!n || k < 0 || k >= n->count) {
    set_error("fs_mpnn_get_sample: sample out of range");
    return -1;
  }
  const size_t N = n->N;
  std::vector<float> h(N * n->W);
  auto get = [&](const float *src, double *dst, size_t m) {
    if (!dst) return;
    FS_HIP_THROW(hipMemcpyAsync(h.data(), src, sizeof(float) * m, hipMemcpyDeviceToHost, n->s));
    FS_HIP_THROW(hipStreamSynchronize(n->s));
    for (size_t i = 0; i < m; ++i) dst[i] = h[i];
  };
  get(n->ds_b.p + k * N, b, N);
  get(n->ds_x.p + k * N, x, N);
  get(n->ds_ac.p + k * N, ac, N);
  get(n->ds_vol.p + k * N, vol, N);
  get(n->ds_anb.p + k * N * n->W, anb, N * n->W);
  get(n->ds_xp.p + k * N, x_prev, N);
  return 0;
  MPNN_GUARD_END
}

int fs_mpnn_loss_grad(fs_mpnn *n, const int *ids, int batch, double *loss, double *grad) {
  MPNN_GUARD_BEGIN
  if (!n || !ids || batch < 1 || !loss) {
    set_error("fs_mpnn_loss_grad: invalid arguments");
    return -1;
  }
  *loss = fs::mpnn_loss_grad(n, ids, batch, grad != nullptr);
  if (grad) {
    std::vector<float> h(n->P);
    FS_HIP(hipMemcpyAsync(h.data(), n->grad.p, sizeof(float) * n->P, hipMemcpyDeviceToHost, n->s));
    FS_HIP(hipStreamSynchronize(n->s));
    for (int i = 0; i < n->P; ++i) grad[i] = h[i];
  }
  return 0;
  MPNN_GUARD_END
}

int fs_mpnn_train_step(fs_mpnn *n, const int *ids, int batch, double lr, double *loss) {
  MPNN_GUARD_BEGIN
  if (!n || !ids || batch < 1 || !(lr > 0.0)) {
    set_error("fs_mpnn_train_step: invalid arguments");
    return -1;
  }
  const double l = fs::mpnn_loss_grad(n, ids, batch, true);
  if (loss) *loss = l;
  n->steps++;
  const float c1 = 1.f - std::pow(0.9f, (float)n->steps), c2 = 1.f - std::pow(0.999f, (float)n->steps);
  hipLaunchKernelGGL(fs::k_adam, dim3(fs::cdiv(n->P, fs::kBlock)), dim3(fs::kBlock), 0, n->s, n->P, n->theta.p,
                     n->grad.p, n->am.p, n->av.p, (float)lr, c1, c2);
  fs::quantize(n);
  FS_HIP(hipStreamSynchronize(n->s));
  return 0;
  MPNN_GUARD_END
}

int fs_mpnn_predict(fs_mpnn *n, const double *b, const double *ac, const double *anb, const double *vol,
                    const double *x_prev, double *x) {
  MPNN_GUARD_BEGIN
  if (!n || !b || !ac || !anb || !vol || !x) {
    set_error("fs_mpnn_predict: NULL argument");
    return -1;
  }
  const size_t N = n->N;
  fs::DevBuf<double> db, dac, danb, dvol, dx, dxp;
  db.upload(b, N, n->s);
  dac.upload(ac, N, n->s);
  danb.upload(anb, N * n->W, n->s);
  dvol.upload(vol, N, n->s);
  if (x_prev) dxp.upload(x_prev, N, n->s);
  dx.alloc(N);
  const int rc = fs::mpnn_predict_device(n, n->s, db.p, dac.p, danb.p, dvol.p, x_prev ? dxp.p : nullptr, dx.p);
  if (rc < 0) return rc;
  FS_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * N, hipMemcpyDeviceToHost, n->s));
  FS_HIP(hipStreamSynchronize(n->s));
  return 0;
  MPNN_GUARD_END
}

} // extern "C"

namespace fs {

int mpnn_predict_device(fs_mpnn *n, hipStream_t stream, const double *b, const double *ac, const double *anb,
                        const double *vol, const double *xp, double *x) {
  // run on the caller's stream
  hipStream_t own = n->s;
  n->s = stream;
  try {
    mpnn_grow(n, 1);
    const int zero = 0;
    FS_HIP_THROW(hipMemcpyAsync(n->ids.p, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
    const float bd = host_bd(n);
    mpnn_forward<double>(n, 1, b, ac, anb, vol, (const double *)nullptr, xp, bd);
    const int nb = std::min(kStatBlocks, cdiv(n->N, kBlock));
    hipLaunchKernelGGL(k_loss, dim3(nb, 1), dim3(kBlock), 0, stream, n->N, n->ymean.p, n->scal.p, n->Y.p,
                       (const float *)nullptr, (float *)nullptr, 0.0, n->part.p, x);
  } catch (...) {
    n->s = own;
    throw;
  }
  n->s = own;
  return 0;
}

bool mpnn_matches(fs_mpnn *n, int n_cells, int width, const std::vector<int> &nbr0) {
  return n && n->N == n_cells && n->W == width && n->h_nbr0 == nbr0;
}

// append one device system (FP64) to the dataset; false when full
bool mpnn_capture(fs_mpnn *n, hipStream_t stream, const double *b, const double *x, const double *ac,
                  const double *anb, const double *vol, const double *xp) {
  if (!n || n->count >= n->cap) return false;
  const size_t N = n->N, off = (size_t)n->count * N;
  auto put = [&](const double *src, float *dst, size_t m) {
    hipLaunchKernelGGL(k_to_f32, dim3(cdiv((long long)m, kBlock)), dim3(kBlock), 0, stream, (long long)m, src, dst);
  };
  put(b, n->ds_b.p + off, N);
  put(x, n->ds_x.p + off, N);
  put(ac, n->ds_ac.p + off, N);
  put(vol, n->ds_vol.p + off, N);
  put(xp, n->ds_xp.p + off, N);
  put(anb, n->ds_anb.p + off * n->W, N * n->W);
  n->count++;
  return true;
}

} // namespace fs
