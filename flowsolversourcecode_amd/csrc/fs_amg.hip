// Aggregation algebraic multigrid on the device.
//
// Setup (host, sequential, run once and then every reAgglomorate calls):
// the greedy seed aggregation of AgglomorateALevelAlgebraicly
// (SolveEquations/AlgebraicMultiGrid/AgglomorateALevelAlgebraicly.for:48-331),
// coarse connectivity in the reference's insertion order (:733-787) and the
// level loop of AgglomorateGrid.for:36-44.  The maps are integer and
// bit-exact with the reference's Parents / Children / ElementParent /
// ElementNeighborMG arrays.
//
// Per call (device): Galerkin operators A_c = P^T A P with the reference's
// summation order (AssembleElementBasedAgglomoratedLHS.for:29-150), one
// thread per coarse matrix entry over a precomputed contribution list, so the
// FP64 coarse operators are bitwise the reference's acMG/anbMG.
//
// Cycle (device): the additive-correction V-cycle of MultiGridCycle.for:32-47
// with restriction = sum of member residuals in fine-cell order
// (AssembleNodeAndElementBasedAgglomoratedRHS.for:25-33) and prolongation =
// injection (CorrectSolutionOnFinerLevel.for:17-24).  The reference's
// sequential SOR smoother becomes multicolour symmetric Gauss-Seidel (or
// damped Jacobi), so cycle iterates differ while the converged solution does
// not.  Each visit of a coarse level starts from a zero correction, the
// standard V-cycle.  The hierarchy can be held in FP32 (FS_OPT_MG_FP32):
// Galerkin sums stay FP64 and are rounded once to FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "fs_common.h"
#include "fs_kernels.h"
#include "fs_solve.h"

namespace fs {

namespace {

inline int gfor(int n) { return std::max(1, (n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------ host setup
struct HostLevel {
  int n = 0, w = 0;
  std::vector<int> nbr;  // 1-based, 0 = none, [j*n + i]
  std::vector<int> nnbr;
  std::vector<double> ac, anb;
  std::vector<int> parent;                 // 1-based, to the next level
  std::vector<int> seed, nchild;           // of this level's aggregates
  std::vector<std::vector<int>> children;  // insertion order
};

int max_children(int ilevel) { // AgglomorateALevelAlgebraicly.for:79-105
  if (ilevel == 2) return 8;
  if (ilevel == 3) return 4;
  if (ilevel == 4) return 3;
  return 2;
}

void aggregate(HostLevel &f, HostLevel &c, int ilevel) {
  const int nf = f.n, maxch = max_children(ilevel);
  const double ratiomin = 0.5, tiny = 1.0e-15;
  std::vector<double> maxanb(nf);
  for (int i = 0; i < nf; ++i) {
    double m = -1.0;
    for (int j = 0; j < f.nnbr[i]; ++j) m = std::fmax(m, -f.anb[(size_t)j * nf + i]);
    maxanb[i] = m;
  }
  std::vector<int> parents(nf, 0);
  std::vector<std::vector<int>> kids(1);
  std::vector<int> seeds(1, 0);
  int ip = 1;
  for (int s1 = 1; s1 <= nf; ++s1) {
    if (parents[s1 - 1] != 0) continue;
    parents[s1 - 1] = ip;
    if ((int)kids.size() <= ip) {
      kids.resize(ip + 1);
      seeds.resize(ip + 1, 0);
    }
    kids[ip].clear();
    int nb = 0;
    for (int l = 0; l < f.nnbr[s1 - 1]; ++l) {
      const int inb = f.nbr[(size_t)l * nf + s1 - 1];
      if (inb == 0 || parents[inb - 1] != 0) continue;
      const double ratio = -f.anb[(size_t)l * nf + s1 - 1] / std::fmax(maxanb[s1 - 1], tiny);
      if (ratio > ratiomin && nb < maxch) {
        parents[inb - 1] = ip;
        kids[ip].push_back(inb);
        ++nb;
      }
    }
    const int nb2 = nb;
    for (int cl = 0; cl < nb2; ++cl) {
      const int ich = kids[ip][cl];
      for (int l = 0; l < f.nnbr[ich - 1]; ++l) {
        const int inb = f.nbr[(size_t)l * nf + ich - 1];
        if (inb == 0 || parents[inb - 1] != 0) continue;
        const double ratio = -f.anb[(size_t)l * nf + ich - 1] / std::fmax(maxanb[ich - 1], tiny);
        if (ratio > ratiomin && nb < maxch) {
          parents[inb - 1] = ip;
          ++nb;
          kids[ip].push_back(inb);
        }
      }
    }
    if (nb == 0) {
      parents[s1 - 1] = 0;
    } else {
      seeds[ip] = s1;
      ++ip;
    }
  }
  for (int io = 1; io <= nf; ++io) {
    if (parents[io - 1] == 0) {
      double strength = 0.0;
      bool placed = false;
      for (int l = 0; l < f.nnbr[io - 1]; ++l) {
        const int inb = f.nbr[(size_t)l * nf + io - 1];
        if (inb == 0) continue;
        if (parents[inb - 1] != 0 && !placed) {
          const double ratio = -f.anb[(size_t)l * nf + io - 1] / maxanb[inb - 1];
          if (strength < ratio) {
            placed = true;
            strength = ratio;
            const int p = parents[inb - 1];
            parents[io - 1] = p;
            kids[p].push_back(io);
          }
        }
      }
    }
    if (parents[io - 1] == 0) { // isolated orphan: aggregate without a seed
      parents[io - 1] = ip;
      if ((int)kids.size() <= ip) {
        kids.resize(ip + 1);
        seeds.resize(ip + 1, 0);
      }
      kids[ip].clear();
      seeds[ip] = 0;
      ++ip;
    }
  }
  const int nc = ip - 1;
  std::vector<std::vector<int>> cn(nc + 1);
  int wmax = 1;
  for (int e = 1; e <= nf; ++e)
    for (int l = 0; l < f.nnbr[e - 1]; ++l) {
      const int inb = f.nbr[(size_t)l * nf + e - 1];
      if (inb == 0) continue;
      const int p = parents[e - 1], q = parents[inb - 1];
      if (p == q) continue;
      auto &lst = cn[p];
      if (std::find(lst.begin(), lst.end(), q) == lst.end()) {
        lst.push_back(q);
        wmax = std::max(wmax, (int)lst.size());
      }
    }
  f.parent = parents;
  c.n = nc;
  c.w = wmax;
  c.nbr.assign((size_t)nc * wmax, 0);
  c.nnbr.assign(nc, 0);
  c.seed.assign(nc, 0);
  c.nchild.assign(nc, 0);
  c.children.assign(nc, {});
  for (int p = 1; p <= nc; ++p) {
    c.seed[p - 1] = seeds[p];
    c.children[p - 1] = kids[p];
    c.nchild[p - 1] = (int)kids[p].size();
    c.nnbr[p - 1] = (int)cn[p].size();
    for (size_t k = 0; k < cn[p].size(); ++k) c.nbr[k * nc + p - 1] = cn[p][k];
  }
  c.ac.assign(nc, 0.0);
  c.anb.assign((size_t)nc * wmax, 0.0);
}

// Galerkin contribution lists of level c in the reference's order and the
// host values (for the aggregation of the next level).
void galerkin_lists(const HostLevel &f, HostLevel &c, std::vector<int> &ptr,
                    std::vector<long long> &src) {
  const int nc = c.n, wc = c.w, nf = f.n;
  std::vector<std::vector<long long>> lists((size_t)nc * (wc + 1));
  for (int p = 1; p <= nc; ++p) {
    auto &dl = lists[p - 1];
    double acc = 0.0;
    const int nmem = c.nchild[p - 1] + 1;
    for (int mm = 0; mm < nmem; ++mm) {
      const int e = (mm == 0) ? c.seed[p - 1] : c.children[p - 1][mm - 1];
      if (e <= 0) continue;
      dl.push_back(e - 1);
      acc = acc + f.ac[e - 1];
      for (int l = 0; l < f.nnbr[e - 1]; ++l) {
        const int inb = f.nbr[(size_t)l * nf + e - 1];
        if (inb <= 0) continue;
        const int q = f.parent[inb - 1];
        if (q <= 0) continue;
        const size_t lin = (size_t)l * nf + e - 1;
        const double a = f.anb[lin];
        if (q == p) {
          dl.push_back((long long)nf + (long long)lin);
          acc = acc + a;
        } else {
          for (int k = 0; k < c.nnbr[p - 1]; ++k)
            if (c.nbr[(size_t)k * nc + p - 1] == q) {
              lists[(size_t)nc + (size_t)k * nc + p - 1].push_back((long long)nf + (long long)lin);
              c.anb[(size_t)k * nc + p - 1] = c.anb[(size_t)k * nc + p - 1] + a;
            }
        }
      }
    }
    c.ac[p - 1] = acc;
  }
  ptr.assign(lists.size() + 1, 0);
  src.clear();
  for (size_t e = 0; e < lists.size(); ++e) {
    ptr[e + 1] = ptr[e] + (int)lists[e].size();
    src.insert(src.end(), lists[e].begin(), lists[e].end());
  }
  if (src.empty()) src.push_back(0);
}

// ------------------------------------------------------------ device kernels

// coarse entry e of level c = sum of its contributions (FP64), then stored
// in both precisions when the hierarchy runs in FP32
__global__ void k_galerkin(int ne, int nc, const int *__restrict__ ptr,
                           const long long *__restrict__ src, int nf,
                           const double *__restrict__ acf, const double *__restrict__ anbf,
                           double *__restrict__ acc_out, double *__restrict__ anbc,
                           float *__restrict__ ac32, float *__restrict__ anb32) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= ne) return;
  double a = 0.0;
  for (int t = ptr[e]; t < ptr[e + 1]; ++t) {
    const long long s = src[t];
    a = a + (s < nf ? acf[s] : anbf[s - nf]);
  }
  if (e < nc) {
    acc_out[e] = a;
    if (ac32) ac32[e] = (float)a;
  } else {
    anbc[e - nc] = a;
    if (anb32) anb32[e - nc] = (float)a;
  }
}

template <class T>
__global__ void k_to(int n, const double *__restrict__ a, T *__restrict__ b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}
template <class T>
__global__ void k_from(int n, const T *__restrict__ a, double *__restrict__ b,
                       const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

template <class T>
__device__ __forceinline__ T lres(int i, int n, int w, const int *__restrict__ nbr,
                                  const T *__restrict__ ac, const T *__restrict__ anb,
                                  const T *__restrict__ b, const T *__restrict__ x) {
  T r = b[i] - ac[i] * x[i];
  for (int j = 0; j < w; ++j) {
    const int k = nbr[(size_t)j * n + i];
    if (k >= 0) r = r - anb[(size_t)j * n + i] * x[k];
  }
  return r;
}

// first damped-Jacobi sweep from a zero start: x = omega*b/ac
template <class T>
__global__ void k_jac0(int n, const T *__restrict__ ac, const T *__restrict__ b, T *__restrict__ x,
                       T omega, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) x[i] = omega * b[i] / ac[i];
}

// x_new = x + omega*(b - A x)/ac
template <class T>
__global__ void k_jac(int n, int w, const int *__restrict__ nbr, const T *__restrict__ ac,
                      const T *__restrict__ anb, const T *__restrict__ b, const T *__restrict__ x,
                      T *__restrict__ xn, T omega, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) xn[i] = x[i] + omega * lres(i, n, w, nbr, ac, anb, b, x) / ac[i];
}

template <class T>
__global__ void k_gsc(int cnt, const int *__restrict__ cells, int n, int w,
                      const int *__restrict__ nbr, const T *__restrict__ ac,
                      const T *__restrict__ anb, const T *__restrict__ b, T *__restrict__ x,
                      const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int i = cells[t];
  T s = b[i];
  for (int j = 0; j < w; ++j) {
    const int k = nbr[(size_t)j * n + i];
    if (k >= 0) s = s - anb[(size_t)j * n + i] * x[k];
  }
  x[i] = s / ac[i];
}

template <class T>
__global__ void k_zero(int n, T *__restrict__ x, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = T(0);
}

template <class T>
__global__ void k_lresid(int n, int w, const int *__restrict__ nbr, const T *__restrict__ ac,
                         const T *__restrict__ anb, const T *__restrict__ b,
                         const T *__restrict__ x, T *__restrict__ r, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) r[i] = lres(i, n, w, nbr, ac, anb, b, x);
}

// b_c[p] = sum of member residuals in increasing fine-cell order
template <class T>
__global__ void k_restrict(int nc, const int *__restrict__ mptr, const int *__restrict__ mem,
                           const T *__restrict__ rf, T *__restrict__ bc,
                           const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nc) return;
  T s = T(0);
  for (int t = mptr[p]; t < mptr[p + 1]; ++t) s = s + rf[mem[t]];
  bc[p] = s;
}

template <class T>
__global__ void k_prolong(int nf, const int *__restrict__ parent, const T *__restrict__ xc,
                          T *__restrict__ xf, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < nf) xf[i] = xf[i] + xc[parent[i]];
}

// AMG-as-solver residual: r = b - A x and partial L1(r)
__global__ void k_amg_res(int n, int w, const int *__restrict__ nbr,
                          const double *__restrict__ ac, const double *__restrict__ anb,
                          const double *__restrict__ b, const double *__restrict__ x,
                          double *__restrict__ r, double *__restrict__ part,
                          const int *__restrict__ flags) {
  if (flags[0]) return;
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double v = b[i] - ac[i] * x[i];
    for (int j = 0; j < w; ++j) {
      const int k = nbr[(size_t)j * n + i];
      if (k >= 0) v = v - anb[(size_t)j * n + i] * x[k];
    }
    r[i] = v;
    s += fabs(v);
  }
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_add2(int n, double *__restrict__ x, const double *__restrict__ c,
                       const int *__restrict__ flags) {
  if (flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + c[i];
}



// ---- lanes-per-row forms for the wide coarse levels: a row's slots are
// spread over LPR consecutive lanes and combined with shuffles, so a row
// costs one dependent-load chain instead of w of them.
template <int LPR, class T>
__device__ __forceinline__ T row_offdiag(int i, int lane, int n, int w, const int *__restrict__ nbr,
                                         const T *__restrict__ anb, const T *__restrict__ x) {
  T s = T(0);
  for (int j = lane; j < w; j += LPR) {
    const int k = nbr[(size_t)j * n + i];
    if (k >= 0) s += anb[(size_t)j * n + i] * x[k];
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
  return s;
}

template <int LPR, class T>
__global__ void k_lresid_lpr(int n, int w, const int *__restrict__ nbr, const T *__restrict__ ac,
                             const T *__restrict__ anb, const T *__restrict__ b,
                             const T *__restrict__ x, T *__restrict__ r, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int i = t / LPR, lane = t % LPR;
  const bool ok = i < n;
  const T s = row_offdiag<LPR, T>(ok ? i : 0, lane, ok ? n : 1, ok ? w : 0, nbr, anb, x);
  if (ok && lane == 0) r[i] = b[i] - ac[i] * x[i] - s;
}

template <int LPR, class T>
__global__ void k_jac_lpr(int n, int w, const int *__restrict__ nbr, const T *__restrict__ ac,
                          const T *__restrict__ anb, const T *__restrict__ b, const T *__restrict__ x,
                          T *__restrict__ xn, T omega, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int i = t / LPR, lane = t % LPR;
  const bool ok = i < n;
  const T s = row_offdiag<LPR, T>(ok ? i : 0, lane, ok ? n : 1, ok ? w : 0, nbr, anb, x);
  if (ok && lane == 0) xn[i] = x[i] + omega * (b[i] - ac[i] * x[i] - s) / ac[i];
}


// ------------------------------------------------------------ small levels
// Every level below the tail threshold is processed by ONE workgroup in one
// launch: restriction, smoothing, the coarsest level and the prolongations of
// those levels, separated by workgroup barriers.  The small levels are
// latency-bound (a few thousand cells, tens of slots), so one resident
// workgroup with its data in L2 replaces dozens of tiny launches.
struct TailLevel {
  int n, w, ncolors;
  const int *nbr;
  const void *ac, *anb;
  void *x, *b, *r, *t;
  const int *parent;              // to the next coarser level, or null
  const int *mptr, *mem;          // members of this level's aggregates in the finer level
  const int *color_ptr, *color_cells;
};

constexpr int kTailThreads = 1024;

constexpr int kTailLPR = 8;

// y_i = b_i - ac_i x_i - sum_j anb_ij x_j for the rows of one level, LPR lanes
// per row (all 1024 threads of the workgroup take part)
template <class T>
__device__ void tail_rows(const TailLevel &L, const T *x, T omega, T *out, int mode) {
  const T *ac = (const T *)L.ac, *anb = (const T *)L.anb, *b = (const T *)L.b;
  const int lane = threadIdx.x % kTailLPR;
  const int rows_per_pass = blockDim.x / kTailLPR;
  for (int base = 0; base < L.n; base += rows_per_pass) {
    const int i = base + threadIdx.x / kTailLPR;
    const bool ok = i < L.n;
    const T s = row_offdiag<kTailLPR, T>(ok ? i : 0, lane, ok ? L.n : 1, ok ? L.w : 0, L.nbr, anb, x);
    if (ok && lane == 0) {
      const T r = b[i] - ac[i] * x[i] - s;
      out[i] = mode == 0 ? r : x[i] + omega * r / ac[i];
    }
  }
}

// damped Jacobi sweeps on one level (x and t alternate; the result ends in x)
template <class T> __device__ void tail_smooth(const TailLevel &L, int sweeps, bool zero, T omega) {
  const T *ac = (const T *)L.ac, *b = (const T *)L.b;
  T *x = (T *)L.x, *t = (T *)L.t;
  int sw = 0;
  if (zero) {
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) x[i] = sweeps > 0 ? omega * b[i] / ac[i] : T(0);
    __syncthreads();
    sw = 1;
  }
  for (; sw < sweeps; ++sw) {
    tail_rows<T>(L, x, omega, t, 1);
    __syncthreads();
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) x[i] = t[i];
    __syncthreads();
  }
}

template <class T>
__global__ void __launch_bounds__(kTailThreads)
k_tail(const TailLevel *__restrict__ lv, int nlev, int pre, int post, int coarse, T omega,
       const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  for (int l = 0; l + 1 < nlev; ++l) {
    const TailLevel &L = lv[l];
    tail_smooth<T>(L, pre, true, omega);
    tail_rows<T>(L, (const T *)L.x, omega, (T *)L.r, 0);
    __syncthreads();
    const TailLevel &C = lv[l + 1];
    const T *r = (const T *)L.r;
    T *bc = (T *)C.b;
    for (int p = threadIdx.x; p < C.n; p += blockDim.x) {
      T s = T(0);
      for (int q = C.mptr[p]; q < C.mptr[p + 1]; ++q) s = s + r[C.mem[q]];
      bc[p] = s;
    }
    __syncthreads();
  }
  tail_smooth<T>(lv[nlev - 1], coarse, true, omega);
  for (int l = nlev - 2; l >= 0; --l) {
    const TailLevel &L = lv[l];
    T *x = (T *)L.x;
    const T *xc = (const T *)lv[l + 1].x;
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) x[i] = x[i] + xc[L.parent[i]];
    __syncthreads();
    tail_smooth<T>(L, post, false, omega);
  }
}

} // namespace

// ------------------------------------------------------------ hierarchy
struct DLevel {
  int n = 0, w = 0;
  const int *nbr = nullptr;      // device, 0-based/-1
  DevBuf<int> nbr_own;
  const double *ac64 = nullptr, *anb64 = nullptr;
  DevBuf<double> ac64_own, anb64_own;
  DevBuf<float> ac32, anb32;
  DevBuf<double> x64, b64, r64, t64;
  DevBuf<float> x32, b32, r32, t32;
  std::vector<int> color_ptr;
  DevBuf<int> color_cells, color_ptr_dev;
  DevBuf<int> parent;            // 0-based parent in the next level
  DevBuf<int> mptr, mem;         // members of this level's aggregates (coarse)
  DevBuf<int> gptr;
  DevBuf<long long> gsrc;
  int gne = 0;
};

struct Hierarchy {
  std::vector<HostLevel> host;   // maps for export / tests
  std::vector<DLevel> lev;
  int calls = 0;
  bool fp32 = false;
  int tail = -1;                 // first level run by the one-workgroup tail
  DevBuf<TailLevel> tail32, tail64;
};

static void build_hierarchy(fs_mesh *m, const double *ac, const double *anb) {
  auto H = std::make_shared<Hierarchy>();
  hipStream_t s = m->stream;
  const int n = m->n, w = m->w;
  HostLevel L0;
  L0.n = n;
  L0.w = w;
  L0.nbr = m->h_nbr;
  L0.nnbr = m->h_nnbr;
  L0.ac.resize(n);
  L0.anb.resize((size_t)n * w);
  FS_HIP_THROW(hipMemcpyAsync(L0.ac.data(), ac, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  FS_HIP_THROW(hipMemcpyAsync(L0.anb.data(), anb, sizeof(double) * n * w, hipMemcpyDeviceToHost, s));
  FS_HIP_THROW(hipStreamSynchronize(s));
  H->host.push_back(std::move(L0));
  std::vector<std::vector<int>> gptrs(1);
  std::vector<std::vector<long long>> gsrcs(1);
  int nl = 1;
  while (nl < m->mg_maxlevels && H->host[nl - 1].n > 2 * m->mg_minparents) {
    HostLevel c;
    aggregate(H->host[nl - 1], c, nl + 1);
    std::vector<int> gp;
    std::vector<long long> gs;
    galerkin_lists(H->host[nl - 1], c, gp, gs);
    H->host.push_back(std::move(c));
    gptrs.push_back(std::move(gp));
    gsrcs.push_back(std::move(gs));
    ++nl;
  }
  H->lev.resize(nl);
  for (int l = 0; l < nl; ++l) {
    DLevel &D = H->lev[l];
    HostLevel &h = H->host[l];
    D.n = h.n;
    D.w = h.w;
    std::vector<int> nbr0((size_t)h.n * h.w);
    for (int i = 0; i < h.n; ++i)
      for (int j = 0; j < h.w; ++j) {
        const int k = h.nbr[(size_t)j * h.n + i];
        nbr0[(size_t)j * h.n + i] = (j < h.nnbr[i] && k > 0) ? k - 1 : -1;
      }
    if (l == 0) {
      D.nbr = m->d_nbr.p;
    } else {
      D.nbr_own.upload(nbr0, s);
      D.nbr = D.nbr_own.p;
      D.ac64_own.alloc(h.n);
      D.anb64_own.alloc((size_t)h.n * h.w);
      D.ac64 = D.ac64_own.p;
      D.anb64 = D.anb64_own.p;
      D.gptr.upload(gptrs[l], s);
      D.gsrc.upload(gsrcs[l], s);
      D.gne = (int)gptrs[l].size() - 1;
      // members in increasing fine order
      const HostLevel &f = H->host[l - 1];
      std::vector<int> cnt(h.n + 1, 0), mem(f.n);
      for (int i = 0; i < f.n; ++i) cnt[f.parent[i]]++;
      for (int p = 0; p < h.n; ++p) cnt[p + 1] += cnt[p];
      std::vector<int> fill(cnt.begin(), cnt.end() - 1);
      for (int i = 0; i < f.n; ++i) mem[fill[f.parent[i] - 1]++] = i;
      D.mptr.upload(cnt, s);
      D.mem.upload(mem, s);
    }
    if (l + 1 < nl) {
      std::vector<int> par0(h.n);
      for (int i = 0; i < h.n; ++i) par0[i] = h.parent[i] - 1;
      D.parent.upload(par0, s);
    }
    std::vector<int> cells;
    greedy_colors(h.n, h.w, nbr0.data(), D.color_ptr, cells);
    D.color_cells.upload(cells, s);
    D.color_ptr_dev.upload(D.color_ptr, s);
    D.x64.alloc(h.n);
    D.b64.alloc(h.n);
    D.r64.alloc(h.n);
    D.t64.alloc(h.n);
    D.x32.alloc(h.n);
    D.b32.alloc(h.n);
    D.r32.alloc(h.n);
    D.t32.alloc(h.n);
    D.ac32.alloc(h.n);
    D.anb32.alloc((size_t)h.n * h.w);
  }
  // levels from the first one at or below the threshold run in the tail
  H->tail = nl; // level 0 always runs grid-wide (it is the caller's operator)
  for (int l = 1; l < nl; ++l)
    if (H->lev[l].n <= m->mg_tail_cells) {
      H->tail = l;
      break;
    }
  if (H->tail < nl) {
    const int nt = nl - H->tail;
    std::vector<TailLevel> d32(nt), d64(nt);
    for (int q = 0; q < nt; ++q) {
      DLevel &D = H->lev[H->tail + q];
      for (int prec = 0; prec < 2; ++prec) {
        TailLevel t{};
        t.n = D.n;
        t.w = D.w;
        t.ncolors = (int)D.color_ptr.size() - 1;
        t.nbr = D.nbr;
        t.ac = prec ? (const void *)D.ac64_own.p : (const void *)D.ac32.p;
        t.anb = prec ? (const void *)D.anb64_own.p : (const void *)D.anb32.p;
        t.x = prec ? (void *)D.x64.p : (void *)D.x32.p;
        t.b = prec ? (void *)D.b64.p : (void *)D.b32.p;
        t.r = prec ? (void *)D.r64.p : (void *)D.r32.p;
        t.t = prec ? (void *)D.t64.p : (void *)D.t32.p;
        t.parent = (H->tail + q + 1 < nl) ? D.parent.p : nullptr;
        t.mptr = q > 0 ? D.mptr.p : nullptr;
        t.mem = q > 0 ? D.mem.p : nullptr;
        t.color_ptr = D.color_ptr_dev.p;
        t.color_cells = D.color_cells.p;
        (prec ? d64 : d32)[q] = t;
      }
    }
    H->tail32.upload(d32, s);
    H->tail64.upload(d64, s);
  }
  FS_HIP_THROW(hipStreamSynchronize(s));
  m->amg = H;
}

void amg_prepare(fs_mesh *m, const double *ac, const double *anb) {
  const bool rebuild = !m->amg || (m->mg_reagg > 0 && m->ncalls_mg % m->mg_reagg == 0) ||
                       m->amg->lev.empty() || m->amg->lev[0].n != m->n;
  if (rebuild) build_hierarchy(m, ac, anb);
  m->ncalls_mg++;
  Hierarchy &H = *m->amg;
  H.fp32 = m->mg_fp32 != 0;
  hipStream_t s = m->stream;
  DLevel &L0 = H.lev[0];
  L0.ac64 = ac;
  L0.anb64 = anb;
  if (H.fp32) {
    hipLaunchKernelGGL(k_to<float>, dim3(gfor(L0.n)), dim3(kBlock), 0, s, L0.n, ac, L0.ac32.p);
    hipLaunchKernelGGL(k_to<float>, dim3(gfor(L0.n * L0.w)), dim3(kBlock), 0, s, L0.n * L0.w, anb,
                       L0.anb32.p);
  }
  for (size_t l = 1; l < H.lev.size(); ++l) {
    DLevel &F = H.lev[l - 1], &C = H.lev[l];
    hipLaunchKernelGGL(k_galerkin, dim3(gfor(C.gne)), dim3(kBlock), 0, s, C.gne, C.n, C.gptr.p,
                       C.gsrc.p, F.n, F.ac64, F.anb64, C.ac64_own.p, C.anb64_own.p,
                       H.fp32 ? C.ac32.p : nullptr, H.fp32 ? C.anb32.p : nullptr);
  }
}

int amg_levels(fs_mesh *m) { return m->amg ? (int)m->amg->lev.size() : 1; }

template <class T> struct LV {
  static T *x(DLevel &d);
  static T *b(DLevel &d);
  static T *r(DLevel &d);
  static T *t(DLevel &d);
  static const T *ac(DLevel &d);
  static const T *anb(DLevel &d);
};
template <> struct LV<double> {
  static double *x(DLevel &d) { return d.x64.p; }
  static double *b(DLevel &d) { return d.b64.p; }
  static double *r(DLevel &d) { return d.r64.p; }
  static double *t(DLevel &d) { return d.t64.p; }
  static const double *ac(DLevel &d) { return d.ac64; }
  static const double *anb(DLevel &d) { return d.anb64; }
};
template <> struct LV<float> {
  static float *x(DLevel &d) { return d.x32.p; }
  static float *b(DLevel &d) { return d.b32.p; }
  static float *r(DLevel &d) { return d.r32.p; }
  static float *t(DLevel &d) { return d.t32.p; }
  static const float *ac(DLevel &d) { return d.ac32.p; }
  static const float *anb(DLevel &d) { return d.anb32.p; }
};

template <class T>
static void launch_jac(fs_mesh *m, DLevel &D, const T *b, const T *x, T *xn, T om, const int *flags) {
  hipStream_t s = m->stream;
  const T *ac = LV<T>::ac(D), *anb = LV<T>::anb(D);
  if (D.w <= 8)
    hipLaunchKernelGGL(k_jac<T>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.w, D.nbr, ac, anb, b, x, xn,
                       om, flags);
  else if (D.w <= 16)
    hipLaunchKernelGGL((k_jac_lpr<4, T>), dim3(gfor(D.n * 4)), dim3(kBlock), 0, s, D.n, D.w, D.nbr, ac,
                       anb, b, x, xn, om, flags);
  else if (D.w <= 32)
    hipLaunchKernelGGL((k_jac_lpr<8, T>), dim3(gfor(D.n * 8)), dim3(kBlock), 0, s, D.n, D.w, D.nbr, ac,
                       anb, b, x, xn, om, flags);
  else
    hipLaunchKernelGGL((k_jac_lpr<16, T>), dim3(gfor(D.n * 16)), dim3(kBlock), 0, s, D.n, D.w, D.nbr,
                       ac, anb, b, x, xn, om, flags);
}

template <class T>
static void launch_resid(fs_mesh *m, DLevel &D, const T *b, const T *x, T *r, const int *flags) {
  hipStream_t s = m->stream;
  const T *ac = LV<T>::ac(D), *anb = LV<T>::anb(D);
  if (D.w <= 8)
    hipLaunchKernelGGL(k_lresid<T>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.w, D.nbr, ac, anb, b, x, r,
                       flags);
  else if (D.w <= 16)
    hipLaunchKernelGGL((k_lresid_lpr<4, T>), dim3(gfor(D.n * 4)), dim3(kBlock), 0, s, D.n, D.w, D.nbr, ac,
                       anb, b, x, r, flags);
  else if (D.w <= 32)
    hipLaunchKernelGGL((k_lresid_lpr<8, T>), dim3(gfor(D.n * 8)), dim3(kBlock), 0, s, D.n, D.w, D.nbr, ac,
                       anb, b, x, r, flags);
  else
    hipLaunchKernelGGL((k_lresid_lpr<16, T>), dim3(gfor(D.n * 16)), dim3(kBlock), 0, s, D.n, D.w, D.nbr,
                       ac, anb, b, x, r, flags);
}

template <class T>
static void smooth(fs_mesh *m, DLevel &D, const T *b, T *&x, T *&tmp, int sweeps, bool zero_start,
                   const int *flags) {
  hipStream_t s = m->stream;
  const T *ac = LV<T>::ac(D), *anb = LV<T>::anb(D);
  const int nc = (int)D.color_ptr.size() - 1;
  // multicolour Gauss-Seidel only where it is cheap: few colours, narrow rows
  const bool gs = m->mg_smoother == 1 && nc <= 4 && D.w <= 8;
  const T om = (T)m->mg_omega;
  if (gs) {
    if (zero_start) hipLaunchKernelGGL(k_zero<T>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, x, flags);
    for (int sw = 0; sw < sweeps; ++sw) {
      for (int c = 0; c < nc; ++c) {
        const int bb = D.color_ptr[c], e = D.color_ptr[c + 1];
        hipLaunchKernelGGL(k_gsc<T>, dim3(gfor(e - bb)), dim3(kBlock), 0, s, e - bb,
                           D.color_cells.p + bb, D.n, D.w, D.nbr, ac, anb, b, x, flags);
      }
      for (int c = nc - 1; c >= 0; --c) {
        const int bb = D.color_ptr[c], e = D.color_ptr[c + 1];
        hipLaunchKernelGGL(k_gsc<T>, dim3(gfor(e - bb)), dim3(kBlock), 0, s, e - bb,
                           D.color_cells.p + bb, D.n, D.w, D.nbr, ac, anb, b, x, flags);
      }
    }
  } else {
    int sw = 0;
    if (zero_start && sweeps > 0) {
      hipLaunchKernelGGL(k_jac0<T>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, ac, b, x, om, flags);
      sw = 1;
    } else if (zero_start) {
      hipLaunchKernelGGL(k_zero<T>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, x, flags);
    }
    for (; sw < sweeps; ++sw) {
      launch_jac<T>(m, D, b, x, tmp, om, flags);
      std::swap(x, tmp);
    }
  }
}

// One V-cycle: x_0 = M^{-1} b_0 on the hierarchy held in precision T.
// Levels 0..tail-1 are swept by grid-wide kernels, levels tail..L-1 by the
// one-workgroup tail kernel.
template <class T> static void vcycle(fs_mesh *m, const T *b0, T *x0, const int *flags) {
  Hierarchy &H = *m->amg;
  hipStream_t s = m->stream;
  const int L = (int)H.lev.size();
  const int tail = std::min(H.tail, L); // >= 1 by construction
  std::vector<T *> X(L), Tm(L);
  std::vector<const T *> B(L);
  for (int l = 0; l < L; ++l) {
    X[l] = LV<T>::x(H.lev[l]);
    Tm[l] = LV<T>::t(H.lev[l]);
    B[l] = LV<T>::b(H.lev[l]);
  }
  B[0] = b0;
  const int top = tail < L ? tail : L - 1; // levels swept by grid-wide kernels: 0..top-1
  for (int l = 0; l < top; ++l) {
    DLevel &D = H.lev[l];
    smooth<T>(m, D, B[l], X[l], Tm[l], m->mg_pre, true, flags);
    launch_resid<T>(m, D, B[l], X[l], LV<T>::r(D), flags);
    DLevel &C = H.lev[l + 1];
    hipLaunchKernelGGL(k_restrict<T>, dim3(gfor(C.n)), dim3(kBlock), 0, s, C.n, C.mptr.p, C.mem.p,
                       LV<T>::r(D), LV<T>::b(C), flags);
  }
  const int coarse = std::max(m->mg_coarse_sweeps, 1);
  if (tail < L) {
    const TailLevel *desc = (sizeof(T) == 4) ? H.tail32.p : H.tail64.p;
    hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(kTailThreads), 0, s, desc, L - tail, m->mg_pre,
                       m->mg_post, coarse, (T)m->mg_omega, flags);
    X[tail] = LV<T>::x(H.lev[tail]); // the tail keeps its result in x
  } else {
    DLevel &D = H.lev[L - 1];
    smooth<T>(m, D, B[L - 1], X[L - 1], Tm[L - 1], coarse, true, flags);
  }
  for (int l = top - 1; l >= 0; --l) {
    DLevel &D = H.lev[l];
    hipLaunchKernelGGL(k_prolong<T>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.parent.p, X[l + 1],
                       X[l], flags);
    smooth<T>(m, D, B[l], X[l], Tm[l], m->mg_post, false, flags);
  }
  if (X[0] != x0) {
    FS_HIP_THROW(hipMemcpyAsync(x0, X[0], sizeof(T) * H.lev[0].n, hipMemcpyDeviceToDevice, s));
  }
}

void amg_apply(fs_mesh *m, const double *rhs, double *out, const int *flags) {
  Hierarchy &H = *m->amg;
  hipStream_t s = m->stream;
  DLevel &L0 = H.lev[0];
  if (H.fp32) {
    hipLaunchKernelGGL(k_to<float>, dim3(gfor(L0.n)), dim3(kBlock), 0, s, L0.n, rhs, L0.b32.p);
    vcycle<float>(m, L0.b32.p, L0.x32.p, flags);
    hipLaunchKernelGGL(k_from<float>, dim3(gfor(L0.n)), dim3(kBlock), 0, s, L0.n, L0.x32.p, out, flags);
  } else {
    vcycle<double>(m, rhs, out, flags);
  }
}

// The reference's multigrid solve loop (AlgebraicMultiGridMainRoutine.for:
// 371-416): cycles while the fine L1 residual exceeds rrf*ires and at most
// MaxMultiGridCycles; each cycle corrects with one V-cycle of the residual.
int amg_solve(fs_mesh *m, const double *ac, const double *anb, const double *bc, double *dphi,
              int solver) {
  (void)solver;
  amg_prepare(m, ac, anb);
  hipStream_t s = m->stream;
  const int n = m->n;
  double *r = m->v[8].p, *c = m->v[9].p;
  int *flags = m->flags.p;
  // flags[2] holds itmax for the Krylov solvers; the multigrid loop is
  // bounded by MaxMultiGridCycles instead
  int itmax_cycles = m->mg_maxcycles;
  FS_HIP_THROW(hipMemcpyAsync(flags + 2, &itmax_cycles, sizeof(int), hipMemcpyHostToDevice, s));
  FS_HIP_THROW(hipMemcpyAsync(r, bc, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  const int gr = std::min(kPartials, gfor(n));
  int fl[4] = {0, 0, 0, 0};
  FS_HIP_THROW(hipMemcpyAsync(fl, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
  FS_HIP_THROW(hipStreamSynchronize(s));
  int launched = 0;
  while (!fl[0] && launched < itmax_cycles) {
    const int batch = std::min(m->check_every, itmax_cycles - launched);
    for (int b = 0; b < batch; ++b) {
      amg_apply(m, r, c, flags);
      hipLaunchKernelGGL(k_add2, dim3(gfor(n)), dim3(kBlock), 0, s, n, dphi, c, flags);
      hipLaunchKernelGGL(k_amg_res, dim3(gr), dim3(kBlock), 0, s, n, m->w, m->d_nbr.p, ac, anb,
                         bc, dphi, r, m->partial.p, flags);
      launch_fin_conv(m, gr);
    }
    launched += batch;
    FS_HIP_THROW(hipMemcpyAsync(fl, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    FS_HIP_THROW(hipStreamSynchronize(s));
  }
  return fl[1];
}

int amg_export(fs_mesh *m, int level, int *dims, int *parent, int *seed, int *nchild,
               int *child_ptr, int *child, int *nbr, int *nnbr, double *ac, double *anb) {
  if (!m->amg || level < 0 || level >= (int)m->amg->host.size()) return -1;
  const HostLevel &h = m->amg->host[level];
  DLevel &D = m->amg->lev[level];
  int nch = 0;
  for (int p = 0; p < (int)h.children.size(); ++p) nch += (int)h.children[p].size();
  if (dims) {
    dims[0] = h.n;
    dims[1] = h.w;
    dims[2] = nch;
    dims[3] = (int)m->amg->lev.size();
    dims[4] = (int)D.color_ptr.size() - 1;
  }
  if (parent && !h.parent.empty()) std::copy(h.parent.begin(), h.parent.end(), parent);
  if (seed && !h.seed.empty()) std::copy(h.seed.begin(), h.seed.end(), seed);
  if (nchild && !h.nchild.empty()) std::copy(h.nchild.begin(), h.nchild.end(), nchild);
  if (child_ptr && child && !h.children.empty()) {
    int t = 0;
    child_ptr[0] = 0;
    for (int p = 0; p < h.n; ++p) {
      for (int c : h.children[p]) child[t++] = c;
      child_ptr[p + 1] = t;
    }
  }
  if (nbr) std::copy(h.nbr.begin(), h.nbr.end(), nbr);
  if (nnbr) std::copy(h.nnbr.begin(), h.nnbr.end(), nnbr);
  // coefficients: the device operators of the last prepare call for coarse
  // levels (level 0 belongs to the caller and may be gone: host copy of setup)
  if (level == 0) {
    if (ac) std::copy(h.ac.begin(), h.ac.end(), ac);
    if (anb) std::copy(h.anb.begin(), h.anb.end(), anb);
  } else {
    if (ac) FS_HIP_THROW(hipMemcpy(ac, D.ac64_own.p, sizeof(double) * h.n, hipMemcpyDeviceToHost));
    if (anb)
      FS_HIP_THROW(hipMemcpy(anb, D.anb64_own.p, sizeof(double) * h.n * h.w, hipMemcpyDeviceToHost));
  }
  return 0;
}

} // namespace fs
