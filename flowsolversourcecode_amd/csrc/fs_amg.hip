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
// injection (CorrectSolutionOnFinerLevel.for:17-24), optionally scaled by an
// over-correction factor alpha (FS_OPT_MG_ALPHA; 1 = the reference).  The
// reference's sequential SOR smoother becomes damped Jacobi (or multicolour
// symmetric Gauss-Seidel), so cycle iterates differ while the converged
// solution does not.  Each visit of a coarse level starts from a zero
// correction, the standard V-cycle.  The coarsest level (at most 64 cells) is
// solved directly in FP64.  The hierarchy can be held in FP32
// (FS_OPT_MG_FP32): Galerkin sums stay FP64 and are rounded once to FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
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

void finish_level(HostLevel &f, HostLevel &c, const std::vector<int> &parents, const std::vector<int> &seeds,
                  const std::vector<std::vector<int>> &kids, int nc);

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
  finish_level(f, c, parents, seeds, kids, ip - 1);
}

// The coarse level of an aggregation: parents (1-based) of the fine cells,
// seed and children (in insertion order) of aggregates 1..nc; the coarse
// neighbour lists in the order the fine cells and their slots reach them.
void finish_level(HostLevel &f, HostLevel &c, const std::vector<int> &parents, const std::vector<int> &seeds,
                  const std::vector<std::vector<int>> &kids, int nc) {
  const int nf = f.n;
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

// Off-diagonal couplings of a level as compressed rows (the diagonal is not
// needed to judge strength).
struct Couplings {
  int n = 0;
  std::vector<int> rp, col;
  std::vector<double> val;
};

Couplings couplings_of(const HostLevel &f) {
  Couplings A;
  A.n = f.n;
  A.rp.assign(f.n + 1, 0);
  for (int i = 0; i < f.n; ++i) {
    for (int l = 0; l < f.nnbr[i]; ++l) {
      const int j = f.nbr[(size_t)l * f.n + i];
      if (j <= 0) continue;
      A.col.push_back(j - 1);
      A.val.push_back(f.anb[(size_t)l * f.n + i]);
    }
    A.rp[i + 1] = (int)A.col.size();
  }
  return A;
}

// One pass of pairwise matching: each unmatched cell, in order, is paired
// with the unmatched neighbour of its strongest negative coupling among those
// at least `beta` times the row's strongest (Notay's pairwise aggregation);
// a cell without one stays alone.  agg[i] = 0-based pair number.
int pair_pass(const Couplings &A, double beta, std::vector<int> &agg) {
  const int n = A.n;
  agg.assign(n, -1);
  int na = 0;
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    double mx = 0.0;
    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) mx = std::fmax(mx, -A.val[k]);
    int best = -1;
    double bv = 0.0;
    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      const int j = A.col[k];
      const double sv = -A.val[k];
      if (j == i || agg[j] >= 0 || !(sv > 0.0) || sv < beta * mx) continue;
      if (sv > bv) {
        bv = sv;
        best = j;
      }
    }
    agg[i] = na;
    if (best >= 0) agg[best] = na;
    ++na;
  }
  return na;
}

// The couplings between pairs: the fine couplings summed over pair members,
// those inside a pair dropped.
Couplings pair_couplings(const Couplings &A, const std::vector<int> &agg, int na) {
  std::vector<std::vector<int>> mem(na);
  for (int i = 0; i < A.n; ++i) mem[agg[i]].push_back(i);
  Couplings B;
  B.n = na;
  B.rp.assign(na + 1, 0);
  std::vector<int> where(na, -1);
  std::vector<int> touched;
  for (int p = 0; p < na; ++p) {
    touched.clear();
    for (int i : mem[p])
      for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int q = agg[A.col[k]];
        if (q == p) continue;
        if (where[q] < 0) {
          where[q] = (int)B.col.size();
          B.col.push_back(q);
          B.val.push_back(0.0);
          touched.push_back(q);
        }
        B.val[where[q]] += A.val[k];
      }
    for (int q : touched) where[q] = -1;
    B.rp[p + 1] = (int)B.col.size();
  }
  return B;
}

// Aggregation by `passes` pairwise passes (aggregates of up to 2^passes
// cells).  Aggregates are numbered in the order of their first fine cell,
// which is also their seed; the other members follow in increasing order.
void aggregate_pairwise(HostLevel &f, HostLevel &c, int passes) {
  const double beta = 0.25;
  const Couplings A = couplings_of(f);
  std::vector<int> agg;
  int nc = pair_pass(A, beta, agg);
  if (passes > 1) {
    const Couplings B = pair_couplings(A, agg, nc);
    std::vector<int> agg2;
    nc = pair_pass(B, beta, agg2);
    for (int i = 0; i < f.n; ++i) agg[i] = agg2[agg[i]];
  }
  std::vector<int> ren(nc, 0);
  int t = 0;
  std::vector<int> parents(f.n);
  for (int i = 0; i < f.n; ++i) {
    if (ren[agg[i]] == 0) ren[agg[i]] = ++t;
    parents[i] = ren[agg[i]];
  }
  std::vector<int> seeds(nc + 1, 0);
  std::vector<std::vector<int>> kids(nc + 1);
  for (int i = 0; i < f.n; ++i) {
    const int p = parents[i];
    if (seeds[p] == 0) seeds[p] = i + 1;
    else kids[p].push_back(i + 1);
  }
  finish_level(f, c, parents, seeds, kids, nc);
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
                           const int *__restrict__ src, int nf,
                           const double *__restrict__ acf, const double *__restrict__ anbf,
                           double *__restrict__ acc_out, double *__restrict__ anbc,
                           float *__restrict__ ac32, float *__restrict__ anb32,
                           const int *__restrict__ csr_pos, double *__restrict__ val64,
                           float *__restrict__ val32) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= ne) return;
  // contributions in list order, loaded in batches (sources, then values)
  const int t0 = ptr[e], t1 = ptr[e + 1];
  double a = 0.0;
  constexpr int B = 4;
  for (int t = t0; t < t1; t += B) {
    int sv[B];
    double v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) sv[u] = src[min(t + u, t1 - 1)];
#pragma unroll
    for (int u = 0; u < B; ++u) v[u] = sv[u] < nf ? acf[sv[u]] : anbf[sv[u] - nf];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (t + u < t1) a = a + v[u];
  }
  if (e < nc) {
    acc_out[e] = a;
    if (ac32) ac32[e] = (float)a;
  } else {
    anbc[e - nc] = a;
    if (anb32) anb32[e - nc] = (float)a;
    // the same value at its place in the row-major CSR copy of the level
    const int q = csr_pos[e - nc];
    if (q >= 0) {
      val64[q] = a;
      if (val32) val32[q] = (float)a;
    }
  }
}

template <class T>
__global__ void k_to(int n, const double *__restrict__ a, T *__restrict__ b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}

// ------------------------------------------------------------ cycle kernels
// Grid-wide kernels for the levels above the tail.  Every kernel returns at
// once when the Krylov loop has set its done flag, so the batches launched
// after convergence cost only their launch.

// The level kernels load their matrix as ordinary (cached) data: a level's
// matrix is re-read by every sweep of a cycle and partly stays in the
// Infinity Cache between them (non-temporal loads here took the 128^3 fine
// sweep from 4.7 to 4.1 TB/s).
constexpr bool kStreamLevel = false;

// sum_j anb_ij x'_j over the slots of row i, spread over LPR consecutive
// lanes and combined with shuffles; x' = x + alpha*xc[parent] when PRO
template <int LPR, class T, bool PRO>
__device__ __forceinline__ T row_offdiag(int i, int lane, int n, int w, EllNbr nbr,
                                         const T *__restrict__ anb, const T *x, const int *__restrict__ parent,
                                         const T *__restrict__ xc, T alpha) {
  T s = T(0);
  if (LPR == 1) {
    // one lane per row: the row's slots in batches (loads first, then the
    // terms in slot order)
    for (int j0 = 0; j0 < w; j0 += kEllBatch) {
      int k[kEllBatch];
      T a[kEllBatch], xv[kEllBatch];
      ell_batch<kEllBatch, kStreamLevel>(nbr, i, n, w, j0, k);
#pragma unroll
      for (int u = 0; u < kEllBatch; ++u) {
        const int kk = k[u] >= 0 ? k[u] : i;
        a[u] = ld_mat<kStreamLevel>(&anb[(size_t)min(j0 + u, w - 1) * n + i]);
        xv[u] = x[kk];
        if (PRO) xv[u] = xv[u] + alpha * xc[parent[kk]];
      }
#pragma unroll
      for (int u = 0; u < kEllBatch; ++u)
        if (j0 + u < w && k[u] >= 0) s += a[u] * xv[u];
    }
    return s;
  }
  for (int j = lane; j < w; j += LPR) {
    const int k = nbr.i32[(size_t)j * n + i];
    if (k >= 0) {
      T xk = x[k];
      if (PRO) xk = xk + alpha * xc[parent[k]];
      s += anb[(size_t)j * n + i] * xk;
    }
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
  return s;
}

// One damped-Jacobi sweep xn = x' + omega*(b - A x')/ac, x' = x (PRO false)
// or x + alpha*xc[parent] (PRO true: the prolongation of the coarse
// correction fused into the first post-smoothing sweep).  TO = double writes
// the FP64 result of the cycle directly.
template <int LPR, class T, class TO, bool PRO>
__global__ void __launch_bounds__(kBlock)
k_sweep(int n, int w, EllNbr nbr, const T *__restrict__ ac, const T *__restrict__ anb,
        const T *__restrict__ b, const T *x, const int *__restrict__ parent, const T *__restrict__ xc,
        T alpha, T omega, TO *xn, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = xcd_block() * kBlock + threadIdx.x;
  const int i = t / LPR, lane = t % LPR;
  const bool ok = i < n;
  const T s = row_offdiag<LPR, T, PRO>(ok ? i : 0, lane, ok ? n : 1, ok ? w : 0, nbr, anb, x, parent, xc, alpha);
  if (ok && lane == 0) {
    T xi = x[i];
    if (PRO) xi = xi + alpha * xc[parent[i]];
    xn[i] = (TO)(xi + omega * (b[i] - ac[i] * xi - s) / ac[i]);
  }
}

// r = b - A x
template <int LPR, class T>
__global__ void __launch_bounds__(kBlock)
k_resid(int n, int w, EllNbr nbr, const T *__restrict__ ac, const T *__restrict__ anb,
        const T *__restrict__ b, const T *__restrict__ x, T *__restrict__ r, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = xcd_block() * kBlock + threadIdx.x;
  const int i = t / LPR, lane = t % LPR;
  const bool ok = i < n;
  const T s = row_offdiag<LPR, T, false>(ok ? i : 0, lane, ok ? n : 1, ok ? w : 0, nbr, anb, x, nullptr,
                                         nullptr, T(0));
  if (ok && lane == 0) r[i] = b[i] - ac[i] * x[i] - s;
}

// ---- coarse levels in row-major CSR (no padding; a row's entries are
// contiguous, so the LPR lanes of a row read one segment)
template <int LPR, class T, bool PRO>
__device__ __forceinline__ T csr_offdiag(int i, int lane, const int *__restrict__ rp, const int *__restrict__ col,
                                         const T *__restrict__ val, const T *x, const int *__restrict__ parent,
                                         const T *__restrict__ xc, T alpha, bool ok) {
  T s = T(0);
  if (ok)
    for (int q = rp[i] + lane; q < rp[i + 1]; q += LPR) {
      const int k = col[q];
      T xk = x[k];
      if (PRO) xk = xk + alpha * xc[parent[k]];
      s += val[q] * xk;
    }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
  return s;
}

template <int LPR, class T, class TO, bool PRO>
__global__ void __launch_bounds__(kBlock)
k_sweep_csr(int n, const int *__restrict__ rp, const int *__restrict__ col, const T *__restrict__ val,
            const T *__restrict__ ac, const T *__restrict__ b, const T *x, const int *__restrict__ parent,
            const T *__restrict__ xc, T alpha, T omega, TO *xn, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = xcd_block() * kBlock + threadIdx.x;
  const int i = t / LPR, lane = t % LPR;
  const bool ok = i < n;
  const T s = csr_offdiag<LPR, T, PRO>(ok ? i : 0, lane, rp, col, val, x, parent, xc, alpha, ok);
  if (ok && lane == 0) {
    T xi = x[i];
    if (PRO) xi = xi + alpha * xc[parent[i]];
    xn[i] = (TO)(xi + omega * (b[i] - ac[i] * xi - s) / ac[i]);
  }
}

// ---- large coarse levels in sliced ELL: slices of 64 rows (one wavefront),
// one row per lane, a slice's entries column-major (entry k of the slice's
// lane l at ptr[s] + 64 k + l, so a wavefront's loads of entry k coalesce),
// padded to the slice's longest row with zero values whose column is the row
// itself.  row[] maps a slot to its row (-1: past the end); rows are sorted by
// length within windows of 256 rows so that a slice's rows have similar
// lengths.  Each lane sums its row's entries in CSR order.
template <class T, bool PRO>
__device__ __forceinline__ T sell_offdiag(int slot, int lane, const int *__restrict__ ptr, const int *__restrict__ col,
                                          const T *__restrict__ val, const T *x, const int *__restrict__ parent,
                                          const T *__restrict__ xc, T alpha) {
  const int sl = slot >> 6, q0 = ptr[sl] + lane, len = (ptr[sl + 1] - ptr[sl]) >> 6;
  constexpr int B = 4; // entries per batch: columns first, then values and gathers
  T s = T(0);
  for (int k0 = 0; k0 < len; k0 += B) {
    int c[B];
    T v[B], xv[B];
#pragma unroll
    for (int u = 0; u < B; ++u) c[u] = col[q0 + (min(k0 + u, len - 1) << 6)];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      v[u] = val[q0 + (min(k0 + u, len - 1) << 6)];
      xv[u] = x[c[u]];
      if (PRO) xv[u] = xv[u] + alpha * xc[parent[c[u]]];
    }
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (k0 + u < len) s += v[u] * xv[u];
  }
  return s;
}

template <class T, class TO, bool PRO>
__global__ void __launch_bounds__(kBlock)
k_sweep_sell(int nslots, const int *__restrict__ ptr, const int *__restrict__ row, const int *__restrict__ col,
             const T *__restrict__ val, const T *__restrict__ ac, const T *__restrict__ b, const T *x,
             const int *__restrict__ parent, const T *__restrict__ xc, T alpha, T omega, TO *xn,
             const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = xcd_block() * kBlock + threadIdx.x;
  if (t >= nslots) return;
  const int i = row[t];
  if (i < 0) return;
  const T s = sell_offdiag<T, PRO>(t, t & 63, ptr, col, val, x, parent, xc, alpha);
  T xi = x[i];
  if (PRO) xi = xi + alpha * xc[parent[i]];
  xn[i] = (TO)(xi + omega * (b[i] - ac[i] * xi - s) / ac[i]);
}

// r = b - A x on a sliced-ELL level
template <class T>
__global__ void __launch_bounds__(kBlock)
k_resid_sell(int nslots, const int *__restrict__ ptr, const int *__restrict__ row, const int *__restrict__ col,
             const T *__restrict__ val, const T *__restrict__ ac, const T *__restrict__ b, const T *__restrict__ x,
             T *__restrict__ r, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = xcd_block() * kBlock + threadIdx.x;
  if (t >= nslots) return;
  const int i = row[t];
  if (i < 0) return;
  const T s = sell_offdiag<T, false>(t, t & 63, ptr, col, val, x, nullptr, nullptr, T(0));
  r[i] = b[i] - ac[i] * x[i] - s;
}

// Residual of a coarse level summed straight into the next level's
// right-hand side: b_c[p] = sum over the members i of p of (b_i - A_i x);
// the coarse level's first sweep as k_restrict (mode 1: x_c = omega
// b_c/ac_c, 2: x_c = 0, 0: untouched).  Each coarse cell has MPR groups of
// LPR lanes (all in one wavefront): group g works through members g, g+MPR,
// ... in increasing order, its LPR lanes through each member row together,
// and the groups' sums are combined in a fixed butterfly, so the result is
// deterministic and a coarse cell waits on ceil(members / MPR) member rows
// instead of all of them.
template <int LPR, int MPR, class T>
__global__ void __launch_bounds__(kBlock)
k_resid_restrict_csr(int nc, const int *__restrict__ mptr, const int *__restrict__ mem, const int *__restrict__ rp,
                     const int *__restrict__ col, const T *__restrict__ val, const T *__restrict__ ac,
                     const T *__restrict__ b, const T *__restrict__ x, T *__restrict__ bc, const T *__restrict__ acc,
                     T *__restrict__ xc, T omega, int mode, const int *__restrict__ flags) {
  constexpr int G = LPR * MPR;
  static_assert(64 % G == 0 && kBlock % 64 == 0, "a coarse cell's lanes must share a wavefront");
  if (flags && flags[0]) return;
  const int t = xcd_block() * kBlock + threadIdx.x;
  const int p = t / G, lane = t % LPR, grp = (t % G) / LPR;
  const bool ok = p < nc;
  const int q0 = ok ? mptr[p] : 0, q1 = ok ? mptr[p + 1] : 0;
  T sum = T(0);
  for (int q = q0 + grp; q < q1; q += MPR) { // the same trip count on the LPR lanes of a group
    const int i = mem[q];
    const T s = csr_offdiag<LPR, T, false>(i, lane, rp, col, val, x, nullptr, nullptr, T(0), true);
    sum = sum + (b[i] - ac[i] * x[i] - s);
  }
#pragma unroll
  for (int off = LPR; off < G; off <<= 1) sum += __shfl_xor(sum, off, G);
  if (ok && lane == 0 && grp == 0) {
    bc[p] = sum;
    if (mode == 1) xc[p] = omega * sum / acc[p];
    else if (mode == 2) xc[p] = T(0);
  }
}

// Entry of the cycle on the finest level: b = (T)rhs (when the hierarchy is
// FP32; b may be null in FP64) and the first sweep from a zero start,
// x = omega*b/ac, or x = 0 when no pre-smoothing sweep is asked for.
template <class T>
__global__ void k_enter(int n, const T *__restrict__ ac, const double *__restrict__ rhs, T *__restrict__ b,
                        T *__restrict__ x, T omega, int sweep, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= n) return;
  const T bi = (T)rhs[i];
  if (b) b[i] = bi;
  x[i] = sweep ? omega * bi / ac[i] : T(0);
}

// Entry of the cycle fused with the residual of its first sweep, for one
// pre-sweep on the finest level: b = (T)rhs, x = omega*b/ac, r = b - A x,
// where the neighbours' x_j = omega*(T)rhs_j/ac_j are formed from their
// right-hand sides (x is pointwise in b), so no second pass is needed.  The
// values are those of k_enter followed by k_resid.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_enter_resid(int n, int w, EllNbr nbr, const T *__restrict__ ac, const T *__restrict__ anb,
              const double *__restrict__ rhs, T *__restrict__ b, T *__restrict__ x, T *__restrict__ r, T omega,
              const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= n) return;
  const T bi = (T)rhs[i];
  if (b) b[i] = bi;
  const T xi = omega * bi / ac[i];
  x[i] = xi;
  T s = T(0);
  for (int j0 = 0; j0 < w; j0 += kEllBatch) {
    int k[kEllBatch];
    T a[kEllBatch], xv[kEllBatch];
    ell_batch<kEllBatch, kStreamLevel>(nbr, i, n, w, j0, k);
#pragma unroll
    for (int u = 0; u < kEllBatch; ++u) {
      const int kk = k[u] >= 0 ? k[u] : i;
      a[u] = ld_mat<kStreamLevel>(&anb[(size_t)min(j0 + u, w - 1) * n + i]);
      xv[u] = omega * (T)rhs[kk] / ac[kk];
    }
#pragma unroll
    for (int u = 0; u < kEllBatch; ++u)
      if (j0 + u < w && k[u] >= 0) s += a[u] * xv[u];
  }
  r[i] = bi - ac[i] * xi - s;
}

// Restriction b_c[p] = sum of the member residuals in increasing fine-cell
// order; mode 1 also takes the coarse level's first sweep from a zero start
// (x_c = omega*b_c/ac_c), mode 2 sets x_c = 0, mode 0 leaves x_c alone.
template <class T>
__global__ void k_restrict(int nc, const int *__restrict__ mptr, const int *__restrict__ mem,
                           const T *__restrict__ rf, T *__restrict__ bc, const T *__restrict__ acc,
                           T *__restrict__ xc, T omega, int mode, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nc) return;
  const int q0 = mptr[p], q1 = mptr[p + 1];
  T s = T(0);
  // members in batches: their indices, then their residuals, then the sum
  // in member order
  constexpr int B = 8;
  for (int q = q0; q < q1; q += B) {
    int id[B];
    T v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) id[u] = mem[min(q + u, q1 - 1)];
#pragma unroll
    for (int u = 0; u < B; ++u) v[u] = rf[id[u]];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (q + u < q1) s = s + v[u];
  }
  bc[p] = s;
  if (mode == 1) xc[p] = omega * s / acc[p];
  else if (mode == 2) xc[p] = T(0);
}

// out = x + alpha*xc[parent] (out may be x)
template <class T, class TO>
__global__ void k_prolong(int nf, const int *__restrict__ parent, const T *__restrict__ xc, const T *x,
                          T alpha, TO *out, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < nf) out[i] = (TO)(x[i] + alpha * xc[parent[i]]);
}

template <class T, class TO>
__global__ void k_convert(int n, const T *__restrict__ x, TO *__restrict__ out, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (TO)x[i];
}

template <class T>
__global__ void k_zero(int n, T *__restrict__ x, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = T(0);
}

// one colour of a multicolour Gauss-Seidel sweep, in place
template <class T>
__global__ void k_gsc(int cnt, const int *__restrict__ cells, int n, int w, const int *__restrict__ nbr,
                      const T *__restrict__ ac, const T *__restrict__ anb, const T *__restrict__ b, T *x,
                      const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int i = cells[t];
  const T s = ell_fold<kEllBatch>(i, n, w, nbr, anb, (const T *)x, b[i], [](T acc, T a, T xk) { return acc - a * xk; });
  x[i] = s / ac[i];
}

// AMG-as-solver residual: r = b - A x and partial L1(r)
__global__ void k_amg_res(int n, int w, const int *__restrict__ nbr, const double *__restrict__ ac,
                          const double *__restrict__ anb, const double *__restrict__ b,
                          const double *__restrict__ x, double *__restrict__ r, double *__restrict__ part,
                          const int *__restrict__ flags) {
  if (flags[0]) return;
  double s = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double v = ell_fold<kEllBatch>(i, n, w, nbr, anb, x, b[i] - ac[i] * x[i],
                                         [](double acc, double a, double xk) { return acc - a * xk; });
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

// ------------------------------------------------------------ small levels
// Every level from the first one below the tail threshold down to the
// coarsest runs in ONE workgroup in one launch, with the levels' operators
// held in LDS.  A byte image of the tail is kept in device memory: the static
// part (CSR columns, row pointers, parents, aggregate members as uint16) is
// written at setup, the coefficients (ac and the CSR values in the
// hierarchy's precision) are packed after each Galerkin update, and the
// coarsest level's FP64 inverse is written there too.  The kernel copies the
// image into LDS with 16-byte loads, runs restriction, smoothing, the
// coarsest solve and the prolongations between workgroup barriers, and
// writes the top tail level's correction back to device memory.
constexpr int kTailThreads = 1024;
constexpr int kTailMaxLev = 8;
constexpr int kTailDirectMax = 64; // coarsest levels up to this size are solved directly

struct TailLvl {
  int n, nnz, lpr;
  int o_ac, o_val, o_col, o_rp, o_par, o_mp, o_mem; // byte offsets of the image sections
  int o_x, o_t, o_r, o_b;                           // byte offsets of the LDS-only vectors
};
struct TailDesc {
  int nlev, img_bytes, lds_bytes, direct, o_g;
  TailLvl L[kTailMaxLev];
};

// one pass over the rows of a tail level: out = b - A x (jac false) or
// x + omega*(b - A x)/ac (jac true); lpr lanes per row, all threads take part
template <class T>
__device__ __forceinline__ void tail_pass(unsigned char *lds, const TailLvl &L, const T *x, T *out, T omega, bool jac) {
  const T *ac = (const T *)(lds + L.o_ac), *val = (const T *)(lds + L.o_val), *b = (const T *)(lds + L.o_b);
  const unsigned short *col = (const unsigned short *)(lds + L.o_col);
  const unsigned short *rp = (const unsigned short *)(lds + L.o_rp);
  const int lpr = L.lpr, lane = threadIdx.x % lpr, rows = blockDim.x / lpr;
  for (int base = 0; base < L.n; base += rows) {
    const int i = base + threadIdx.x / lpr;
    T s = T(0);
    if (i < L.n)
      for (int q = rp[i] + lane; q < rp[i + 1]; q += lpr) s += val[q] * x[col[q]];
    for (int off = lpr / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, lpr);
    if (i < L.n && lane == 0) {
      const T r = b[i] - ac[i] * x[i] - s;
      out[i] = jac ? x[i] + omega * r / ac[i] : r;
    }
  }
}

// damped-Jacobi sweeps on a tail level; the result ends in x
template <class T> __device__ __forceinline__ void tail_smooth(unsigned char *lds, const TailLvl &L, int sweeps, bool zero, T omega) {
  T *x = (T *)(lds + L.o_x), *t = (T *)(lds + L.o_t);
  const T *ac = (const T *)(lds + L.o_ac), *b = (const T *)(lds + L.o_b);
  int sw = 0;
  if (zero) {
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) x[i] = sweeps > 0 ? omega * b[i] / ac[i] : T(0);
    __syncthreads();
    sw = 1;
  }
  T *src = x, *dst = t;
  for (; sw < sweeps; ++sw) {
    tail_pass<T>(lds, L, src, dst, omega, true);
    __syncthreads();
    T *tmp = src;
    src = dst;
    dst = tmp;
  }
  if (src != x) {
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) x[i] = src[i];
    __syncthreads();
  }
}

template <class T>
__global__ void __launch_bounds__(kTailThreads)
k_tail(const TailDesc *__restrict__ d, const unsigned char *__restrict__ img, const T *__restrict__ b_top,
       T *__restrict__ x_top, int pre, int post, int coarse, T omega, T alpha, const int *__restrict__ flags) {
  if (flags && flags[0]) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const int nlev = d->nlev;
  {
    const uint4 *src = (const uint4 *)img;
    uint4 *dst = (uint4 *)lds;
    const int nv = d->img_bytes / 16;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) dst[i] = src[i];
    const TailLvl &L = d->L[0];
    T *b = (T *)(lds + L.o_b);
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) b[i] = b_top[i];
  }
  __syncthreads();
  for (int l = 0; l + 1 < nlev; ++l) {
    const TailLvl &L = d->L[l];
    tail_smooth<T>(lds, L, pre, true, omega);
    tail_pass<T>(lds, L, (const T *)(lds + L.o_x), (T *)(lds + L.o_r), omega, false);
    __syncthreads();
    const TailLvl &C = d->L[l + 1];
    const T *r = (const T *)(lds + L.o_r);
    T *bc = (T *)(lds + C.o_b);
    const unsigned short *mp = (const unsigned short *)(lds + C.o_mp);
    const unsigned short *mem = (const unsigned short *)(lds + C.o_mem);
    for (int p = threadIdx.x; p < C.n; p += blockDim.x) {
      T s = T(0);
      for (int q = mp[p]; q < mp[p + 1]; ++q) s = s + r[mem[q]];
      bc[p] = s;
    }
    __syncthreads();
  }
  const TailLvl &Lc = d->L[nlev - 1];
  if (d->direct) {
    // coarsest level: x = G b with the FP64 (pseudo-)inverse G
    const double *G = (const double *)(lds + d->o_g);
    const T *b = (const T *)(lds + Lc.o_b);
    T *x = (T *)(lds + Lc.o_x);
    for (int i = threadIdx.x; i < Lc.n; i += blockDim.x) {
      double s = 0.0;
      for (int j = 0; j < Lc.n; ++j) s += G[(size_t)i * Lc.n + j] * (double)b[j];
      x[i] = (T)s;
    }
    __syncthreads();
  } else {
    tail_smooth<T>(lds, Lc, coarse, true, omega);
  }
  for (int l = nlev - 2; l >= 0; --l) {
    const TailLvl &L = d->L[l];
    T *x = (T *)(lds + L.o_x);
    const T *xc = (const T *)(lds + d->L[l + 1].o_x);
    const unsigned short *par = (const unsigned short *)(lds + L.o_par);
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) x[i] = x[i] + alpha * xc[par[i]];
    __syncthreads();
    tail_smooth<T>(lds, L, post, false, omega);
  }
  const TailLvl &L0 = d->L[0];
  const T *x0 = (const T *)(lds + L0.o_x);
  for (int i = threadIdx.x; i < L0.n; i += blockDim.x) x_top[i] = x0[i];
}

// coefficients of one tail level into its image sections (after Galerkin)
template <class T>
__global__ void k_tail_pack(int n, int nnz, const double *__restrict__ ac, const double *__restrict__ anb,
                            const int *__restrict__ lin, T *__restrict__ img_ac, T *__restrict__ img_val) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < n) img_ac[t] = (T)ac[t];
  if (t < nnz) img_val[t] = (T)anb[lin[t]];
}

// FP64 inverse of the coarsest operator by Gauss-Jordan elimination without
// pivoting (the coarse operators are diagonally dominant), one workgroup,
// n <= kTailDirectMax.  A singular operator with zero row sums (the pure
// Neumann p' problem, null space = constants) is solved with its last
// unknown pinned to zero: G is the inverse of the leading (n-1)x(n-1) block,
// bordered by zeros, a symmetric generalised inverse.  A zero or non-finite
// pivot leaves G = 0 (no coarse correction) rather than garbage.
constexpr int kInvThreads = 256;
__global__ void __launch_bounds__(kInvThreads)
k_coarse_inverse(int n, int w, const int *__restrict__ nbr, const double *__restrict__ ac,
                 const double *__restrict__ anb, double *__restrict__ G) {
  __shared__ double M[kTailDirectMax * kTailDirectMax];
  __shared__ double rowk[kTailDirectMax], colk[kTailDirectMax];
  __shared__ int s_m, s_fail;
  const int nn = n * n;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const int i = e / n, k = e % n;
    M[e] = (i == k) ? ac[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // off-diagonals in slot order; singularity test on the row sums
    bool zero_rows = true;
    for (int i = 0; i < n; ++i) {
      double rs = ac[i];
      for (int j = 0; j < w; ++j) {
        const int k = nbr[(size_t)j * n + i];
        if (k >= 0) {
          M[i * n + k] += anb[(size_t)j * n + i];
          rs += anb[(size_t)j * n + i];
        }
      }
      if (!(fabs(rs) <= 1.0e-9 * fabs(ac[i]))) zero_rows = false;
    }
    s_m = zero_rows ? n - 1 : n;
    s_fail = 0;
  }
  __syncthreads();
  const int m = s_m;
  for (int k = 0; k < m; ++k) {
    const double p = M[k * n + k];
    if (!(fabs(p) > 0.0) || !isfinite(p)) {
      if (threadIdx.x == 0) s_fail = 1;
      break; // p is read from LDS by every thread: the branch is uniform
    }
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      rowk[j] = (j == k) ? 1.0 / p : M[k * n + j] / p;
      colk[j] = M[j * n + k];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = e / m, j = e % m;
      if (i == k) M[k * n + j] = rowk[j];
      else M[i * n + j] = ((j == k) ? 0.0 : M[i * n + j]) - colk[i] * rowk[j];
    }
    __syncthreads();
  }
  __syncthreads();
  const bool fail = s_fail != 0;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    const int i = e / n, j = e % n;
    G[e] = (!fail && i < m && j < m) ? M[i * n + j] : 0.0;
  }
}

} // namespace

// ------------------------------------------------------------ hierarchy
struct DLevel {
  int n = 0, w = 0;
  const int *nbr = nullptr;      // device, 0-based/-1
  const int *nbr16 = nullptr;    // level 0: the mesh's packed 16-bit offsets, if any
  DevBuf<int> nbr_own;
  const double *ac64 = nullptr, *anb64 = nullptr;
  DevBuf<double> ac64_own, anb64_own;
  DevBuf<float> ac32, anb32;
  DevBuf<double> x64, b64, r64, t64;
  DevBuf<float> x32, b32, r32, t32;
  std::vector<int> color_ptr;
  DevBuf<int> color_cells;
  DevBuf<int> parent;            // 0-based parent in the next level
  DevBuf<int> mptr, mem;         // members of this level's aggregates (coarse)
  std::vector<int> h_mptr, h_mem;
  DevBuf<int> gptr;
  DevBuf<int> gsrc;               // 32-bit: fine n + n w stays below 2^31
  int gne = 0;
  // levels >= 1: row-major CSR copy of the off-diagonals for the grid-wide
  // Jacobi kernels (values written by k_galerkin through csr_pos)
  long long nnz = 0;
  DevBuf<int> csr_rp, csr_col, csr_pos;
  DevBuf<double> val64;
  DevBuf<float> val32;
  // large levels: sliced ELL instead (val64/val32 then hold the sliced
  // layout and csr_pos points into it)
  bool sell = false;
  int sell_slots = 0;
  DevBuf<int> sell_ptr, sell_row, sell_col;
};

constexpr int kSellWindow = 256; // rows sorted by length within windows of this size

// Sliced-ELL image of a CSR level: slot order, slice offsets, columns, and
// the map from CSR entry to sliced entry.
static void build_sell(int n, const std::vector<int> &rp, const std::vector<int> &col, std::vector<int> &ptr,
                       std::vector<int> &row, std::vector<int> &scol, std::vector<int> &csr2sell) {
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int w0 = 0; w0 < n; w0 += kSellWindow) {
    const int w1 = std::min(n, w0 + kSellWindow);
    std::stable_sort(order.begin() + w0, order.begin() + w1,
                     [&](int a, int b) { return rp[a + 1] - rp[a] > rp[b + 1] - rp[b]; });
  }
  const int ns = (n + 63) / 64;
  row.assign((size_t)ns * 64, -1);
  ptr.assign(ns + 1, 0);
  for (int s = 0; s < ns; ++s) {
    int len = 0;
    for (int l = 0; l < 64 && s * 64 + l < n; ++l) {
      const int r = order[s * 64 + l];
      row[(size_t)s * 64 + l] = r;
      len = std::max(len, rp[r + 1] - rp[r]);
    }
    ptr[s + 1] = ptr[s] + 64 * len;
  }
  scol.assign(std::max(ptr[ns], 1), 0);
  csr2sell.assign(std::max<size_t>(col.size(), 1), -1);
  for (int s = 0; s < ns; ++s) {
    const int len = (ptr[s + 1] - ptr[s]) / 64;
    for (int l = 0; l < 64; ++l) {
      const int r = row[(size_t)s * 64 + l];
      for (int k = 0; k < len; ++k) {
        const int q = ptr[s] + 64 * k + l;
        if (r < 0) scol[q] = 0;
        else if (k < rp[r + 1] - rp[r]) {
          scol[q] = col[rp[r] + k];
          csr2sell[rp[r] + k] = q;
        } else {
          scol[q] = r; // padding: zero value times the row's own x
        }
      }
    }
  }
}

// The tail of one precision: its first level, descriptor and byte image.
struct TailImage {
  int ts = -1;                   // first tail level; == levels when there is no tail
  TailDesc desc{};
  DevBuf<unsigned char> img;
  DevBuf<TailDesc> ddesc;
  DevBuf<int> lin;               // CSR entry -> ELL linear index, all tail levels
  std::vector<int> lin_off;      // per tail level
};

struct Hierarchy {
  std::vector<HostLevel> host;   // maps for export / tests
  std::vector<DLevel> lev;
  int calls = 0;
  bool fp32 = false;
  TailImage tail32, tail64;
};

static size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

static int lds_budget() {
  static int budget = -1;
  if (budget < 0) {
    int dev = 0, v = 0;
    budget = 64 * 1024;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      budget = v;
    budget = std::min(budget, 160 * 1024);
    // the tail kernels ask for more than the default dynamic LDS window
    if (hipFuncSetAttribute((const void *)k_tail<float>, hipFuncAttributeMaxDynamicSharedMemorySize, budget) !=
            hipSuccess ||
        hipFuncSetAttribute((const void *)k_tail<double>, hipFuncAttributeMaxDynamicSharedMemorySize, budget) !=
            hipSuccess) {
      (void)hipGetLastError();
      budget = 64 * 1024;
    }
  }
  return budget;
}

// lanes per row for a tail level: fewest dependent steps per sweep
static int tail_lpr(int n, int nnz) {
  int best = 1;
  double best_cost = 1e30;
  const double avg = n > 0 ? (double)nnz / n : 0.0;
  for (int lpr = 1; lpr <= 16; lpr *= 2) {
    const int rows = kTailThreads / lpr;
    const double passes = std::ceil((double)n / rows);
    int lg = 0;
    while ((1 << lg) < lpr) ++lg;
    const double cost = passes * (std::ceil(avg / lpr) + lg);
    if (cost < best_cost) {
      best_cost = cost;
      best = lpr;
    }
  }
  return best;
}

// Lays out the tail of precision tsz (4 or 8 bytes) starting at the first
// level at or below mg_tail_cells; moves the start down while the layout
// does not fit the LDS budget, the level count or the uint16 index range.
static void build_tail(fs_mesh *m, Hierarchy &H, TailImage &ti, size_t tsz) {
  hipStream_t s = m->stream;
  const int nl = (int)H.lev.size();
  ti.ts = nl;
  if (nl < 2) return;
  int ts = nl;
  for (int l = 1; l < nl; ++l)
    if (H.lev[l].n <= m->mg_tail_cells) {
      ts = l;
      break;
    }
  const size_t budget = (size_t)lds_budget();
  TailDesc d{};
  for (; ts < nl; ++ts) {
    const int nt = nl - ts;
    if (nt > kTailMaxLev) continue;
    d = TailDesc{};
    d.nlev = nt;
    size_t off = 0;
    bool ok = true;
    for (int q = 0; q < nt; ++q) {
      const HostLevel &h = H.host[ts + q];
      int nnz = 0;
      for (int i = 0; i < h.n; ++i)
        for (int j = 0; j < h.nnbr[i]; ++j)
          if (h.nbr[(size_t)j * h.n + i] > 0) ++nnz;
      if (h.n >= 65535 || nnz >= 65535 || (q > 0 && H.host[ts + q - 1].n >= 65535)) ok = false;
      TailLvl &L = d.L[q];
      L.n = h.n;
      L.nnz = nnz;
      L.lpr = tail_lpr(h.n, nnz);
      L.o_ac = (int)off; off += align16(tsz * h.n);
      L.o_val = (int)off; off += align16(tsz * std::max(nnz, 1));
      L.o_col = (int)off; off += align16(2 * std::max(nnz, 1));
      L.o_rp = (int)off; off += align16(2 * (h.n + 1));
      L.o_par = (int)off; if (q + 1 < nt) off += align16(2 * h.n);
      L.o_mp = (int)off; if (q > 0) off += align16(2 * (h.n + 1));
      L.o_mem = (int)off; if (q > 0) off += align16(2 * H.host[ts + q - 1].n);
    }
    const int nc = d.L[nt - 1].n;
    d.direct = nc <= kTailDirectMax ? 1 : 0;
    d.o_g = (int)off;
    if (d.direct) off += align16(sizeof(double) * nc * nc);
    d.img_bytes = (int)off;
    for (int q = 0; q < nt; ++q) {
      TailLvl &L = d.L[q];
      L.o_x = (int)off; off += align16(tsz * L.n);
      L.o_t = (int)off; off += align16(tsz * L.n);
      L.o_r = (int)off; off += align16(tsz * L.n);
      L.o_b = (int)off; off += align16(tsz * L.n);
    }
    d.lds_bytes = (int)off;
    if (ok && off <= budget) break;
  }
  ti.ts = ts;
  if (ts >= nl) return;
  // static sections of the image
  std::vector<unsigned char> img(d.img_bytes, 0);
  std::vector<int> lin;
  ti.lin_off.assign(d.nlev, 0);
  for (int q = 0; q < d.nlev; ++q) {
    const HostLevel &h = H.host[ts + q];
    const TailLvl &L = d.L[q];
    unsigned short *col = (unsigned short *)(img.data() + L.o_col);
    unsigned short *rp = (unsigned short *)(img.data() + L.o_rp);
    ti.lin_off[q] = (int)lin.size();
    int e = 0;
    for (int i = 0; i < h.n; ++i) {
      rp[i] = (unsigned short)e;
      for (int j = 0; j < h.nnbr[i]; ++j) {
        const int k = h.nbr[(size_t)j * h.n + i];
        if (k > 0) {
          col[e++] = (unsigned short)(k - 1);
          lin.push_back(j * h.n + i);
        }
      }
    }
    rp[h.n] = (unsigned short)e;
    if (q + 1 < d.nlev) {
      unsigned short *par = (unsigned short *)(img.data() + L.o_par);
      for (int i = 0; i < h.n; ++i) par[i] = (unsigned short)(h.parent[i] - 1);
    }
    if (q > 0) {
      const DLevel &D = H.lev[ts + q];
      unsigned short *mp = (unsigned short *)(img.data() + L.o_mp);
      unsigned short *mem = (unsigned short *)(img.data() + L.o_mem);
      for (int p = 0; p <= h.n; ++p) mp[p] = (unsigned short)D.h_mptr[p];
      for (size_t t = 0; t < D.h_mem.size(); ++t) mem[t] = (unsigned short)D.h_mem[t];
    }
  }
  if (lin.empty()) lin.push_back(0);
  ti.desc = d;
  ti.img.upload(img, s);
  ti.ddesc.upload(&ti.desc, 1, s);
  ti.lin.upload(lin, s);
}

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
    if (m->mg_agg == 0) aggregate(H->host[nl - 1], c, nl + 1);
    else aggregate_pairwise(H->host[nl - 1], c, m->mg_agg);
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
      D.nbr16 = m->d_nbr16.p;
    } else {
      D.nbr_own.upload(nbr0, s);
      D.nbr = D.nbr_own.p;
      D.ac64_own.alloc(h.n);
      D.anb64_own.alloc((size_t)h.n * h.w);
      D.ac64 = D.ac64_own.p;
      D.anb64 = D.anb64_own.p;
      D.gptr.upload(gptrs[l], s);
      {
        std::vector<int> g32(gsrcs[l].size());
        for (size_t t = 0; t < g32.size(); ++t) {
          if (gsrcs[l][t] > INT32_MAX) throw std::runtime_error("AMG level too large for 32-bit Galerkin sources");
          g32[t] = (int)gsrcs[l][t];
        }
        D.gsrc.upload(g32, s);
      }
      D.gne = (int)gptrs[l].size() - 1;
      // members in increasing fine order
      const HostLevel &f = H->host[l - 1];
      D.h_mptr.assign(h.n + 1, 0);
      D.h_mem.assign(f.n, 0);
      for (int i = 0; i < f.n; ++i) D.h_mptr[f.parent[i]]++;
      for (int p = 0; p < h.n; ++p) D.h_mptr[p + 1] += D.h_mptr[p];
      std::vector<int> fill(D.h_mptr.begin(), D.h_mptr.end() - 1);
      for (int i = 0; i < f.n; ++i) D.h_mem[fill[f.parent[i] - 1]++] = i;
      D.mptr.upload(D.h_mptr, s);
      D.mem.upload(D.h_mem, s);
      std::vector<int> rp(h.n + 1, 0), col, pos((size_t)h.n * h.w, -1);
      for (int i = 0; i < h.n; ++i) {
        for (int j = 0; j < h.w; ++j) {
          const int k = nbr0[(size_t)j * h.n + i];
          if (k < 0) continue;
          pos[(size_t)j * h.n + i] = (int)col.size();
          col.push_back(k);
        }
        rp[i + 1] = (int)col.size();
      }
      D.nnz = (long long)col.size();
      // a level that runs grid-wide (larger than the tail) and has at least
      // mg_sell_rows rows keeps its operator in sliced ELL
      D.sell = m->mg_sell_rows > 0 && h.n >= m->mg_sell_rows && h.n > m->mg_tail_cells;
      size_t nval = std::max<size_t>(col.size(), 1);
      if (D.sell) {
        std::vector<int> sptr, srow, scol, c2s;
        build_sell(h.n, rp, col, sptr, srow, scol, c2s);
        for (auto &q : pos)
          if (q >= 0) q = c2s[q];
        D.sell_slots = (int)srow.size();
        D.sell_ptr.upload(sptr, s);
        D.sell_row.upload(srow, s);
        D.sell_col.upload(scol, s);
        nval = scol.size();
      }
      if (col.empty()) col.push_back(0);
      D.csr_rp.upload(rp, s);
      D.csr_col.upload(col, s);
      D.csr_pos.upload(pos, s);
      D.val64.alloc(nval);
      D.val32.alloc(nval);
      // padded sliced entries are never written by the Galerkin kernel
      FS_HIP_THROW(hipMemsetAsync(D.val64.p, 0, sizeof(double) * nval, s));
      FS_HIP_THROW(hipMemsetAsync(D.val32.p, 0, sizeof(float) * nval, s));
    }
    if (l + 1 < nl) {
      std::vector<int> par0(h.n);
      for (int i = 0; i < h.n; ++i) par0[i] = h.parent[i] - 1;
      D.parent.upload(par0, s);
    }
    std::vector<int> cells;
    greedy_colors(h.n, h.w, nbr0.data(), D.color_ptr, cells);
    D.color_cells.upload(cells, s);
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
  build_tail(m, *H, H->tail32, sizeof(float));
  build_tail(m, *H, H->tail64, sizeof(double));
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
                       H.fp32 ? C.ac32.p : nullptr, H.fp32 ? C.anb32.p : nullptr, C.csr_pos.p, C.val64.p,
                       H.fp32 ? C.val32.p : nullptr);
  }
  // the tail of the precision in use: coefficients and the coarsest inverse
  TailImage &ti = H.fp32 ? H.tail32 : H.tail64;
  const int nl = (int)H.lev.size();
  if (ti.ts < nl) {
    const TailDesc &d = ti.desc;
    for (int q = 0; q < d.nlev; ++q) {
      DLevel &D = H.lev[ti.ts + q];
      const TailLvl &L = d.L[q];
      const int cnt = std::max(L.n, L.nnz);
      unsigned char *base = ti.img.p;
      if (H.fp32)
        hipLaunchKernelGGL(k_tail_pack<float>, dim3(gfor(cnt)), dim3(kBlock), 0, s, L.n, L.nnz, D.ac64, D.anb64,
                           ti.lin.p + ti.lin_off[q], (float *)(base + L.o_ac), (float *)(base + L.o_val));
      else
        hipLaunchKernelGGL(k_tail_pack<double>, dim3(gfor(cnt)), dim3(kBlock), 0, s, L.n, L.nnz, D.ac64,
                           D.anb64, ti.lin.p + ti.lin_off[q], (double *)(base + L.o_ac),
                           (double *)(base + L.o_val));
    }
    if (d.direct) {
      DLevel &Dc = H.lev[nl - 1];
      hipLaunchKernelGGL(k_coarse_inverse, dim3(1), dim3(kInvThreads), 0, s, Dc.n, Dc.w, Dc.nbr, Dc.ac64,
                         Dc.anb64, (double *)(ti.img.p + d.o_g));
    }
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
  static const T *val(DLevel &d);
  static TailImage &tail(Hierarchy &h);
};
template <> struct LV<double> {
  static double *x(DLevel &d) { return d.x64.p; }
  static double *b(DLevel &d) { return d.b64.p; }
  static double *r(DLevel &d) { return d.r64.p; }
  static double *t(DLevel &d) { return d.t64.p; }
  static const double *ac(DLevel &d) { return d.ac64; }
  static const double *anb(DLevel &d) { return d.anb64; }
  static const double *val(DLevel &d) { return d.val64.p; }
  static TailImage &tail(Hierarchy &h) { return h.tail64; }
};
template <> struct LV<float> {
  static float *x(DLevel &d) { return d.x32.p; }
  static float *b(DLevel &d) { return d.b32.p; }
  static float *r(DLevel &d) { return d.r32.p; }
  static float *t(DLevel &d) { return d.t32.p; }
  static const float *ac(DLevel &d) { return d.ac32.p; }
  static const float *anb(DLevel &d) { return d.anb32.p; }
  static const float *val(DLevel &d) { return d.val32.p; }
  static TailImage &tail(Hierarchy &h) { return h.tail32; }
};

// lanes per row of a coarse level's CSR kernels by its mean row length
static int csr_lpr(const DLevel &D) {
  const double a = D.n ? (double)D.nnz / D.n : 0.0;
  return a <= 6.0 ? 4 : a <= 16.0 ? 8 : 16;
}

// member groups per coarse cell in the fused residual-restriction: up to
// kRestrictGroups, as many as fit one wavefront with LPR lanes each
constexpr int kRestrictGroups = 8;
// rows wider than 16 entries are restricted with 8 lanes per member row and
// 8 member groups rather than 16 lanes and 4 groups
constexpr bool kRestrictNarrow = true;
template <int LPR> constexpr int restrict_groups() { return kRestrictGroups < 64 / LPR ? kRestrictGroups : 64 / LPR; }

// residual of coarse level D restricted into C (plus C's first sweep)
template <class T>
static void launch_resid_restrict(fs_mesh *m, DLevel &D, DLevel &C, const T *b, const T *x, T om, int mode,
                                  const int *flags) {
  hipStream_t s = m->stream;
  const int *rp = D.csr_rp.p, *col = D.csr_col.p;
  const T *val = LV<T>::val(D), *ac = LV<T>::ac(D);
  constexpr int M4 = restrict_groups<4>(), M8 = restrict_groups<8>(), M16 = restrict_groups<16>();
  switch (csr_lpr(D)) {
  case 4:
    hipLaunchKernelGGL((k_resid_restrict_csr<4, M4, T>), dim3(gfor(C.n * 4 * M4)), dim3(kBlock), 0, s, C.n,
                       C.mptr.p, C.mem.p, rp, col, val, ac, b, x, LV<T>::b(C), LV<T>::ac(C), LV<T>::x(C), om, mode,
                       flags);
    break;
  case 8:
    hipLaunchKernelGGL((k_resid_restrict_csr<8, M8, T>), dim3(gfor(C.n * 8 * M8)), dim3(kBlock), 0, s, C.n,
                       C.mptr.p, C.mem.p, rp, col, val, ac, b, x, LV<T>::b(C), LV<T>::ac(C), LV<T>::x(C), om, mode,
                       flags);
    break;
  default:
    if (kRestrictNarrow) {
      hipLaunchKernelGGL((k_resid_restrict_csr<8, M8, T>), dim3(gfor(C.n * 8 * M8)), dim3(kBlock), 0, s, C.n,
                         C.mptr.p, C.mem.p, rp, col, val, ac, b, x, LV<T>::b(C), LV<T>::ac(C), LV<T>::x(C), om, mode,
                         flags);
      break;
    }
    hipLaunchKernelGGL((k_resid_restrict_csr<16, M16, T>), dim3(gfor(C.n * 16 * M16)), dim3(kBlock), 0, s, C.n,
                       C.mptr.p, C.mem.p, rp, col, val, ac, b, x, LV<T>::b(C), LV<T>::ac(C), LV<T>::x(C), om, mode,
                       flags);
  }
}

// one damped-Jacobi sweep on a grid-wide level, lanes per row by row width;
// sweeps on the finest level go through the kernel clock (site
// FS_CLOCK_FINE_SWEEP) with their algorithmic bytes: ac, b, x, the result, the
// neighbour indices and the off-diagonal values, plus parent and the coarse
// correction when the prolongation is fused in
template <class T, class TO, bool PRO>
static void launch_sweep(fs_mesh *m, DLevel &D, int level, const T *b, const T *x, const int *parent, const T *xc,
                         int nc, T alpha, T om, TO *xn, const int *flags) {
  hipStream_t s = m->stream;
  const T *ac = LV<T>::ac(D), *anb = LV<T>::anb(D);
  if (level > 0 && D.sell) {
    hipLaunchKernelGGL((k_sweep_sell<T, TO, PRO>), dim3(gfor(D.sell_slots)), dim3(kBlock), 0, s, D.sell_slots,
                       D.sell_ptr.p, D.sell_row.p, D.sell_col.p, LV<T>::val(D), ac, b, x, parent, xc, alpha, om, xn,
                       flags);
    return;
  }
  if (level > 0) {
    const int *rp = D.csr_rp.p, *col = D.csr_col.p;
    const T *val = LV<T>::val(D);
    switch (csr_lpr(D)) {
    case 4:
      hipLaunchKernelGGL((k_sweep_csr<4, T, TO, PRO>), dim3(gfor(D.n * 4)), dim3(kBlock), 0, s, D.n, rp, col, val,
                         ac, b, x, parent, xc, alpha, om, xn, flags);
      break;
    case 8:
      hipLaunchKernelGGL((k_sweep_csr<8, T, TO, PRO>), dim3(gfor(D.n * 8)), dim3(kBlock), 0, s, D.n, rp, col, val,
                         ac, b, x, parent, xc, alpha, om, xn, flags);
      break;
    default:
      hipLaunchKernelGGL((k_sweep_csr<16, T, TO, PRO>), dim3(gfor(D.n * 16)), dim3(kBlock), 0, s, D.n, rp, col,
                         val, ac, b, x, parent, xc, alpha, om, xn, flags);
    }
    return;
  }
  const int site = level == 0 ? FS_CLOCK_FINE_SWEEP : 0;
  double bytes = 0.0;
  if (site)
    bytes = (double)D.n * (3.0 * sizeof(T) + sizeof(TO)) + (D.nbr16 ? 4.0 * ((D.w + 1) / 2) : 4.0 * D.w) * D.n +
            (double)sizeof(T) * m->nnz + (PRO ? 4.0 * D.n + (double)sizeof(T) * nc : 0.0);
  const EllNbr nb(D.nbr, D.nbr16);
  if (D.w <= 8)
    timed_launch(m, site, bytes, k_sweep<1, T, TO, PRO>, dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.w, nb, ac, anb,
                 b, x, parent, xc, alpha, om, xn, flags);
  else if (D.w <= 16)
    timed_launch(m, site, bytes, k_sweep<4, T, TO, PRO>, dim3(gfor(D.n * 4)), dim3(kBlock), 0, s, D.n, D.w, nb, ac,
                 anb, b, x, parent, xc, alpha, om, xn, flags);
  else if (D.w <= 32)
    timed_launch(m, site, bytes, k_sweep<8, T, TO, PRO>, dim3(gfor(D.n * 8)), dim3(kBlock), 0, s, D.n, D.w, nb, ac,
                 anb, b, x, parent, xc, alpha, om, xn, flags);
  else
    timed_launch(m, site, bytes, k_sweep<16, T, TO, PRO>, dim3(gfor(D.n * 16)), dim3(kBlock), 0, s, D.n, D.w,
                 D.nbr, ac, anb, b, x, parent, xc, alpha, om, xn, flags);
}

template <class T>
static void launch_resid(fs_mesh *m, DLevel &D, const T *b, const T *x, T *r, const int *flags) {
  hipStream_t s = m->stream;
  const T *ac = LV<T>::ac(D), *anb = LV<T>::anb(D);
  if (D.sell) {
    hipLaunchKernelGGL((k_resid_sell<T>), dim3(gfor(D.sell_slots)), dim3(kBlock), 0, s, D.sell_slots, D.sell_ptr.p,
                       D.sell_row.p, D.sell_col.p, LV<T>::val(D), ac, b, x, r, flags);
    return;
  }
  const EllNbr nb(D.nbr, D.nbr16);
  if (D.w <= 8)
    hipLaunchKernelGGL((k_resid<1, T>), dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.w, nb, ac, anb, b, x, r, flags);
  else if (D.w <= 16)
    hipLaunchKernelGGL((k_resid<4, T>), dim3(gfor(D.n * 4)), dim3(kBlock), 0, s, D.n, D.w, nb, ac, anb, b, x, r,
                       flags);
  else if (D.w <= 32)
    hipLaunchKernelGGL((k_resid<8, T>), dim3(gfor(D.n * 8)), dim3(kBlock), 0, s, D.n, D.w, nb, ac, anb, b, x, r,
                       flags);
  else
    hipLaunchKernelGGL((k_resid<16, T>), dim3(gfor(D.n * 16)), dim3(kBlock), 0, s, D.n, D.w, nb, ac, anb, b, x, r,
                       flags);
}

// multicolour symmetric Gauss-Seidel only where it is cheap: few colours,
// narrow rows
static bool use_gs(const fs_mesh *m, const DLevel &D) {
  return m->mg_smoother == 1 && (int)D.color_ptr.size() - 1 <= 4 && D.w <= 8;
}

template <class T>
static void gs_sweeps(fs_mesh *m, DLevel &D, const T *b, T *x, int sweeps, const int *flags) {
  hipStream_t s = m->stream;
  const T *ac = LV<T>::ac(D), *anb = LV<T>::anb(D);
  const int nc = (int)D.color_ptr.size() - 1;
  // symmetric sweeps; a colour is an independent set, so a colour swept
  // twice in a row (the last of a forward pass, the first of the next
  // forward pass) would recompute the same values and is swept once
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int c = sw == 0 ? 0 : 1; c < nc; ++c) {
      const int bb = D.color_ptr[c], e = D.color_ptr[c + 1];
      hipLaunchKernelGGL(k_gsc<T>, dim3(gfor(e - bb)), dim3(kBlock), 0, s, e - bb, D.color_cells.p + bb, D.n,
                         D.w, D.nbr, ac, anb, b, x, flags);
    }
    for (int c = nc - 2; c >= 0; --c) {
      const int bb = D.color_ptr[c], e = D.color_ptr[c + 1];
      hipLaunchKernelGGL(k_gsc<T>, dim3(gfor(e - bb)), dim3(kBlock), 0, s, e - bb, D.color_cells.p + bb, D.n,
                         D.w, D.nbr, ac, anb, b, x, flags);
    }
  }
}

// Remaining pre-smoothing sweeps after the first (which the entry or the
// restriction kernel took); returns the buffer that holds the result.
template <class T>
static T *pre_smooth(fs_mesh *m, DLevel &D, int level, const T *b, int sweeps, T om, T *&other, const int *flags) {
  T *cur = LV<T>::x(D), *oth = LV<T>::t(D);
  if (use_gs(m, D)) {
    gs_sweeps<T>(m, D, b, cur, sweeps, flags); // x was zeroed by the caller
  } else {
    for (int k = 1; k < sweeps; ++k) {
      launch_sweep<T, T, false>(m, D, level, b, cur, nullptr, nullptr, 0, T(0), om, oth, flags);
      std::swap(cur, oth);
    }
  }
  other = oth;
  return cur;
}

// Prolongation of the coarse correction xc into x plus the post-smoothing
// sweeps; the last operation writes dst (type TO), or a level buffer when dst
// is null.  Returns the buffer holding the result when dst is null.
template <class T, class TO>
static T *post_smooth(fs_mesh *m, DLevel &D, int level, int nc, const T *b, T *x, T *other, const T *xc, int sweeps,
                      T om, T al, TO *dst, const int *flags) {
  hipStream_t s = m->stream;
  if (use_gs(m, D)) {
    hipLaunchKernelGGL((k_prolong<T, T>), dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.parent.p, xc, x, al, x,
                       flags);
    gs_sweeps<T>(m, D, b, x, sweeps, flags);
    if (dst) hipLaunchKernelGGL((k_convert<T, TO>), dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, x, dst, flags);
    return x;
  }
  if (sweeps == 0) {
    if (dst) {
      hipLaunchKernelGGL((k_prolong<T, TO>), dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.parent.p, xc, x, al, dst,
                         flags);
    } else {
      hipLaunchKernelGGL((k_prolong<T, T>), dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.parent.p, xc, x, al, x,
                         flags);
    }
    return x;
  }
  // the prolongation fused into the first sweep (each row adds the coarse
  // value of itself and of its neighbours), or as a pass of its own first
  // (FS_OPT_MG_FUSE_PROLONG 0: cheaper where the neighbours' parent lookups
  // are scattered)
  const bool fuse = m->mg_fuse_prolong != 0;
  if (!fuse)
    hipLaunchKernelGGL((k_prolong<T, T>), dim3(gfor(D.n)), dim3(kBlock), 0, s, D.n, D.parent.p, xc, x, al, x,
                       flags);
  T *cur = x, *oth = other;
  for (int k = 0; k < sweeps; ++k) {
    const bool last = k + 1 == sweeps, pro = fuse && k == 0;
    if (last && dst) {
      if (pro) launch_sweep<T, TO, true>(m, D, level, b, cur, D.parent.p, xc, nc, al, om, dst, flags);
      else launch_sweep<T, TO, false>(m, D, level, b, cur, nullptr, nullptr, 0, T(0), om, dst, flags);
      return cur;
    }
    if (pro) launch_sweep<T, T, true>(m, D, level, b, cur, D.parent.p, xc, nc, al, om, oth, flags);
    else launch_sweep<T, T, false>(m, D, level, b, cur, nullptr, nullptr, 0, T(0), om, oth, flags);
    std::swap(cur, oth);
  }
  return cur;
}

// out = M^{-1} rhs: one V-cycle on the hierarchy held in precision T.  The
// first sweep of each level is fused into the kernel that produces its
// right-hand side, the prolongation into the first post-smoothing sweep, the
// FP64 output into the last one; the levels from the tail's first level down
// run in the one-workgroup tail kernel.  out must not alias rhs.
template <class T>
static void vcycle(fs_mesh *m, const double *rhs, double *out, double alpha, const int *flags) {
  Hierarchy &H = *m->amg;
  hipStream_t s = m->stream;
  const int L = (int)H.lev.size();
  TailImage &ti = LV<T>::tail(H);
  const int ts = std::min(ti.ts, L);
  const T om = (T)m->mg_omega, al = (T)alpha;
  const int pre0 = m->mg_pre, post0 = m->mg_post;
  const int cpre = m->mg_cpre >= 0 ? m->mg_cpre : pre0, cpost = m->mg_cpost >= 0 ? m->mg_cpost : post0;
  const int coarse = std::max(m->mg_coarse_sweeps, 1);
  const bool fp64 = sizeof(T) == 8;

  DLevel &D0 = H.lev[0];
  const T *b0 = fp64 ? (const T *)rhs : LV<T>::b(D0);
  // level 0 with no coarser level: smoothing only
  if (L == 1) {
    const bool gs = use_gs(m, D0);
    hipLaunchKernelGGL(k_enter<T>, dim3(gfor(D0.n)), dim3(kBlock), 0, s, D0.n, LV<T>::ac(D0), rhs,
                       fp64 ? nullptr : LV<T>::b(D0), LV<T>::x(D0), om, gs ? 0 : 1, flags);
    T *oth = nullptr;
    T *x = pre_smooth<T>(m, D0, 0, b0, coarse, om, oth, flags);
    hipLaunchKernelGGL((k_convert<T, double>), dim3(gfor(D0.n)), dim3(kBlock), 0, s, D0.n, x, out, flags);
    return;
  }
  std::vector<T *> res(L, nullptr), oth(L, nullptr);
  // levels above the bottom: grid-wide kernels
  const int top = ts < L ? ts : L - 1;
  // the restriction from level l into l + 1 also starts level l + 1: mode 0
  // hands it to the tail kernel, 1 makes the first Jacobi sweep from zero,
  // 2 zeroes it (Gauss-Seidel)
  auto mode_from = [&](int l) {
    if (l + 1 == ts) return 0;                                // the tail kernel takes over
    if (l + 1 == L - 1) return 1;                             // grid-wide coarsest: `coarse` sweeps
    return (!use_gs(m, H.lev[l + 1]) && cpre > 0) ? 1 : 2;
  };
  auto restrict_from = [&](int l, const T *b, const T *x) {
    DLevel &D = H.lev[l], &C = H.lev[l + 1];
    const int mode = mode_from(l);
    if (l > 0 && !use_gs(m, D) && !D.sell) {
      launch_resid_restrict<T>(m, D, C, b, x, om, mode, flags);
    } else {
      launch_resid<T>(m, D, b, x, LV<T>::r(D), flags);
      hipLaunchKernelGGL(k_restrict<T>, dim3(gfor(C.n)), dim3(kBlock), 0, s, C.n, C.mptr.p, C.mem.p, LV<T>::r(D),
                         LV<T>::b(C), LV<T>::ac(C), LV<T>::x(C), om, mode, flags);
    }
  };
  // the levels below the grid-wide ones
  auto bottom = [&]() {
    if (ts < L) {
      DLevel &Dt = H.lev[ts];
      hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(kTailThreads), ti.desc.lds_bytes, s, ti.ddesc.p, ti.img.p,
                         LV<T>::b(Dt), LV<T>::x(Dt), cpre, cpost, coarse, om, al, flags);
      FS_HIP_THROW(hipGetLastError()); // a refused launch (LDS request) must not pass silently
      res[ts] = LV<T>::x(Dt);
    } else {
      DLevel &Dc = H.lev[L - 1];
      // the restriction took the first of the coarse sweeps
      T *o = nullptr;
      res[L - 1] = pre_smooth<T>(m, Dc, L - 1, LV<T>::b(Dc), coarse, om, o, flags);
    }
  };
  // prolongation and post-smoothing from the bottom up to level stop (level
  // 0 writes out)
  auto ascend = [&](int stop) {
    for (int l = top - 1; l >= stop; --l) {
      DLevel &D = H.lev[l];
      const T *b = l == 0 ? b0 : LV<T>::b(D);
      const int npost = l == 0 ? post0 : cpost;
      const int nc = H.lev[l + 1].n;
      if (l == 0) {
        post_smooth<T, double>(m, D, 0, nc, b, res[0], oth[0], res[1], npost, om, al, out, flags);
      } else {
        res[l] = post_smooth<T, T>(m, D, l, nc, b, res[l], oth[l], res[l + 1], npost, om, al, (T *)nullptr, flags);
        oth[l] = res[l] == LV<T>::x(D) ? LV<T>::t(D) : LV<T>::x(D);
      }
    }
  };
  // a second coarse-grid correction of level l1 >= 1 (F and W cycles):
  // pre-sweeps from its current iterate, then a fresh descent (every coarser
  // level from zero) to the bottom
  auto redescend = [&](int l1) {
    DLevel &D = H.lev[l1];
    const T *b = LV<T>::b(D);
    T *cur = res[l1], *o = oth[l1];
    if (use_gs(m, D)) {
      gs_sweeps<T>(m, D, b, cur, cpre, flags);
    } else {
      for (int k = 0; k < cpre; ++k) {
        launch_sweep<T, T, false>(m, D, l1, b, cur, nullptr, nullptr, 0, T(0), om, o, flags);
        std::swap(cur, o);
      }
    }
    res[l1] = cur;
    oth[l1] = o;
    restrict_from(l1, b, cur);
    for (int l = l1 + 1; l < top; ++l) {
      DLevel &Dl = H.lev[l];
      res[l] = pre_smooth<T>(m, Dl, l, LV<T>::b(Dl), cpre, om, oth[l], flags);
      restrict_from(l, LV<T>::b(Dl), res[l]);
    }
    bottom();
  };

  // one Jacobi pre-sweep on the finest level: the entry and the residual of
  // that sweep in one kernel
  const bool fuse0 = !use_gs(m, D0) && pre0 == 1 && top > 0;
  if (!fuse0) {
    const bool gs = use_gs(m, D0);
    hipLaunchKernelGGL(k_enter<T>, dim3(gfor(D0.n)), dim3(kBlock), 0, s, D0.n, LV<T>::ac(D0), rhs,
                       fp64 ? nullptr : LV<T>::b(D0), LV<T>::x(D0), om, (!gs && pre0 > 0) ? 1 : 0, flags);
  }
  for (int l = 0; l < top; ++l) {
    DLevel &D = H.lev[l];
    const T *b = l == 0 ? b0 : LV<T>::b(D);
    const int npre = l == 0 ? pre0 : cpre;
    if (l == 0 && fuse0) {
      DLevel &C = H.lev[1];
      hipLaunchKernelGGL(k_enter_resid<T>, dim3(gfor(D0.n)), dim3(kBlock), 0, s, D0.n, D0.w,
                         EllNbr(D0.nbr, D0.nbr16), LV<T>::ac(D0), LV<T>::anb(D0), rhs, fp64 ? nullptr : LV<T>::b(D0),
                         LV<T>::x(D0), LV<T>::r(D0), om, flags);
      res[0] = LV<T>::x(D0);
      oth[0] = LV<T>::t(D0);
      hipLaunchKernelGGL(k_restrict<T>, dim3(gfor(C.n)), dim3(kBlock), 0, s, C.n, C.mptr.p, C.mem.p, LV<T>::r(D0),
                         LV<T>::b(C), LV<T>::ac(C), LV<T>::x(C), om, mode_from(0), flags);
      continue;
    }
    res[l] = pre_smooth<T>(m, D, l, b, npre, om, oth[l], flags);
    restrict_from(l, b, res[l]);
  }
  bottom();
  // F and W cycles (FS_OPT_MG_CYCLE 1, 2) in the pattern of the reference's
  // fMultiGridCycle and wMultiGridCycle (MultiGridCycle.for:84-127, 162-220):
  // back up to level l1 and down again, for l1 from the deepest grid-wide
  // level up to level 1, and for W once more from level 2 down to the
  // deepest; the levels from the tail's first level down count as one
  if (m->mg_cycle >= 1)
    for (int l1 = top - 1; l1 >= 1; --l1) {
      ascend(l1);
      redescend(l1);
    }
  if (m->mg_cycle == 2)
    for (int l1 = 2; l1 <= top - 1; ++l1) {
      ascend(l1);
      redescend(l1);
    }
  ascend(0);
}

void amg_apply(fs_mesh *m, const double *rhs, double *out, const int *flags, double alpha) {
  if (m->amg->fp32) vcycle<float>(m, rhs, out, alpha, flags);
  else vcycle<double>(m, rhs, out, alpha, flags);
}

// The reference's multigrid solve loop (AlgebraicMultiGridMainRoutine.for:
// 371-416): cycles while the fine L1 residual exceeds rrf*ires and at most
// MaxMultiGridCycles; each cycle corrects with one V-cycle of the residual.
int amg_solve(fs_mesh *m, const double *ac, const double *anb, const double *bc, double *dphi,
              int solver, bool warm) {
  (void)solver;
  amg_prepare(m, ac, anb);
  hipStream_t s = m->stream;
  const int n = m->n;
  double *r = m->wk.v[8], *c = m->wk.v[9];
  int *flags = m->wk.flags;
  // additive correction as in the reference unless asked otherwise
  const double alpha = m->mg_alpha > 0.0 ? m->mg_alpha : 1.0;
  // flags[2] holds itmax for the Krylov solvers; the multigrid loop is
  // bounded by MaxMultiGridCycles instead
  int itmax_cycles = m->mg_maxcycles;
  FS_HIP_THROW(hipMemcpyAsync(flags + 2, &itmax_cycles, sizeof(int), hipMemcpyHostToDevice, s));
  const int gr = std::min(kPartials, gfor(n));
  // the first cycle's right-hand side: bc (dphi = 0), or the residual of the
  // initial guess
  if (warm)
    hipLaunchKernelGGL(k_amg_res, dim3(gr), dim3(kBlock), 0, s, n, m->w, m->d_nbr.p, ac, anb, bc, dphi, r,
                       m->wk.partial, (const int *)flags);
  else
    FS_HIP_THROW(hipMemcpyAsync(r, bc, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  int fl[4] = {0, 0, 0, 0};
  FS_HIP_THROW(hipMemcpyAsync(fl, flags, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
  FS_HIP_THROW(hipStreamSynchronize(s));
  int launched = 0;
  while (!fl[0] && launched < itmax_cycles) {
    const int batch = std::min(m->check_every, itmax_cycles - launched);
    for (int b = 0; b < batch; ++b) {
      amg_apply(m, r, c, flags, alpha);
      hipLaunchKernelGGL(k_add2, dim3(gfor(n)), dim3(kBlock), 0, s, n, dphi, c, flags);
      hipLaunchKernelGGL(k_amg_res, dim3(gr), dim3(kBlock), 0, s, n, m->w, m->d_nbr.p, ac, anb,
                         bc, dphi, r, m->wk.partial, flags);
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
