// Linear-solve path on the device: the reference's SolveEquation
// (SolveEquations/AlgebraicSolvers/SolveAlegbraicEquations.for:4) with its
// residual norms, PBCG, Gauss-Seidel and D-ILU solvers, and the Krylov
// solvers that carry the AMG preconditioner.
//
// Layout: every cell-indexed vector is FP64 in HBM; the ELL matrix keeps the
// reference's column-major anb(N,W), so slot j of consecutive cells is one
// contiguous stream and every thread's loads coalesce.  Gathers of x through
// the neighbour table are the only irregular accesses.
//
// Reductions are deterministic: each block reduces with wavefront shuffles in
// a fixed order into one of kPartials partials (grid-stride loop over a fixed
// grid), and a one-block kernel combines the partials in index order.  Per
// cell, every sum is formed in the reference's slot order, so a sweep,
// residual or product of one cell is bitwise the reference's.
//
// Convergence never needs the host inside a batch: a one-block kernel
// updates the device flags (done, iteration count), every iteration kernel
// returns immediately once done is set, and the host reads the flags only
// every check_every iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <stdexcept>

#include "fs_common.h"
#include "fs_kernels.h"
#include "fs_solve.h"

namespace fs {

// ---------------------------------------------------------------- scalars
enum Slot {
  S_IRES = 0, S_FRES, S_RRF, S_BKNUM, S_BKDEN, S_AKDEN, S_BK, S_AK,
  S_RHO, S_ALPHA, S_OMEGA, S_BETA, S_TS, S_TT, S_RV, S_RES_ABS, S_RES_MAX,
  S_RES_RMS, S_RES_SCALED, S_ACPHI, S_GAMMA, S_NSLOTS
};
enum Flag {
  F_DONE = 0, F_ITER = 1, F_ITMAX = 2, F_REJECT = 3, F_INNER = 4, F_GATE = 5,
  F_BAD = 6,     // the input was refused: 2 = a non-finite ac or bc, 1 = a diagonal <= 0
                 // for a solver that divides by it
  F_NEEDPOS = 7, // set by the host when the solver divides by the diagonal
  F_NFLAGS = 8
};
// F_INNER: the GMRES cycle in progress has finished its Arnoldi steps;
// F_GATE = F_DONE | F_INNER, the flag the cycle's kernels (and the V-cycle
// they call) test

constexpr int kFinBlock = 1024; // threads of the one-workgroup finalise kernels

static inline int grid_for(int n) {
  int g = (n + kBlock - 1) / kBlock;
  return g < 1 ? 1 : g;
}
static inline int grid_red(int n) {
  int g = (n + kBlock - 1) / kBlock;
  if (g > kPartials) g = kPartials;
  return g < 1 ? 1 : g;
}

// The Krylov products stream the FP64 matrix (read twice per iteration, too
// large to stay cached beside the vectors): non-temporal matrix loads.
constexpr bool kStreamMatrix = true;

// y = A*x in the order of dsprsax (PBCGMethod.for:324-329): diagonal first,
// then the off-diagonals in slot order.
__device__ __forceinline__ double ell_ax(int i, int n, int w, EllNbr nbr,
                                         const double *__restrict__ ac,
                                         const double *__restrict__ anb,
                                         const double *__restrict__ x) {
  const double s = ld_mat<kStreamMatrix>(&ac[i]) * x[i];
  return ell_fold<kEllBatch, kStreamMatrix>(i, n, w, nbr, anb, x, s,
                             [](double acc, double a, double xk) { return acc + a * xk; });
}

// r = b - A*x in the order of CheckConvergence (CheckConvergenceRoutine.for:
// 23-34): b - ac*x first, then each off-diagonal subtracted in slot order.
__device__ __forceinline__ double ell_res(int i, int n, int w, EllNbr nbr,
                                          const double *__restrict__ ac,
                                          const double *__restrict__ anb,
                                          const double *__restrict__ b,
                                          const double *__restrict__ x) {
  const double r = b[i] - ld_mat<kStreamMatrix>(&ac[i]) * x[i];
  return ell_fold<kEllBatch, kStreamMatrix>(i, n, w, nbr, anb, x, r,
                             [](double acc, double a, double xk) { return acc - a * xk; });
}

// y = A^T*x in the order of dsprstx (PBCGMethod.for:359-368): diagonal first,
// then the contributions of source rows in increasing row order.  tlin/tsrc
// list, per target cell, the linear anb index and source row of each
// transposed entry, pre-sorted by (row, slot).
__device__ __forceinline__ double ell_atx(int i, int n, int wt, const int *__restrict__ tlin,
                                          const int *__restrict__ tsrc,
                                          const double *__restrict__ ac,
                                          const double *__restrict__ anb,
                                          const double *__restrict__ x) {
  double s = ac[i] * x[i];
  for (int j = 0; j < wt; ++j) {
    const int l = tlin[(size_t)j * n + i];
    if (l >= 0) s = s + anb[l] * x[tsrc[(size_t)j * n + i]];
  }
  return s;
}

// ---------------------------------------------------------------- kernels

// Initial residual norms of SolveEquation (CalculateResidualsRoutine.for:
// 19-67) with dphi = 0: r_i = bc_i.  Partials: sum|r|, max|r|, sum r^2,
// max|ac*fit|, and the input check: the largest of 0 (row fine), 1 (ac <= 0)
// and 2 (ac or bc not finite).
__global__ void k_init_norms(int n, const double *__restrict__ ac, const double *__restrict__ bc,
                             const double *__restrict__ fit, double *__restrict__ part) {
  double s1 = 0.0, m1 = -1.0, s2 = 0.0, m2 = -1.0, bad = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double a = ac[i], b = bc[i];
    const double r = fabs(b);
    s1 += r;
    m1 = fmax(m1, r);
    s2 += r * r;
    m2 = fmax(m2, fabs(a * fit[i]));
    bad = fmax(bad, !(isfinite(a) && isfinite(b)) ? 2.0 : (a > 0.0 ? 0.0 : 1.0));
  }
  s1 = block_sum<kBlock>(s1);
  m1 = block_max<kBlock>(m1);
  s2 = block_sum<kBlock>(s2);
  m2 = block_max<kBlock>(m2);
  bad = block_max<kBlock>(bad);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[kPartials + blockIdx.x] = m1;
    part[2 * kPartials + blockIdx.x] = s2;
    part[3 * kPartials + blockIdx.x] = m2;
    part[4 * kPartials + blockIdx.x] = bad;
  }
}

// dphi = 0 (SolveAlegbraicEquations.for:47), unless the input was refused
__global__ void k_zero_start(int n, double *__restrict__ x, const int *__restrict__ flags) {
  if (flags[F_BAD]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// x = 0 by a kernel (no memset node, so the sequence can be graph-captured)
__global__ void k_zero(int n, double *__restrict__ x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// L1 norm of the true residual bc - A*x (CheckConvergence).
__global__ void k_resid_l1(int n, int w, EllNbr nbr, const double *__restrict__ ac,
                           const double *__restrict__ anb, const double *__restrict__ b,
                           const double *__restrict__ x, double *__restrict__ part,
                           const int *__restrict__ flags) {
  if (flags && flags[F_DONE]) return;
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    s += fabs(ell_res(i, n, w, nbr, ac, anb, b, x));
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// r = b - A*x, also copied to r0; partial L1 of r.
// x *= gamma (the L2-optimal scaling of an initial guess)
__global__ void k_warm_scale(int n, double *__restrict__ x, const double *__restrict__ scal,
                             const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = scal[S_GAMMA] * x[i];
}

// a rejected initial guess becomes the zero start: x = 0
__global__ void k_warm_fix(int n, double *__restrict__ x, const int *__restrict__ flags) {
  if (!flags[F_REJECT]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// p = z (first iteration) or z + beta*p
__global__ void k_cg_p(int n, const double *__restrict__ z, double *__restrict__ p,
                       const double *__restrict__ scal, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const bool first = flags[F_ITER] == 0;
  const double be = scal[S_BETA];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = first ? z[i] : z[i] + be * p[i];
}

// x += alpha p, r -= alpha q, partial L1(r)
__global__ void k_cg_update(int n, double *__restrict__ x, double *__restrict__ r,
                            const double *__restrict__ p, const double *__restrict__ q,
                            const double *__restrict__ scal, double *__restrict__ part,
                            const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const double al = scal[S_ALPHA];
  double a = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] = x[i] + al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    a += fabs(ri);
  }
  a = block_sum<kBlock>(a);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void k_resid_vec(int n, int w, EllNbr nbr, const double *__restrict__ ac,
                            const double *__restrict__ anb, const double *__restrict__ b,
                            const double *__restrict__ x, double *__restrict__ r,
                            double *__restrict__ r0, double *__restrict__ part) {
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double v = ell_res(i, n, w, nbr, ac, anb, b, x);
    r[i] = v;
    if (r0) r0[i] = v;
    s += fabs(v);
  }
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_spmv(int n, int w, const int *__restrict__ nbr, const double *__restrict__ ac,
                       const double *__restrict__ anb, const double *__restrict__ x,
                       double *__restrict__ y) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) y[i] = ell_ax(i, n, w, nbr, ac, anb, x);
}

// PBCG start (PBCGMethod.for:147-161): r = bc - A*dphi, rr = r.
__global__ void k_pbcg_start(int n, const double *__restrict__ bc, double *__restrict__ r,
                             double *__restrict__ rr) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const double v = bc[i] - r[i];
    r[i] = v;
    rr[i] = v;
  }
}

// bknum partials: sum (r*dc)*rr (PBCGMethod.for:163-170).
// z = dc*r (the reference's D-ILU asolve) or, with dc null, the
// preconditioned vector zp given (one AMG V-cycle)
__global__ void k_pbcg_bknum(int n, const double *__restrict__ r, const double *__restrict__ rr,
                             const double *__restrict__ dc, const double *__restrict__ zp,
                             double *__restrict__ part, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    s += (dc ? r[i] * dc[i] : zp[i]) * rr[i];
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// p = bk*p + r*dc, pp = bk*pp + rr*dc; on the first iteration p = r*dc,
// pp = rr*dc (PBCGMethod.for:172-192).
__global__ void k_pbcg_p(int n, const double *__restrict__ r, const double *__restrict__ rr,
                         const double *__restrict__ dc, const double *__restrict__ zp,
                         const double *__restrict__ zzp, double *__restrict__ p,
                         double *__restrict__ pp, const double *__restrict__ scal,
                         const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double z = dc ? r[i] * dc[i] : zp[i], zz = dc ? rr[i] * dc[i] : zzp[i];
  if (flags[F_ITER] == 0) {
    p[i] = z;
    pp[i] = zz;
  } else {
    const double bk = scal[S_BK];
    p[i] = bk * p[i] + z;
    pp[i] = bk * pp[i] + zz;
  }
}

// z = A*p, zz = A^T*pp, akden partials = sum z*pp (PBCGMethod.for:195-205).
__global__ void k_pbcg_ap(int n, int w, int wt, const int *__restrict__ nbr,
                          const int *__restrict__ tlin, const int *__restrict__ tsrc,
                          const double *__restrict__ ac, const double *__restrict__ anb,
                          const double *__restrict__ p, const double *__restrict__ pp,
                          double *__restrict__ z, double *__restrict__ zz,
                          double *__restrict__ part, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double zi = ell_ax(i, n, w, nbr, ac, anb, p);
    z[i] = zi;
    zz[i] = ell_atx(i, n, wt, tlin, tsrc, ac, anb, pp);
    s += zi * pp[i];
  }
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// dphi += ak*p, r -= ak*z, rr -= ak*zz (PBCGMethod.for:207-213).
__global__ void k_pbcg_axpy(int n, double *__restrict__ dphi, double *__restrict__ r,
                            double *__restrict__ rr, const double *__restrict__ p,
                            const double *__restrict__ z, const double *__restrict__ zz,
                            const double *__restrict__ scal, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double ak = scal[S_AK];
  dphi[i] = dphi[i] + ak * p[i];
  r[i] = r[i] - ak * z[i];
  rr[i] = rr[i] - ak * zz[i];
}

// Exact natural-order D-ILU diagonal, one dependency level per launch
// (FactorizeUsingILU.for:23-75 in pull form): dc_j = 1/(ac_j - sum over the
// lower couplings i of anb(j,i)*dc_i*anb(i,j)), terms in the reference's
// push order.
__global__ void k_ilu_level(int cnt, const int *__restrict__ cells, const int *__restrict__ ptr,
                            const int *__restrict__ ta, const int *__restrict__ ti,
                            const int *__restrict__ tb, const double *__restrict__ ac,
                            const double *__restrict__ anb, double *__restrict__ dc) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int j = cells[t];
  double d = ac[j];
  for (int e = ptr[j]; e < ptr[j + 1]; ++e) d = d - anb[ta[e]] * dc[ti[e]] * anb[tb[e]];
  dc[j] = 1.0 / d;
}

// One colour of a multicolour Gauss-Seidel sweep: dphi_i = (bc_i - sum
// anb*dphi_nb)/ac_i for the cells of the colour (SORMethod.for:22-36 per cell).
__global__ void k_gs_color(int cnt, const int *__restrict__ cells, int n, int w,
                           EllNbr nbr, const double *__restrict__ ac,
                           const double *__restrict__ anb, const double *__restrict__ bc,
                           double *__restrict__ x, const int *__restrict__ flags) {
  if (flags && flags[F_DONE]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int i = cells[t];
  const double s = ell_fold<kEllBatch>(i, n, w, nbr, anb, (const double *)x, bc[i],
                                       [](double acc, double a, double xk) { return acc - a * xk; });
  x[i] = s / ac[i];
}

// ---- generic vector kernels for the Krylov solvers

// partial dots: out0 = sum a*b, and if c: out1 = sum c*d
__global__ void k_dot2(int n, const double *__restrict__ a, const double *__restrict__ b,
                       const double *__restrict__ c, const double *__restrict__ d,
                       double *__restrict__ part, const int *__restrict__ flags) {
  if (flags && flags[F_DONE]) return;
  double s0 = 0.0, s1 = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    s0 += a[i] * b[i];
    if (c) s1 += c[i] * d[i];
  }
  s0 = block_sum<kBlock>(s0);
  s1 = block_sum<kBlock>(s1);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kPartials + blockIdx.x] = s1;
  }
}

// BiCGSTAB: p = r + beta*(p - omega*v)
__global__ void k_bicg_p(int n, const double *__restrict__ r, const double *__restrict__ v,
                         double *__restrict__ p, const double *__restrict__ scal,
                         const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (flags[F_ITER] == 0) {
    p[i] = r[i];
  } else {
    p[i] = r[i] + scal[S_BETA] * (p[i] - scal[S_OMEGA] * v[i]);
  }
}

// v = A*phat and partial (r0, v)
__global__ void k_spmv_dot(int n, int w, EllNbr nbr, const double *__restrict__ ac,
                           const double *__restrict__ anb, const double *__restrict__ x,
                           double *__restrict__ y, const double *__restrict__ d,
                           double *__restrict__ part, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double yi = ell_ax(i, n, w, nbr, ac, anb, x);
    y[i] = yi;
    s += yi * d[i];
  }
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// s = r - alpha*v
__global__ void k_bicg_s(int n, const double *__restrict__ r, const double *__restrict__ v,
                         double *__restrict__ s, const double *__restrict__ scal,
                         const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) s[i] = r[i] - scal[S_ALPHA] * v[i];
}

// t = A*shat, partials (t,s) and (t,t)
__global__ void k_spmv_dot_ts(int n, int w, EllNbr nbr,
                              const double *__restrict__ ac, const double *__restrict__ anb,
                              const double *__restrict__ x, double *__restrict__ t,
                              const double *__restrict__ s, double *__restrict__ part,
                              const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  double a0 = 0.0, a1 = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double ti = ell_ax(i, n, w, nbr, ac, anb, x);
    t[i] = ti;
    a0 += ti * s[i];
    a1 += ti * ti;
  }
  a0 = block_sum<kBlock>(a0);
  a1 = block_sum<kBlock>(a1);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a0;
    part[kPartials + blockIdx.x] = a1;
  }
}

// x += alpha*phat + omega*shat; r = s - omega*t; partial L1(r)
// x += alpha phat + omega shat, r = s - omega t, with the partials of |r|_1
// and of the next iteration's rho = (r0, r) (the sums k_dot2 would form)
__global__ void k_bicg_update(int n, double *__restrict__ x, double *__restrict__ r,
                              const double *__restrict__ phat, const double *__restrict__ shat,
                              const double *__restrict__ s, const double *__restrict__ t,
                              const double *__restrict__ r0, const double *__restrict__ scal,
                              double *__restrict__ part, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const double al = scal[S_ALPHA], om = scal[S_OMEGA];
  double a = 0.0, d = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] = x[i] + al * phat[i] + om * shat[i];
    const double ri = s[i] - om * t[i];
    r[i] = ri;
    a += fabs(ri);
    d += r0[i] * ri;
  }
  a = block_sum<kBlock>(a);
  d = block_sum<kBlock>(d);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kPartials + blockIdx.x] = d;
  }
}

// ---- the reference's ILU iteration (ILUMethod.for:4): rc = bc - A dphi with
// the diagonal first, then level-scheduled forward and backward substitution
// with the operation order of the sequential loops
__global__ void k_ilu_resid(int n, int w, const int *__restrict__ nbr, const double *__restrict__ ac,
                            const double *__restrict__ anb, const double *__restrict__ bc,
                            const double *__restrict__ dphi, double *__restrict__ rc, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= n) return;
  rc[i] = ell_fold<kEllBatch>(i, n, w, nbr, anb, dphi, -ac[i] * dphi[i] + bc[i],
                              [](double acc, double a, double xk) { return acc - a * xk; });
}

__global__ void k_ilu_fwd(int cnt, const int *__restrict__ cells, const int *__restrict__ ptr,
                          const int *__restrict__ lin, const int *__restrict__ src, const double *__restrict__ anb,
                          const double *__restrict__ dc, double *rc, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int j = cells[t];
  double v = rc[j];
  for (int q = ptr[j]; q < ptr[j + 1]; ++q) {
    const int i = src[q];
    const double p1 = dc[i] * rc[i];
    v = v - anb[lin[q]] * p1;
  }
  rc[j] = v;
}

__global__ void k_ilu_bwd(int cnt, const int *__restrict__ cells, const int *__restrict__ ptr,
                          const int *__restrict__ lin, const int *__restrict__ src, const double *__restrict__ anb,
                          const double *__restrict__ dc, double *rc, double *__restrict__ dphi,
                          const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= cnt) return;
  const int i = cells[t];
  double v = rc[i];
  for (int q = ptr[i]; q < ptr[i + 1]; ++q) v = v - anb[lin[q]] * rc[src[q]];
  const double p1 = dc[i] * v;
  rc[i] = p1;
  dphi[i] = dphi[i] + p1;
}

// ---- restarted GMRES with right preconditioning (matrixSolverOMP.f90:48 is
// the reference's GMRES; this one keeps its Krylov basis and Hessenberg
// matrix on the device).  Classical Gram-Schmidt applied twice, all dot
// products FP64 over fixed partials.  gm[] holds the small state:
// H[(m+1) x m] column-major, then cs[m], sn[m], g[m+1], y[m], and the slots
// below.
enum GmSlot { GM_BETA = 0, GM_JEND, GM_TOL2, GM_NSLOT };
inline __host__ __device__ size_t gm_h(int) { return 0; }
inline __host__ __device__ size_t gm_cs(int m) { return (size_t)(m + 1) * m; }
inline __host__ __device__ size_t gm_sn(int m) { return gm_cs(m) + m; }
inline __host__ __device__ size_t gm_g(int m) { return gm_sn(m) + m; }
inline __host__ __device__ size_t gm_y(int m) { return gm_g(m) + m + 1; }
inline __host__ __device__ size_t gm_slot(int m) { return gm_y(m) + m; }
inline __host__ __device__ size_t gm_hpass(int m) { return gm_slot(m) + GM_NSLOT; }
inline __host__ __device__ size_t gm_size(int m) { return gm_hpass(m) + m + 1; }

__global__ void k_gm_spmv(int n, int w, const int *__restrict__ nbr, const double *__restrict__ ac,
                          const double *__restrict__ anb, const double *__restrict__ x, double *__restrict__ y,
                          const int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i < n) y[i] = ell_ax(i, n, w, nbr, ac, anb, x);
}

// part[j*kPartials + block] = (V_j, w); blockIdx.y = j
__global__ void k_gm_dots(int n, const double *__restrict__ V, size_t ldv, const double *__restrict__ wv,
                          double *__restrict__ part, const int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  const double *vj = V + (size_t)blockIdx.y * ldv;
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += vj[i] * wv[i];
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * kPartials + blockIdx.x] = s;
}

__device__ double reduce_partials(const double *part, int np);

// hpass[j] = (V_j, w) from the partials; H[j][col] += hpass[j]; blockIdx.x = j
__global__ void k_gm_hfin(int np, int m, int col, const double *__restrict__ part, double *__restrict__ gm,
                          const int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  const int j = blockIdx.x;
  const double a = reduce_partials(part + (size_t)j * kPartials, np);
  if (threadIdx.x == 0) {
    gm[gm_hpass(m) + j] = a;
    gm[gm_h(m) + (size_t)col * (m + 1) + j] += a;
  }
}

// w -= sum_{j < nv} hpass[j] V_j
__global__ void k_gm_maxpy(int n, int nv, int m, const double *__restrict__ V, size_t ldv,
                           const double *__restrict__ gm, double *__restrict__ wv, const int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = wv[i];
  for (int j = 0; j < nv; ++j) s -= gm[gm_hpass(m) + j] * V[(size_t)j * ldv + i];
  wv[i] = s;
}

// partials of (w, w)
__global__ void k_gm_norm2(int n, const double *__restrict__ wv, double *__restrict__ part,
                           const int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  double s = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += wv[i] * wv[i];
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Column `col` of the Arnoldi relation: H[col+1][col] = ||w||, the previous
// Givens rotations, a new one, the residual estimate |g[col+1]|; ends the
// cycle's Arnoldi steps when the estimate meets the L2 target, the basis is
// full or w vanished.
__global__ void k_gm_col(int np, int m, int col, const double *__restrict__ part, double *__restrict__ gm,
                         int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  const double hn = sqrt(reduce_partials(part, np));
  if (threadIdx.x != 0) return;
  double *H = gm + gm_h(m) + (size_t)col * (m + 1);
  double *cs = gm + gm_cs(m), *sn = gm + gm_sn(m), *g = gm + gm_g(m);
  H[col + 1] = hn;
  for (int i = 0; i < col; ++i) {
    const double a = H[i], b = H[i + 1];
    H[i] = cs[i] * a + sn[i] * b;
    H[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = H[col], b = H[col + 1];
  const double d = sqrt(a * a + b * b);
  const double c = d > 0.0 ? a / d : 1.0, s = d > 0.0 ? b / d : 0.0;
  cs[col] = c;
  sn[col] = s;
  H[col] = d;
  H[col + 1] = 0.0;
  const double gj = g[col];
  g[col] = c * gj;
  g[col + 1] = -s * gj;
  gm[gm_slot(m) + GM_JEND] = col + 1;
  gm[gm_hpass(m) + m] = hn; // the norm for the scaling of the next basis vector
  flags[F_ITER] = flags[F_ITER] + 1;
  if (!(fabs(g[col + 1]) > gm[gm_slot(m) + GM_TOL2]) || col + 1 == m || !(hn > 0.0) ||
      flags[F_ITER] >= flags[F_ITMAX]) {
    flags[F_INNER] = 1;
    flags[F_GATE] = 1;
  }
}

// v = w / ||w|| for the next basis vector
__global__ void k_gm_scale(int n, int m, const double *wv, double *v,
                           const double *__restrict__ gm, const int *__restrict__ flags) {
  if (flags[F_GATE]) return;
  const double inv = 1.0 / gm[gm_hpass(m) + m];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = wv[i] * inv;
}

// y = H^{-1} g by back substitution over the cycle's jend columns
__global__ void k_gm_solve(int m, double *__restrict__ gm, const int *__restrict__ flags) {
  if (flags[F_DONE] || threadIdx.x != 0) return;
  const int je = (int)gm[gm_slot(m) + GM_JEND];
  const double *H = gm + gm_h(m), *g = gm + gm_g(m);
  double *y = gm + gm_y(m);
  for (int i = je - 1; i >= 0; --i) {
    double s = g[i];
    for (int k = i + 1; k < je; ++k) s -= H[(size_t)k * (m + 1) + i] * y[k];
    const double d = H[(size_t)i * (m + 1) + i];
    y[i] = d != 0.0 ? s / d : 0.0;
  }
}

// u = sum_{j < jend} y_j V_j
__global__ void k_gm_combine(int n, int m, const double *__restrict__ V, size_t ldv, const double *__restrict__ gm,
                             double *__restrict__ u, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int je = (int)gm[gm_slot(m) + GM_JEND];
  double s = 0.0;
  for (int j = 0; j < je; ++j) s += gm[gm_y(m) + j] * V[(size_t)j * ldv + i];
  u[i] = s;
}

__global__ void k_gm_addx(int n, double *__restrict__ x, const double *__restrict__ z, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + z[i];
}

// r = b - A x with L1 and L2 partials
__global__ void k_gm_resid(int n, int w, const int *__restrict__ nbr, const double *__restrict__ ac,
                           const double *__restrict__ anb, const double *__restrict__ b,
                           const double *__restrict__ x, double *__restrict__ r, double *__restrict__ part,
                           const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  double s1 = 0.0, s2 = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double v = b[i] - ell_ax(i, n, w, nbr, ac, anb, x);
    r[i] = v;
    s1 += fabs(v);
    s2 += v * v;
  }
  s1 = block_sum<kBlock>(s1);
  s2 = block_sum<kBlock>(s2);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[kPartials + blockIdx.x] = s2;
  }
}

// start of a cycle: the reference's stop rule on the true L1 residual
// (CheckConvergenceRoutine.for semantics: fres <= rrf * ires), then
// beta = ||r||_2, g = (beta, 0, ...), H = 0, and the cycle's L2 target: the
// L2 residual scaled by the L1 reduction still needed, halved for safety
__global__ void k_gm_restart(int np, int m, const double *__restrict__ part, double *__restrict__ gm,
                             double *__restrict__ scal, int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const double l1 = reduce_partials(part, np);
  const double l2 = sqrt(reduce_partials(part + kPartials, np));
  if (threadIdx.x != 0) return;
  scal[S_FRES] = l1;
  const bool done = !(l1 > scal[S_RRF] * scal[S_IRES] && flags[F_ITER] < flags[F_ITMAX]);
  flags[F_DONE] = done;
  flags[F_INNER] = 0;
  flags[F_GATE] = done;
  for (size_t k = 0; k < gm_cs(m); ++k) gm[gm_h(m) + k] = 0.0;
  double *g = gm + gm_g(m);
  g[0] = l2;
  for (int k = 1; k <= m; ++k) g[k] = 0.0;
  gm[gm_slot(m) + GM_BETA] = l2;
  gm[gm_slot(m) + GM_JEND] = 0;
  gm[gm_slot(m) + GM_TOL2] = l1 > 0.0 ? 0.5 * l2 * scal[S_RRF] * scal[S_IRES] / l1 : 0.0;
  if (!(l2 > 0.0)) {
    flags[F_DONE] = 1;
    flags[F_GATE] = 1;
  }
}

// v_0 = r / beta
__global__ void k_gm_v0(int n, int m, const double *__restrict__ r, double *__restrict__ v,
                        const double *__restrict__ gm, const int *__restrict__ flags) {
  if (flags[F_DONE]) return;
  const double inv = 1.0 / gm[gm_slot(m) + GM_BETA];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = r[i] * inv;
}

// ---------------------------------------------------------------- finalisers
enum FinOp {
  FIN_INIT_NORMS = 0, // res4, ires, done-before-start
  FIN_PBCG_BK,        // bknum, bk, bkden
  FIN_PBCG_AK,        // ak
  FIN_CONV,           // fres from partials, iteration count, done
  FIN_BICG_RHO,       // rho, beta
  FIN_BICG_ALPHA,     // alpha = rho/(r0,v)
  FIN_BICG_OMEGA,     // omega = (t,s)/(t,t)
  FIN_TRUE_RES,       // fres only
  FIN_CG_RHO,         // rz = (r,z), beta
  FIN_CG_ALPHA,       // alpha = rz/(p,q)
  FIN_WARM,           // reject an initial guess whose L1 residual exceeds ires
  FIN_WARM_SCALE,     // gamma = (A x0, b)/(A x0, A x0)
  FIN_CONV_BICG_RHO   // FIN_CONV, then (unless done) FIN_BICG_RHO of the next iteration
};

// one workgroup of kFinBlock threads combines up to kPartials partials
// np <= kPartials = kFinR * kFinBlock: a thread's partials i, i + kFinBlock,
// ... are loaded together, then folded in index order (the slots past np add
// +0.0, or take the max with the starting -1.0, which changes nothing)
constexpr int kFinR = kPartials / kFinBlock;
static_assert(kFinR * kFinBlock == kPartials, "partials per finalise thread");
__device__ double reduce_partials(const double *part, int np) {
  double v[kFinR];
#pragma unroll
  for (int u = 0; u < kFinR; ++u) {
    const int i = threadIdx.x + u * kFinBlock;
    v[u] = i < np ? part[i] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < kFinR; ++u) s += v[u];
  return block_sum<kFinBlock>(s);
}
__device__ double reduce_partials_max(const double *part, int np) {
  double v[kFinR];
#pragma unroll
  for (int u = 0; u < kFinR; ++u) {
    const int i = threadIdx.x + u * kFinBlock;
    v[u] = i < np ? part[i] : -1.0;
  }
  double s = -1.0;
#pragma unroll
  for (int u = 0; u < kFinR; ++u) s = fmax(s, v[u]);
  return block_max<kFinBlock>(s);
}

// rho = (r0, r) and beta of BiCGSTAB.  Breakdown (rho or the last omega
// zero, or a non-finite beta, as on a singular system whose right-hand side
// carries round-off outside the range): stop with the current iterate.
__device__ void bicg_rho(double a, double *__restrict__ scal, int *__restrict__ flags) {
  if (flags[F_ITER] > 0) {
    const double beta = (a / scal[S_RHO]) * (scal[S_ALPHA] / scal[S_OMEGA]);
    if (scal[S_OMEGA] == 0.0 || !isfinite(beta)) {
      flags[F_DONE] = 1;
      return;
    }
    scal[S_BETA] = beta;
  }
  scal[S_RHO] = a;
  if (a == 0.0) flags[F_DONE] = 1;
}

__global__ void k_fin(int op, int np, int n, int nf, const double *__restrict__ part,
                      double *__restrict__ scal, int *__restrict__ flags) {
  if (op != FIN_INIT_NORMS && op != FIN_TRUE_RES && flags[F_DONE]) return;
  const double a = reduce_partials(part, np);
  double b = 0.0, c = 0.0, d = 0.0, e = 0.0;
  if (op == FIN_INIT_NORMS) {
    b = reduce_partials_max(part + kPartials, np);
    c = reduce_partials(part + 2 * kPartials, np);
    d = reduce_partials_max(part + 3 * kPartials, np);
    e = reduce_partials_max(part + 4 * kPartials, np);
  } else if (op == FIN_BICG_OMEGA || op == FIN_WARM_SCALE || op == FIN_CONV_BICG_RHO) {
    b = reduce_partials(part + kPartials, np);
  }
  if (threadIdx.x != 0) return;
  switch (op) {
  case FIN_INIT_NORMS: {
    scal[S_RES_ABS] = a;
    scal[S_RES_MAX] = b;
    scal[S_RES_RMS] = sqrt(c / n);
    scal[S_RES_SCALED] = (nf == 4) ? 0.0 : b / (d + 1.0e-15);
    scal[S_IRES] = a;
    scal[S_FRES] = a;
    flags[F_ITER] = 0;
    flags[F_DONE] = !(a > scal[S_RRF] * a && 0 < flags[F_ITMAX]);
    // refused input: nothing runs, the caller's arrays stay as they were
    flags[F_BAD] = e >= 2.0 ? 2 : (e >= 1.0 && flags[F_NEEDPOS]) ? 1 : 0;
    if (flags[F_BAD]) flags[F_DONE] = 1;
    break;
  }
  case FIN_PBCG_BK: {
    scal[S_BKNUM] = a;
    if (flags[F_ITER] > 0) scal[S_BK] = a / (scal[S_BKDEN] + 1.0e-10);
    scal[S_BKDEN] = a;
    break;
  }
  case FIN_PBCG_AK:
    scal[S_AKDEN] = a;
    scal[S_AK] = scal[S_BKNUM] / (a + 1.0e-10);
    break;
  case FIN_CONV: {
    scal[S_FRES] = a;
    const int it = flags[F_ITER] + 1;
    flags[F_ITER] = it;
    // NaN residuals end the loop as well
    flags[F_DONE] = !(a > scal[S_RRF] * scal[S_IRES] && it < flags[F_ITMAX]);
    break;
  }
  case FIN_BICG_RHO:
    bicg_rho(a, scal, flags);
    break;
  case FIN_CONV_BICG_RHO: {
    scal[S_FRES] = a;
    const int it = flags[F_ITER] + 1;
    flags[F_ITER] = it;
    flags[F_DONE] = !(a > scal[S_RRF] * scal[S_IRES] && it < flags[F_ITMAX]);
    if (!flags[F_DONE]) bicg_rho(b, scal, flags);
    break;
  }
  case FIN_BICG_ALPHA:
    scal[S_RV] = a;
    scal[S_ALPHA] = (a != 0.0) ? scal[S_RHO] / a : 0.0;
    break;
  case FIN_BICG_OMEGA:
    scal[S_TS] = a;
    scal[S_TT] = b;
    scal[S_OMEGA] = (b != 0.0) ? a / b : 0.0;
    break;
  case FIN_TRUE_RES:
    scal[S_FRES] = a;
    break;
  case FIN_CG_RHO: {
    if (flags[F_ITER] > 0) {
      const double beta = a / scal[S_RHO];
      if (!isfinite(beta)) {
        flags[F_DONE] = 1; // breakdown: keep the iterate
        break;
      }
      scal[S_BETA] = beta;
    }
    scal[S_RHO] = a;
    if (a == 0.0) flags[F_DONE] = 1;
    break;
  }
  case FIN_WARM:
    flags[F_REJECT] = !(a <= scal[S_IRES]);
    break;
  case FIN_WARM_SCALE:
    scal[S_GAMMA] = (b > 0.0 && isfinite(a / b)) ? a / b : 0.0;
    break;
  case FIN_CG_ALPHA:
    scal[S_RV] = a;
    scal[S_ALPHA] = (a != 0.0) ? scal[S_RHO] / a : 0.0;
    if (a == 0.0 || !isfinite(scal[S_ALPHA])) {
      scal[S_ALPHA] = 0.0;
      flags[F_DONE] = 1;
    }
    break;
  }
}

// ---------------------------------------------------------------- host side

static void precompute_transpose(fs_mesh *m, std::vector<int> &tlin, std::vector<int> &tsrc, int &wt) {
  const int n = m->n, w = m->w;
  std::vector<int> deg(n, 0);
  for (int i = 0; i < n; ++i)
    for (int s = 0; s < m->h_nnbr[i]; ++s) {
      const int k = m->h_nbr[(size_t)s * n + i];
      if (k > 0 && k != i + 1) deg[k - 1]++;
    }
  wt = 1;
  for (int i = 0; i < n; ++i) wt = std::max(wt, deg[i]);
  tlin.assign((size_t)wt * n, -1);
  tsrc.assign((size_t)wt * n, 0);
  std::fill(deg.begin(), deg.end(), 0);
  for (int i = 0; i < n; ++i) // source rows ascending, slots ascending
    for (int s = 0; s < m->h_nnbr[i]; ++s) {
      const int k = m->h_nbr[(size_t)s * n + i];
      if (k > 0 && k != i + 1) {
        const int j = k - 1, t = deg[j]++;
        tlin[(size_t)t * n + j] = s * n + i;
        tsrc[(size_t)t * n + j] = i;
      }
    }
  (void)w;
}

// Emulates the loops of FactorizeUsingILU.for:29-75 to list, per cell, the
// terms of its D-ILU diagonal in the order the reference applies them, and
// groups the cells into dependency levels.
static void precompute_ilu(fs_mesh *m) {
  const int n = m->n, w = m->w;
  const size_t total = (size_t)n * w;
  std::vector<std::vector<int>> ta(n), ti(n), tb(n);
  for (int i1 = 1; i1 <= n; ++i1) {
    if (i1 == n - 1) continue; // FactorizeUsingILU.for:36
    const int i3 = m->h_nnbr[i1 - 1];
    for (int j2 = 1; j2 <= i3; ++j2) {
      const int j4 = m->h_nbr[(size_t)(j2 - 1) * n + (i1 - 1)];
      if (!(j4 > i1 && j4 <= n)) continue;
      const int j3 = m->h_nnbr[j4 - 1];
      int i2 = 0, i4 = -1;
      while (i2 <= j3 && i4 != i1) {
        i2 = i2 + 1;
        const size_t lin = (size_t)(i2 - 1) * n + (j4 - 1);
        i4 = lin < total ? m->h_nbr[lin] : 0;
      }
      if (i4 == i1) {
        ta[j4 - 1].push_back((i2 - 1) * n + (j4 - 1));
        ti[j4 - 1].push_back(i1 - 1);
        tb[j4 - 1].push_back((j2 - 1) * n + (i1 - 1));
      }
    }
  }
  std::vector<int> level(n, 0), ptr(n + 1, 0), A, I, B;
  int maxl = 0;
  for (int j = 0; j < n; ++j) {
    int l = 0;
    for (int i : ti[j]) l = std::max(l, level[i] + 1);
    level[j] = l;
    maxl = std::max(maxl, l);
    ptr[j + 1] = ptr[j] + (int)ta[j].size();
    A.insert(A.end(), ta[j].begin(), ta[j].end());
    I.insert(I.end(), ti[j].begin(), ti[j].end());
    B.insert(B.end(), tb[j].begin(), tb[j].end());
  }
  m->ilu_nlevels = maxl + 1;
  m->ilu_level_ptr.assign(maxl + 2, 0);
  for (int j = 0; j < n; ++j) m->ilu_level_ptr[level[j] + 1]++;
  for (int l = 0; l <= maxl; ++l) m->ilu_level_ptr[l + 1] += m->ilu_level_ptr[l];
  std::vector<int> cells(n), fill(m->ilu_level_ptr.begin(), m->ilu_level_ptr.end() - 1);
  for (int j = 0; j < n; ++j) cells[fill[level[j]]++] = j;
  if (A.empty()) {
    A.push_back(0);
    I.push_back(0);
    B.push_back(0);
  }
  m->d_ilu_cells.upload(cells, m->stream);
  m->d_ilu_ptr.upload(ptr, m->stream);
  m->d_ilu_a.upload(A, m->stream);
  m->d_ilu_i.upload(I, m->stream);
  m->d_ilu_b.upload(B, m->stream);
}

// Greedy colouring in cell order: each cell takes the smallest colour not
// used by an already coloured neighbour.
void greedy_colors(int n, int w, const int *nbr0, std::vector<int> &ptr, std::vector<int> &cells) {
  std::vector<int> col(n, -1);
  int nc = 0;
  std::vector<int> used;
  for (int i = 0; i < n; ++i) {
    used.assign(nc + 1, 0);
    for (int j = 0; j < w; ++j) {
      const int k = nbr0[(size_t)j * n + i];
      if (k >= 0 && col[k] >= 0) used[col[k]] = 1;
    }
    int c = 0;
    while (c < nc && used[c]) ++c;
    col[i] = c;
    if (c == nc) ++nc;
  }
  ptr.assign(nc + 1, 0);
  for (int i = 0; i < n; ++i) ptr[col[i] + 1]++;
  for (int c = 0; c < nc; ++c) ptr[c + 1] += ptr[c];
  cells.resize(n);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < n; ++i) cells[fill[col[i]]++] = i;
}

// Level-scheduled triangular sweeps of the reference's ILU iteration
// (ILUMethod.for:45-140): per cell the contributions in the reference's
// order, cells grouped in dependency levels.
struct IluSweep {
  bool ready = false;
  std::vector<int> fwd_level_ptr, bwd_level_ptr;
  DevBuf<int> fwd_cells, fwd_ptr, fwd_lin, fwd_src, bwd_cells, bwd_ptr, bwd_lin, bwd_src;
};

struct SolveWork {
  IluSweep ilu;
  int gm_m = 0;
  DevBuf<double> gm_V, gm_state, gm_part;
  int wt = 1;
  DevBuf<int> tlin, tsrc;
  std::vector<int> color_ptr;
  DevBuf<int> color_cells;
  DevBuf<double> dc;
};

static SolveWork &work(fs_mesh *m) {
  static_assert(sizeof(SolveWork) > 0, "");
  return *reinterpret_cast<SolveWork *>(m->solve_work);
}

// Lists of the forward substitution (ILUMethod.for:45-93: rc(j) -=
// anb(j, slot of i) * dc(i) * rc(i) for the lower neighbours i of j, i
// increasing) and of the backward one (:97-138: rc(i) -= anb(i, j2) * rc(j)
// for the upper neighbours j of i in slot order), with their levels.
static void levels_of(int n, const std::vector<int> &ptr, const std::vector<int> &src, bool up,
                      std::vector<int> &level_ptr, std::vector<int> &cells) {
  std::vector<int> level(n, 0);
  int maxl = 0;
  for (int q = 0; q < n; ++q) {
    const int i = up ? n - 1 - q : q; // dependencies are processed first
    int l = 0;
    for (int t = ptr[i]; t < ptr[i + 1]; ++t) l = std::max(l, level[src[t]] + 1);
    level[i] = l;
    maxl = std::max(maxl, l);
  }
  level_ptr.assign(maxl + 2, 0);
  for (int i = 0; i < n; ++i) level_ptr[level[i] + 1]++;
  for (int l = 0; l <= maxl; ++l) level_ptr[l + 1] += level_ptr[l];
  cells.assign(n, 0);
  std::vector<int> fill(level_ptr.begin(), level_ptr.end() - 1);
  for (int i = 0; i < n; ++i) cells[fill[level[i]]++] = i;
}

static void precompute_ilu_sweep(fs_mesh *m, IluSweep &S) {
  const int n = m->n;
  const std::vector<int> &nb = m->h_nbr, &nn = m->h_nnbr;
  std::vector<std::vector<int>> flin(n), fsrc(n);
  for (int i1 = 1; i1 <= n; ++i1) {
    for (int j2 = 1; j2 <= nn[i1 - 1]; ++j2) {
      const int j1 = nb[(size_t)(j2 - 1) * n + (i1 - 1)];
      if (!(j1 != 0 && j1 > i1 && j1 <= n)) continue;
      const int j3 = nn[j1 - 1];
      int i2 = 0, k = 0;
      while (i2 + 1 <= j3 && k != i1) {
        i2 = i2 + 1;
        k = nb[(size_t)(i2 - 1) * n + (j1 - 1)];
      }
      if (k != 0 && k == i1) {
        flin[j1 - 1].push_back((i2 - 1) * n + (j1 - 1));
        fsrc[j1 - 1].push_back(i1 - 1);
      }
    }
  }
  std::vector<int> fptr(n + 1, 0), fl, fs, bptr(n + 1, 0), bl, bs;
  for (int j = 0; j < n; ++j) {
    fptr[j + 1] = fptr[j] + (int)flin[j].size();
    fl.insert(fl.end(), flin[j].begin(), flin[j].end());
    fs.insert(fs.end(), fsrc[j].begin(), fsrc[j].end());
  }
  for (int i1 = 1; i1 <= n; ++i1) {
    if (i1 < n)
      for (int j2 = 1; j2 <= nn[i1 - 1]; ++j2) {
        const int j = nb[(size_t)(j2 - 1) * n + (i1 - 1)];
        if (j > i1) {
          bl.push_back((j2 - 1) * n + (i1 - 1));
          bs.push_back(j - 1);
        }
      }
    bptr[i1] = (int)bl.size();
  }
  std::vector<int> fcells, bcells;
  levels_of(n, fptr, fs, false, S.fwd_level_ptr, fcells);
  levels_of(n, bptr, bs, true, S.bwd_level_ptr, bcells);
  for (auto *v : {&fl, &fs, &bl, &bs})
    if (v->empty()) v->push_back(0);
  S.fwd_cells.upload(fcells, m->stream);
  S.fwd_ptr.upload(fptr, m->stream);
  S.fwd_lin.upload(fl, m->stream);
  S.fwd_src.upload(fs, m->stream);
  S.bwd_cells.upload(bcells, m->stream);
  S.bwd_ptr.upload(bptr, m->stream);
  S.bwd_lin.upload(bl, m->stream);
  S.bwd_src.upload(bs, m->stream);
  S.ready = true;
}

int mesh_prepare_solver(fs_mesh *m) {
  auto *sw = new SolveWork();
  m->solve_work = sw;
  std::vector<int> tlin, tsrc;
  precompute_transpose(m, tlin, tsrc, sw->wt);
  sw->tlin.upload(tlin, m->stream);
  sw->tsrc.upload(tsrc, m->stream);
  precompute_ilu(m);
  std::vector<int> nbr0((size_t)m->n * m->w);
  for (size_t t = 0; t < nbr0.size(); ++t) nbr0[t] = m->h_nbr_dev[t];
  std::vector<int> cells;
  greedy_colors(m->n, m->w, nbr0.data(), sw->color_ptr, cells);
  sw->color_cells.upload(cells, m->stream);
  sw->dc.alloc(m->n);
  for (auto &v : m->v) v.alloc(m->n);
  m->partial.alloc(5 * kPartials);
  m->scal.alloc(S_NSLOTS);
  m->flags.alloc(F_NFLAGS);
  for (int k = 0; k < 12; ++k) m->wk.v[k] = m->v[k].p;
  m->wk.dc = sw->dc.p;
  m->wk.partial = m->partial.p;
  m->wk.scal = m->scal.p;
  m->wk.flags = m->flags.p;
  m->nnz = 0;
  for (int i = 0; i < m->n; ++i)
    for (int j = 0; j < m->w; ++j)
      if (m->h_nbr_dev[(size_t)j * m->n + i] >= 0) ++m->nnz;
  FS_HIP_THROW(hipStreamSynchronize(m->stream));
  return 0;
}

int kernel_clock_start(fs_mesh *m, int site_mask, int capacity) {
  KernelClock &c = m->kclock;
  const int valid = (1 << FS_CLOCK_SPMV) | (1 << FS_CLOCK_FINE_SWEEP);
  if (site_mask & ~valid) {
    set_error("fs_kernel_clock_start: unknown site in mask " + std::to_string(site_mask));
    return -1;
  }
  if (site_mask == 0) {
    c.sites = 0;
    return 0;
  }
  if (capacity < 1 || capacity > (1 << 20)) {
    set_error("fs_kernel_clock_start: capacity must lie in [1, 2^20]");
    return -1;
  }
  while ((int)c.ev.size() < 2 * capacity) {
    hipEvent_t e;
    FS_HIP(hipEventCreate(&e));
    c.ev.push_back(e);
  }
  c.sites = site_mask;
  c.cap = capacity;
  c.used = 0;
  for (int &d : c.dropped) d = 0;
  for (long long &k : c.seen) k = 0;
  c.solve_first = 0;
  c.iter.assign(capacity, -1);
  c.site.assign(capacity, 0);
  c.bytes.assign(capacity, 0.0);
  c.live.assign(capacity, 0);
  return 0;
}

int kernel_clock_read(fs_mesh *m, int site, double *out) {
  KernelClock &c = m->kclock;
  if (site < 1 || site > 3) {
    set_error("fs_kernel_clock_read: unknown site " + std::to_string(site));
    return -1;
  }
  FS_HIP(hipStreamSynchronize(m->stream));
  long long k = 0, idle = 0;
  double ms = 0.0, by = 0.0;
  for (int u = 0; u < c.used; ++u) {
    if (c.site[u] != site) continue;
    if (!c.live[u]) {
      ++idle;
      continue;
    }
    float t = 0.f;
    FS_HIP(hipEventElapsedTime(&t, c.ev[2 * u], c.ev[2 * u + 1]));
    ms += t;
    by += c.bytes[u];
    ++k;
  }
  out[0] = (double)k;
  out[1] = k ? ms / k : 0.0;
  out[2] = k ? by / k : 0.0;
  out[3] = (double)(c.dropped[site & 3] + idle);
  return 0;
}

int kernel_clock_clear(fs_mesh *m) {
  KernelClock &c = m->kclock;
  FS_HIP(hipStreamSynchronize(m->stream));
  c.used = 0;
  for (int &d : c.dropped) d = 0;
  for (long long &k : c.seen) k = 0;
  c.solve_first = 0;
  return 0;
}

int kernel_clock_stride(fs_mesh *m, int stride) {
  if (stride < 1 || stride > 1024) {
    set_error("fs_kernel_clock_stride: stride must lie in [1, 1024]");
    return -1;
  }
  m->kclock.stride = stride;
  for (long long &k : m->kclock.seen) k = 0;
  return 0;
}

void kernel_clock_release(fs_mesh *m) {
  for (hipEvent_t e : m->kclock.ev) (void)hipEventDestroy(e);
  m->kclock = KernelClock{};
}

void mesh_release_solver(fs_mesh *m) {
  kernel_clock_release(m);
  delete reinterpret_cast<SolveWork *>(m->solve_work);
  m->solve_work = nullptr;
}

// D-ILU diagonal on the device (exact reference order).
static void ilu_factor(fs_mesh *m, const double *ac, const double *anb, double *dc) {
  for (int l = 0; l < m->ilu_nlevels; ++l) {
    const int b = m->ilu_level_ptr[l], e = m->ilu_level_ptr[l + 1];
    if (e > b)
      hipLaunchKernelGGL(k_ilu_level, dim3(grid_for(e - b)), dim3(kBlock), 0, m->stream, e - b,
                         m->d_ilu_cells.p + b, m->d_ilu_ptr.p, m->d_ilu_a.p, m->d_ilu_i.p,
                         m->d_ilu_b.p, ac, anb, dc);
  }
}

static void host_flags(fs_mesh *m, int *fl) {
  FS_HIP_THROW(hipMemcpyAsync(fl, m->wk.flags, sizeof(int) * F_NFLAGS, hipMemcpyDeviceToHost, m->stream));
  FS_HIP_THROW(hipStreamSynchronize(m->stream));
}

// Runs `body` in batches until the device sets done.  The first batch is the
// iteration count of the previous solve of the same variable and solver (the
// SIMPLE loop solves nearly the same system every outer iteration; at most
// 4 check_every), later batches grow 2, 4, ... up to check_every; without a
// previous solve every batch is check_every long.  Iterations launched past
// convergence return at once, so the batch sizes only trade idle launches
// against host checks.
template <class F> static int run_batches(fs_mesh *m, int itmax, int key, F body) {
  int fl[F_NFLAGS] = {0};
  int launched = 0;
  int &pred = m->iter_pred[key & 255];
  int next = pred > 0 ? std::min(pred, 4 * m->check_every) : m->check_every, grow = 2;
  m->kclock.solve_first = m->kclock.used;
  while (!fl[F_DONE] && launched < itmax) {
    const int batch = std::min(next, itmax - launched);
    for (int b = 0; b < batch; ++b) {
      m->kclock.cur_iter = launched + b;
      body();
    }
    launched += batch;
    host_flags(m, fl);
    next = pred > 0 ? std::min(grow, m->check_every) : m->check_every;
    // stop doubling at check_every: an unbounded doubling overflows int
    // after 31 batches
    if (grow < m->check_every) grow *= 2;
  }
  pred = fl[F_ITER];
  // launches of iterations past the converged one returned at once
  KernelClock &c = m->kclock;
  for (int u = c.solve_first; u < c.used; ++u)
    if (c.iter[u] >= fl[F_ITER]) c.live[u] = 0;
  c.cur_iter = -1;
  return fl[F_ITER];
}

// x0 *= (A x0, b)/(A x0, A x0): the scaling of an initial guess with the
// smallest L2 residual (0 when A x0 = 0), so the guess is never worse than
// a zero start in that norm.  v is scratch.
static void scale_guess(fs_mesh *m, const double *ac, const double *anb, const double *bc, double *x, double *v) {
  const int n = m->n, gr = grid_red(n);
  hipStream_t s = m->stream;
  hipLaunchKernelGGL(k_spmv_dot_ts, dim3(gr), dim3(kBlock), 0, s, n, m->w, EllNbr(m->d_nbr.p, m->d_nbr16.p), ac, anb, x,
                     v, bc, m->wk.partial, m->wk.flags);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_WARM_SCALE, gr, n, 4, m->wk.partial, m->wk.scal,
                     m->wk.flags);
  hipLaunchKernelGGL(k_warm_scale, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, m->wk.scal, m->wk.flags);
}

// ---------------------------------------------------------------- projection
// Projection of the right-hand side onto earlier solutions (Fischer, "Projection
// techniques for iterative solution of Ax = b with successive right-hand
// sides", CMAME 163, 1998): x0 = sum_a c_a x_a with the c that minimise
// |b - A x0|_2, i.e. G c = h with G_ab = (A x_a, A x_b), h_a = (A x_a, b).

constexpr int kProjGrid = 2048; // blocks of the Gram pass: 14 partial rows fit the partial buffer

struct ProjVecs {
  const double *p[kProjMax];
};
struct ProjCoef {
  double c[kProjMax];
};

// Partials of the upper triangle of G and of h, row u of the order
// (G_00, h_0, G_01, ..., G_11, h_1, ...) at part[u * gridDim.x + block].
template <int K>
__global__ void __launch_bounds__(kBlock) k_proj_gram(int n, ProjVecs y, const double *__restrict__ b,
                                                      double *__restrict__ part) {
  constexpr int NS = K * (K + 1) / 2 + K;
  double s[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) s[u] = 0.0;
  for (int i = xcd_block() * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double v[K];
#pragma unroll
    for (int a = 0; a < K; ++a) v[a] = y.p[a][i];
    const double bi = b[i];
    int u = 0;
#pragma unroll
    for (int a = 0; a < K; ++a) {
#pragma unroll
      for (int c = a; c < K; ++c) s[u++] += v[a] * v[c];
      s[u++] += v[a] * bi;
    }
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const double t = block_sum<kBlock>(s[u]);
    if (threadIdx.x == 0) part[(size_t)u * gridDim.x + blockIdx.x] = t;
  }
}

// out[u] = sum of row u of the partials, in a fixed order
__global__ void __launch_bounds__(kBlock) k_proj_fin(int ns, int gp, const double *__restrict__ part,
                                                     double *__restrict__ out) {
  for (int u = 0; u < ns; ++u) {
    double t = 0.0;
    for (int j = threadIdx.x; j < gp; j += kBlock) t += part[(size_t)u * gp + j];
    t = block_sum<kBlock>(t);
    if (threadIdx.x == 0) out[u] = t;
  }
}

__global__ void k_proj_comb(int n, int k, ProjVecs x, ProjCoef c, double *__restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int a = 0; a < k; ++a) s += c.c[a] * x.p[a][i];
  out[i] = s;
}

int project_guess(fs_mesh *m, const double *ac, const double *anb, const double *bc, const double *const *X, int k,
                  double *x0) {
  if (k > kProjMax) k = kProjMax;
  if (k <= 0) return 0;
  const int n = m->n;
  hipStream_t s = m->stream;
  ProjVecs yv{}, xv{};
  for (int a = 0; a < k; ++a) {
    yv.p[a] = m->wk.v[a];
    xv.p[a] = X[a];
    hipLaunchKernelGGL(k_spmv, dim3(grid_for(n)), dim3(kBlock), 0, s, n, m->w, m->d_nbr.p, ac, anb, X[a],
                       m->wk.v[a]);
  }
  const int ns = k * (k + 1) / 2 + k;
  double *part = m->wk.partial, *out = m->wk.partial + (size_t)ns * kProjGrid;
  switch (k) {
  case 1: hipLaunchKernelGGL(k_proj_gram<1>, dim3(kProjGrid), dim3(kBlock), 0, s, n, yv, bc, part); break;
  case 2: hipLaunchKernelGGL(k_proj_gram<2>, dim3(kProjGrid), dim3(kBlock), 0, s, n, yv, bc, part); break;
  case 3: hipLaunchKernelGGL(k_proj_gram<3>, dim3(kProjGrid), dim3(kBlock), 0, s, n, yv, bc, part); break;
  default: hipLaunchKernelGGL(k_proj_gram<4>, dim3(kProjGrid), dim3(kBlock), 0, s, n, yv, bc, part); break;
  }
  hipLaunchKernelGGL(k_proj_fin, dim3(1), dim3(kBlock), 0, s, ns, kProjGrid, part, out);
  double g[kProjMax * (kProjMax + 1) / 2 + kProjMax];
  FS_HIP_THROW(hipMemcpyAsync(g, out, sizeof(double) * ns, hipMemcpyDeviceToHost, s));
  FS_HIP_THROW(hipStreamSynchronize(s));
  double G[kProjMax][kProjMax] = {}, h[kProjMax] = {};
  for (int a = 0, u = 0; a < k; ++a) {
    for (int c = a; c < k; ++c) G[a][c] = G[c][a] = g[u++];
    h[a] = g[u++];
  }
  // Cholesky of G over the vectors in order; a vector whose part outside
  // the span of those before it is below 1e-10 of its norm (or not finite)
  // is left out, so the earliest vectors (the latest solutions) are kept
  double L[kProjMax][kProjMax] = {}, z[kProjMax] = {};
  bool use[kProjMax] = {};
  int used = 0;
  for (int j = 0; j < k; ++j) {
    double d = G[j][j];
    for (int p = 0; p < j; ++p) d -= L[j][p] * L[j][p];
    if (!(G[j][j] > 0.0) || !std::isfinite(G[j][j]) || !(d > 1e-10 * G[j][j])) continue;
    use[j] = true;
    ++used;
    L[j][j] = std::sqrt(d);
    for (int i = j + 1; i < k; ++i) {
      double t = G[i][j];
      for (int p = 0; p < j; ++p) t -= L[i][p] * L[j][p];
      L[i][j] = t / L[j][j];
    }
  }
  for (int j = 0; j < k; ++j) {
    if (!use[j]) continue;
    double t = h[j];
    for (int p = 0; p < j; ++p) t -= L[j][p] * z[p];
    z[j] = t / L[j][j];
  }
  ProjCoef cf{};
  for (int j = k - 1; j >= 0; --j) {
    if (!use[j]) continue;
    double t = z[j];
    for (int i = j + 1; i < k; ++i) t -= L[i][j] * cf.c[i];
    cf.c[j] = t / L[j][j];
  }
  hipLaunchKernelGGL(k_proj_comb, dim3(grid_for(n)), dim3(kBlock), 0, s, n, k, xv, cf, x0);
  return used;
}

int solve_device(fs_mesh *m, int nf, double *fit, double rrf, int itmax, int solver, int lmg,
                 const double *ac, const double *anb, const double *bc, double *dphi, double *res4,
                 fs_solve_stats *st, bool warm) {
  const int n = m->n, w = m->w;
  hipStream_t s = m->stream;
  auto t0 = std::chrono::steady_clock::now();
  const int gr = grid_red(n), ge = grid_for(n);
  double *part = m->wk.partial, *scal = m->wk.scal;
  int *flags = m->wk.flags;
  const int *nbr = m->d_nbr.p;
  const EllNbr nb(m->d_nbr.p, m->d_nbr16.p); // the row loops of the hot kernels
  SolveWork &sw = work(m);

  double h_sc[S_NSLOTS] = {0};
  h_sc[S_RRF] = rrf;
  if (solver < FS_SOR || solver > FS_CG || solver == FS_DIRECT) {
    set_error("unsupported solver code " + std::to_string(solver) +
              (solver == FS_DIRECT ? " ('direct' is the host's dense LU)" : ""));
    return -3;
  }
  int h_fl[F_NFLAGS] = {0, 0, itmax, 0};
  // Gauss-Seidel, D-ILU (also PBCG's preconditioner) and every multigrid
  // smoother divide by the diagonal
  h_fl[F_NEEDPOS] = (lmg != FS_MG_OFF || solver == FS_SOR || solver == FS_ILU || solver == FS_PBCG) ? 1 : 0;
  FS_HIP(hipMemcpyAsync(scal, h_sc, sizeof(h_sc), hipMemcpyHostToDevice, s));
  FS_HIP(hipMemcpyAsync(flags, h_fl, sizeof(h_fl), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_init_norms, dim3(gr), dim3(kBlock), 0, s, n, ac, bc, fit, part);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_INIT_NORMS, gr, n, nf, part, scal, flags);
  if (!warm) hipLaunchKernelGGL(k_zero_start, dim3(ge), dim3(kBlock), 0, s, n, dphi, (const int *)flags);
  auto refuse = [](int bad) {
    set_error(bad == 2 ? "a coefficient ac or source bc is not finite"
                       : "a diagonal coefficient ac <= 0, which this solver divides by");
    return bad == 2 ? -4 : -1;
  };
  if (lmg != FS_MG_OFF) {
    // the hierarchy is built from the operator: refuse bad input before that
    int fl0[F_NFLAGS];
    host_flags(m, fl0);
    if (fl0[F_BAD]) return refuse(fl0[F_BAD]);
  }

  int iters = 0;
  double *V0 = m->wk.v[0], *V1 = m->wk.v[1], *V2 = m->wk.v[2], *V3 = m->wk.v[3], *V4 = m->wk.v[4],
         *V5 = m->wk.v[5], *V6 = m->wk.v[6], *V7 = m->wk.v[7];

  // An initial guess, whatever the solver: scaled to its multiple with the
  // smallest L2 residual, then dropped for the zero start when its L1
  // residual still exceeds that of the zero start (|bc|_1).  Every solver
  // below starts from dphi as it stands, so the guess can only change the
  // iteration count, never the stop rule.
  if (warm) {
    scale_guess(m, ac, anb, bc, dphi, V3);
    hipLaunchKernelGGL(k_resid_l1, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, part, flags);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_WARM, gr, n, nf, part, scal, flags);
    hipLaunchKernelGGL(k_warm_fix, dim3(ge), dim3(kBlock), 0, s, n, dphi, (const int *)flags);
  }

  if (lmg == FS_MG_SOLVER) {
    iters = amg_solve(m, ac, anb, bc, dphi, solver, warm);
  } else if (solver == FS_PBCG) {
    // r=V0 rr=V1 p=V2 pp=V3 z=V4 zz=V5 (PBCG1.F90:4)
    ilu_factor(m, ac, anb, m->wk.dc);
    hipLaunchKernelGGL(k_spmv, dim3(ge), dim3(kBlock), 0, s, n, w, nbr, ac, anb, dphi, V0);
    hipLaunchKernelGGL(k_pbcg_start, dim3(ge), dim3(kBlock), 0, s, n, bc, V0, V1);
    hipLaunchKernelGGL(k_spmv, dim3(ge), dim3(kBlock), 0, s, n, w, nbr, ac, anb, V0, V1);
    // with FS_MG_PRECOND the D-ILU asolve is replaced by one AMG V-cycle for
    // both r and rr (the transpose preconditioner is the cycle itself: exact
    // for symmetric operators such as p', an approximation otherwise)
    const bool pre = (lmg == FS_MG_PRECOND);
    if (pre) amg_prepare(m, ac, anb);
    const double alpha = m->mg_alpha > 0.0 ? m->mg_alpha : 1.7;
    const double *dc = pre ? nullptr : m->wk.dc;
    iters = run_batches(m, itmax, nf * 8 + solver, [&]() {
      if (pre) {
        amg_apply(m, V0, V6, flags, alpha);
        amg_apply(m, V1, V7, flags, alpha);
      }
      hipLaunchKernelGGL(k_pbcg_bknum, dim3(gr), dim3(kBlock), 0, s, n, V0, V1, dc, V6, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_PBCG_BK, gr, n, nf, part, scal, flags);
      hipLaunchKernelGGL(k_pbcg_p, dim3(ge), dim3(kBlock), 0, s, n, V0, V1, dc, V6, V7, V2, V3, scal, flags);
      hipLaunchKernelGGL(k_pbcg_ap, dim3(gr), dim3(kBlock), 0, s, n, w, sw.wt, nbr, sw.tlin.p,
                         sw.tsrc.p, ac, anb, V2, V3, V4, V5, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_PBCG_AK, gr, n, nf, part, scal, flags);
      hipLaunchKernelGGL(k_pbcg_axpy, dim3(ge), dim3(kBlock), 0, s, n, dphi, V0, V1, V2, V4, V5, scal, flags);
      hipLaunchKernelGGL(k_resid_l1, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CONV, gr, n, nf, part, scal, flags);
    });
  } else if (solver == FS_SOR) {
    // multicolour symmetric Gauss-Seidel: colours forward then backward.  A
    // colour is an independent set, so its update reads only the other
    // colours, and a colour swept twice in a row would recompute the same
    // values bitwise: the backward pass starts at colour nc-2, and every
    // forward pass after the first at colour 1 (the residual check between
    // iterations leaves x alone).
    bool first = true;
    iters = run_batches(m, itmax, nf * 8 + solver, [&]() {
      const int nc = (int)sw.color_ptr.size() - 1;
      for (int c = first ? 0 : 1; c < nc; ++c) {
        const int b = sw.color_ptr[c], e = sw.color_ptr[c + 1];
        hipLaunchKernelGGL(k_gs_color, dim3(grid_for(e - b)), dim3(kBlock), 0, s, e - b,
                           sw.color_cells.p + b, n, w, nb, ac, anb, bc, dphi, flags);
      }
      for (int c = nc - 2; c >= 0; --c) {
        const int b = sw.color_ptr[c], e = sw.color_ptr[c + 1];
        hipLaunchKernelGGL(k_gs_color, dim3(grid_for(e - b)), dim3(kBlock), 0, s, e - b,
                           sw.color_cells.p + b, n, w, nb, ac, anb, bc, dphi, flags);
      }
      first = false;
      hipLaunchKernelGGL(k_resid_l1, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CONV, gr, n, nf, part, scal, flags);
    });
  } else if (solver == FS_BICGS) {
    // right-preconditioned BiCGSTAB; r=V0 r0=V1 p=V2 v=V3 s=V4 t=V5 phat=V6 shat=V7
    const bool pre = (lmg == FS_MG_PRECOND);
    if (pre) amg_prepare(m, ac, anb);
    const double alpha = m->mg_alpha > 0.0 ? m->mg_alpha : 1.7;
    // algorithmic bytes of y = A x with the (r0, y) partials: ac, the
    // neighbour words (4 per slot, or 4 per pair of slots), anb, x, y, r0
    const double ib = nb.d16 ? 4.0 * ((w + 1) / 2) : 4.0 * w;
    const double spmv_bytes = 8.0 * n + ib * n + 8.0 * m->nnz + 24.0 * n;
    hipLaunchKernelGGL(k_resid_vec, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, V0, V1, part);
    hipLaunchKernelGGL(k_zero, dim3(ge), dim3(kBlock), 0, s, n, V3);
    // rho of the first iteration; every later rho comes with the update
    hipLaunchKernelGGL(k_dot2, dim3(gr), dim3(kBlock), 0, s, n, V1, V0, nullptr, nullptr, part, flags);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_BICG_RHO, gr, n, nf, part, scal, flags);
    iters = run_batches(m, itmax, nf * 8 + solver, [&]() {
      hipLaunchKernelGGL(k_bicg_p, dim3(ge), dim3(kBlock), 0, s, n, V0, V3, V2, scal, flags);
      if (pre) amg_apply(m, V2, V6, flags, alpha);
      const double *ph = pre ? V6 : V2;
      timed_launch(m, FS_CLOCK_SPMV, spmv_bytes, k_spmv_dot, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, ph,
                   V3, (const double *)V1, part, (const int *)flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_BICG_ALPHA, gr, n, nf, part, scal, flags);
      hipLaunchKernelGGL(k_bicg_s, dim3(ge), dim3(kBlock), 0, s, n, V0, V3, V4, scal, flags);
      if (pre) amg_apply(m, V4, V7, flags, alpha);
      const double *sh = pre ? V7 : V4;
      hipLaunchKernelGGL(k_spmv_dot_ts, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, sh, V5, V4, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_BICG_OMEGA, gr, n, nf, part, scal, flags);
      hipLaunchKernelGGL(k_bicg_update, dim3(gr), dim3(kBlock), 0, s, n, dphi, V0, ph, sh, V4, V5, (const double *)V1,
                         scal, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CONV_BICG_RHO, gr, n, nf, part, scal,
                         flags);
    });
  } else if (solver == FS_CG) {
    // preconditioned conjugate gradient; r=V0 z=V1 p=V2 q=V3
    const bool pre = (lmg == FS_MG_PRECOND);
    if (pre) amg_prepare(m, ac, anb);
    const double alpha = m->mg_alpha > 0.0 ? m->mg_alpha : 1.7;
    // y = A p with the (p, y) partials: ac, nbr, anb, p, y
    const double spmv_bytes = 8.0 * n + (nb.d16 ? 4.0 * ((w + 1) / 2) : 4.0 * w) * n + 8.0 * m->nnz + 16.0 * n;
    hipLaunchKernelGGL(k_resid_vec, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, V0, nullptr,
                       part);
    iters = run_batches(m, itmax, nf * 8 + solver, [&]() {
      const double *z = V0;
      if (pre) {
        amg_apply(m, V0, V1, flags, alpha);
        z = V1;
      }
      hipLaunchKernelGGL(k_dot2, dim3(gr), dim3(kBlock), 0, s, n, V0, z, nullptr, nullptr, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CG_RHO, gr, n, nf, part, scal, flags);
      hipLaunchKernelGGL(k_cg_p, dim3(ge), dim3(kBlock), 0, s, n, z, V2, scal, flags);
      timed_launch(m, FS_CLOCK_SPMV, spmv_bytes, k_spmv_dot, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, V2,
                   V3, (const double *)V2, part, (const int *)flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CG_ALPHA, gr, n, nf, part, scal, flags);
      hipLaunchKernelGGL(k_cg_update, dim3(gr), dim3(kBlock), 0, s, n, dphi, V0, V2, V3, scal, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CONV, gr, n, nf, part, scal, flags);
    });
  } else if (solver == FS_GMRES) {
    // restarted GMRES(m), right-preconditioned: basis V (m+1 vectors), z = M^-1 v_j
    // (V10), u = V y (V11), r (V0); the cycle ends when its L2 estimate meets the
    // target, the basis is full or itmax is reached; the stop rule is checked on
    // the true L1 residual at every restart
    const bool pre = (lmg == FS_MG_PRECOND);
    if (pre) amg_prepare(m, ac, anb);
    const double alpha = m->mg_alpha > 0.0 ? m->mg_alpha : 1.7;
    const int gmm = std::max(1, std::min(64, m->gmres_m));
    if (sw.gm_m != gmm || sw.gm_V.n < (size_t)(gmm + 1) * n) {
      sw.gm_V.alloc((size_t)(gmm + 1) * n);
      sw.gm_state.alloc(gm_size(gmm));
      sw.gm_part.alloc((size_t)(gmm + 1) * kPartials);
      sw.gm_m = gmm;
    }
    double *V = sw.gm_V.p, *gm = sw.gm_state.p, *gp = sw.gm_part.p;
    double *Z = m->wk.v[10], *U = m->wk.v[11];
    const size_t ldv = (size_t)n;
    auto restart = [&]() {
      hipLaunchKernelGGL(k_gm_resid, dim3(gr), dim3(kBlock), 0, s, n, w, nbr, ac, anb, bc, dphi, V0, part, flags);
      hipLaunchKernelGGL(k_gm_restart, dim3(1), dim3(kFinBlock), 0, s, gr, gmm, part, gm, scal, flags);
      hipLaunchKernelGGL(k_gm_v0, dim3(ge), dim3(kBlock), 0, s, n, gmm, V0, V, gm, flags);
    };
    restart();
    int fl[F_NFLAGS];
    host_flags(m, fl);
    m->kclock.solve_first = m->kclock.used;
    while (!fl[F_DONE]) {
      for (int j = 0; j < gmm; ++j) {
        const double *vj = V + (size_t)j * ldv;
        double *wv = V + (size_t)(j + 1) * ldv;
        const double *zj = vj;
        if (pre) {
          amg_apply(m, vj, Z, flags + F_GATE, alpha);
          zj = Z;
        }
        hipLaunchKernelGGL(k_gm_spmv, dim3(ge), dim3(kBlock), 0, s, n, w, nbr, ac, anb, zj, wv, flags);
        for (int pass = 0; pass < 2; ++pass) {
          hipLaunchKernelGGL(k_gm_dots, dim3(gr, j + 1), dim3(kBlock), 0, s, n, V, ldv, wv, gp, flags);
          hipLaunchKernelGGL(k_gm_hfin, dim3(j + 1), dim3(kFinBlock), 0, s, gr, gmm, j, gp, gm, flags);
          hipLaunchKernelGGL(k_gm_maxpy, dim3(ge), dim3(kBlock), 0, s, n, j + 1, gmm, V, ldv, gm, wv, flags);
        }
        hipLaunchKernelGGL(k_gm_norm2, dim3(gr), dim3(kBlock), 0, s, n, wv, gp, flags);
        hipLaunchKernelGGL(k_gm_col, dim3(1), dim3(kFinBlock), 0, s, gr, gmm, j, gp, gm, flags);
        hipLaunchKernelGGL(k_gm_scale, dim3(ge), dim3(kBlock), 0, s, n, gmm, wv, wv, gm, flags);
      }
      hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, s, gmm, gm, flags);
      hipLaunchKernelGGL(k_gm_combine, dim3(ge), dim3(kBlock), 0, s, n, gmm, V, ldv, gm, U, flags);
      const double *zu = U;
      if (pre) {
        amg_apply(m, U, Z, flags, alpha);
        zu = Z;
      }
      hipLaunchKernelGGL(k_gm_addx, dim3(ge), dim3(kBlock), 0, s, n, dphi, zu, flags);
      restart();
      host_flags(m, fl);
    }
    iters = fl[F_ITER];
  } else if (solver == FS_ILU) {
    // the reference's ILU iteration: exact D-ILU diagonal, then per iteration
    // the residual and the two level-scheduled substitutions
    if (!sw.ilu.ready) precompute_ilu_sweep(m, sw.ilu);
    ilu_factor(m, ac, anb, m->wk.dc);
    IluSweep &S = sw.ilu;
    double *rcv = m->wk.v[0];
    iters = run_batches(m, itmax, nf * 8 + solver, [&]() {
      hipLaunchKernelGGL(k_ilu_resid, dim3(ge), dim3(kBlock), 0, s, n, w, nbr, ac, anb, bc, dphi, rcv, flags);
      for (size_t l = 0; l + 1 < S.fwd_level_ptr.size(); ++l) {
        const int b0 = S.fwd_level_ptr[l], b1 = S.fwd_level_ptr[l + 1];
        if (b1 > b0)
          hipLaunchKernelGGL(k_ilu_fwd, dim3(grid_for(b1 - b0)), dim3(kBlock), 0, s, b1 - b0, S.fwd_cells.p + b0,
                             S.fwd_ptr.p, S.fwd_lin.p, S.fwd_src.p, anb, m->wk.dc, rcv, flags);
      }
      for (size_t l = 0; l + 1 < S.bwd_level_ptr.size(); ++l) {
        const int b0 = S.bwd_level_ptr[l], b1 = S.bwd_level_ptr[l + 1];
        if (b1 > b0)
          hipLaunchKernelGGL(k_ilu_bwd, dim3(grid_for(b1 - b0)), dim3(kBlock), 0, s, b1 - b0, S.bwd_cells.p + b0,
                             S.bwd_ptr.p, S.bwd_lin.p, S.bwd_src.p, anb, m->wk.dc, rcv, dphi, flags);
      }
      hipLaunchKernelGGL(k_resid_l1, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, part, flags);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_CONV, gr, n, nf, part, scal, flags);
    });
  } else {
    set_error("unsupported solver code " + std::to_string(solver));
    return -3;
  }
  (void)V5;
  (void)V6;
  (void)V7;

  // true final residual, fit += dphi (:180)
  int fl[F_NFLAGS];
  host_flags(m, fl);
  if (fl[F_BAD]) return refuse(fl[F_BAD]);
  hipLaunchKernelGGL(k_resid_l1, dim3(gr), dim3(kBlock), 0, s, n, w, nb, ac, anb, bc, dphi, part, nullptr);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, s, (int)FIN_TRUE_RES, gr, n, nf, part, scal, flags);
  add_inplace(m, fit, dphi);
  double h_out[S_NSLOTS];
  FS_HIP(hipMemcpyAsync(h_out, scal, sizeof(h_out), hipMemcpyDeviceToHost, s));
  FS_HIP(hipStreamSynchronize(s));
  if (res4) {
    res4[0] = h_out[S_RES_ABS];
    res4[1] = h_out[S_RES_MAX];
    res4[2] = h_out[S_RES_RMS];
    res4[3] = h_out[S_RES_SCALED];
  }
  if (lmg != FS_MG_SOLVER) iters = fl[F_ITER];
  if (st) {
    st->iterations = iters;
    st->levels = (lmg != FS_MG_OFF) ? amg_levels(m) : 1;
    st->initial_l1 = h_out[S_IRES];
    st->final_l1 = h_out[S_FRES];
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  const bool converged = !(h_out[S_FRES] > rrf * h_out[S_IRES]);
  if (std::isnan(h_out[S_FRES])) {
    set_error("residual is NaN");
    return -4;
  }
  return converged ? 0 : 1;
}

__global__ void k_add(int n, double *__restrict__ a, const double *__restrict__ b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = a[i] + b[i];
}

void launch_fin_conv(fs_mesh *m, int np) {
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, m->stream, (int)FIN_CONV, np, m->n, 4,
                     m->wk.partial, m->wk.scal, m->wk.flags);
}

void add_inplace(fs_mesh *m, double *a, const double *b) {
  hipLaunchKernelGGL(k_add, dim3(grid_for(m->n)), dim3(kBlock), 0, m->stream, m->n, a, b);
}

} // namespace fs

// ---- kernel timing for the roofline figure: the ELL product y = A*x of the
// Krylov loop (k_spmv_dot), launched reps times between two events on the
// mesh's stream.
namespace fs {
int time_spmv(fs_mesh *m, const double *ac, const double *anb, const double *x, const double *dot_with,
              double *y, int reps, double *avg_ms) {
  const int n = m->n, w = m->w, gr = grid_red(n);
  hipEvent_t e0, e1;
  FS_HIP(hipEventCreate(&e0));
  FS_HIP(hipEventCreate(&e1));
  int h_fl[F_NFLAGS] = {0, 0, 1, 0};
  FS_HIP(hipMemcpyAsync(m->wk.flags, h_fl, sizeof(h_fl), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_spmv_dot, dim3(gr), dim3(kBlock), 0, m->stream, n, w, EllNbr(m->d_nbr.p, m->d_nbr16.p), ac, anb, x,
                     y, dot_with, m->wk.partial, m->wk.flags);
  FS_HIP(hipEventRecord(e0, m->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_spmv_dot, dim3(gr), dim3(kBlock), 0, m->stream, n, w, EllNbr(m->d_nbr.p, m->d_nbr16.p), ac, anb,
                       x, y, dot_with, m->wk.partial, m->wk.flags);
  FS_HIP(hipEventRecord(e1, m->stream));
  FS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FS_HIP(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}
} // namespace fs

// ---- the solve without allocation: its scratch in one static device array.
// Layout (doubles, every slice 256-byte aligned): 8 work vectors and the D-ILU
// diagonal of kNoallocMaxCells each, the reduction partials, the scalar slots,
// the flags.
namespace fs {
constexpr int kNoallocMaxCells = FS_NOALLOC_MAX_CELLS;
constexpr size_t kNoallocVec = (size_t)kNoallocMaxCells;
constexpr size_t kNoallocPartial = 5 * (size_t)kPartials;
constexpr size_t kNoallocScal = 64, kNoallocFlags = 64;
static_assert(S_NSLOTS <= (int)kNoallocScal && F_NFLAGS * sizeof(int) <= kNoallocFlags * sizeof(double),
              "scalar and flag slices");
static_assert(kNoallocVec % 32 == 0 && kNoallocPartial % 32 == 0, "256-byte slices");
constexpr size_t kNoallocDoubles = 9 * kNoallocVec + kNoallocPartial + kNoallocScal + kNoallocFlags;
static_assert(kNoallocDoubles * sizeof(double) == FS_NOALLOC_SCRATCH_BYTES, "size stated in fsamd.h");
} // namespace fs

static __device__ __align__(256) double scratch[fs::kNoallocDoubles];

namespace fs {
// The device address of `scratch` on each device, taken once per device with
// that device current (each GPU holds its own copy of the array).
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64] = {nullptr};

static double *scratch_on(int dev) {
  if (dev < 0 || dev >= 64) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    double *ptr = nullptr;
    FS_HIP_THROW(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

int solve_device_noalloc(fs_mesh *m, int nf, double *fit, double rrf, int itmax, int solver, int lmg,
                         const double *ac, const double *anb, const double *bc, double *dphi, double *res4,
                         fs_solve_stats *st) {
  if (m->n > kNoallocMaxCells) {
    set_error("fs_solve_equation_device_noalloc: " + std::to_string(m->n) + " cells exceed the scratch array's " +
              std::to_string(kNoallocMaxCells));
    return -1;
  }
  if (lmg != FS_MG_OFF) {
    set_error("fs_solve_equation_device_noalloc: LMultigrid must be 0 (the AMG hierarchy is allocated)");
    return -1;
  }
  if (solver != FS_SOR && solver != FS_PBCG && solver != FS_BICGS && solver != FS_CG) {
    set_error("fs_solve_equation_device_noalloc: solver must be SOR, PBCG, BICGS or CG (ILU builds its "
              "substitution schedule and GMRES its basis on first use)");
    return -3;
  }
  int cur = 0;
  FS_HIP(hipGetDevice(&cur));
  if (cur != m->device) {
    set_error("fs_solve_equation_device_noalloc: the mesh's device " + std::to_string(m->device) +
              " is not the current device " + std::to_string(cur));
    return -1;
  }
  double *base = scratch_on(cur);
  fs_mesh::Scratch saved = m->wk, sc{};
  for (int k = 0; k < 8; ++k) sc.v[k] = base + (size_t)k * kNoallocVec;
  for (int k = 8; k < 12; ++k) sc.v[k] = nullptr; // the multigrid and GMRES vectors are not used
  sc.dc = base + 8 * kNoallocVec;
  sc.partial = base + 9 * kNoallocVec;
  sc.scal = sc.partial + kNoallocPartial;
  sc.flags = reinterpret_cast<int *>(sc.scal + kNoallocScal);
  m->wk = sc;
  int rc;
  try {
    rc = solve_device(m, nf, fit, rrf, itmax, solver, lmg, ac, anb, bc, dphi, res4, st, false);
  } catch (...) {
    m->wk = saved;
    throw;
  }
  m->wk = saved;
  return rc;
}
} // namespace fs

extern "C" int fs_solve_equation_device_noalloc(fs_mesh *m, int nf, double *d_fit, double rrf, int itmax, int solver,
                                                int lmultigrid, const double *d_ac, const double *d_anb,
                                                const double *d_bc, double *d_dphi, double *res4,
                                                fs_solve_stats *stats) {
  if (!m || !d_fit || !d_ac || !d_anb || !d_bc || !d_dphi) {
    fs::set_error("fs_solve_equation_device_noalloc: NULL argument");
    return -1;
  }
  try {
    return fs::solve_device_noalloc(m, nf, d_fit, rrf, itmax, solver, lmultigrid, d_ac, d_anb, d_bc, d_dphi, res4,
                                    stats);
  } catch (const std::exception &e) {
    fs::set_error(e.what());
    return -2;
  }
}
