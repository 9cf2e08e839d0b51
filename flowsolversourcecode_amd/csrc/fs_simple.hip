// One steady SIMPLE outer iteration on the device (MainProgram.for:257-348):
// SolveMomentumX/Y/Z (Momentum/SolveXMomentumEquation.for:175-299 and twins)
// and SolveContinuity (ContinuityEquation/SolveContinuityEquation.for:37-114)
// for laminar incompressible constant-property flow bounded by no-slip walls
// and symmetry planes.
//
// Face loops run face-ordered over structure-of-arrays geometry (coalesced);
// every face quantity the reference accumulates (FluxCf/FluxFf/FluxTf, the
// Green-Gauss face values) is formed by one thread in the reference's order
// and written once.  Cells then gather their faces through the per-slot face
// code (face*4 + role) in slot order, which is the reference's face-loop
// order, so the scatter AssembleGlobalMatrixFaceFluxes becomes an atomic-free
// gather that is bitwise the reference's per-cell sum.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "fs_common.h"
#include "fs_geom.h"
#include "fs_hr.h"
#include "fs_kernels.h"
#include "fs_solve.h"
#include "fs_mpnn.h"
#include "fs_simple.h"

namespace fs {

namespace {

enum { ROLE_NONE = 0, ROLE_OWNER = 1, ROLE_NEIGH = 2, ROLE_BOUND = 3 };

inline int blocks(int n) { return std::max(1, (n + kBlock - 1) / kBlock); }

struct GradArgs { // up to 4 fields at once
  const double *phi[4];
  const double *bval[4]; // boundary values per boundary face, or nullptr =
                         // the owner cell's own value
  double *grad[4];       // [gx | gy | gz], 3N each
  const double *fv[4];   // face values (interior)
};

// Face kernels stream the per-face geometry once per pass (non-temporal loads)
// and gather per-cell values, which the hint leaves in the caches.
constexpr bool kStreamFaces = true;

// Green-Gauss face value phi_f = phi_C g + phi_F (1-g)
// (StandardGreenGaussCorrection1Gradient.for:51-53)
// Face values live over all F faces (field k at fp[k F + f]): boundary faces
// carry the boundary value (or the owner's own value when the field has
// none), so the cell sums read one array whatever the face's role.
template <int NF>
__global__ void k_gg_face0(int Fi, int F, const int *__restrict__ own, const int *__restrict__ nei,
                           const double *__restrict__ g, GradArgs a, double *__restrict__ fp) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= F) return;
  const int i = own[f];
  if (f >= Fi) {
#pragma unroll
    for (int k = 0; k < NF; ++k) fp[(size_t)k * F + f] = a.bval[k] ? a.bval[k][f - Fi] : a.phi[k][i];
    return;
  }
  const int j = nei[f];
  const double gf = g[f];
#pragma unroll
  for (int k = 0; k < NF; ++k) fp[(size_t)k * F + f] = a.phi[k][i] * gf + a.phi[k][j] * (1. - gf);
}

// corrected face value at the face centroid (:108-147):
// phi_f = phi_f' + sum_d (grad_d C g + grad_d F (1-g)) (r_f - r_f')_d
template <int NF>
__global__ void k_gg_facecorr(int Fi, int F, int N, const int *__restrict__ own,
                              const int *__restrict__ nei, const double *__restrict__ g,
                              const double *__restrict__ kx, const double *__restrict__ ky,
                              const double *__restrict__ kz, const double *__restrict__ fp,
                              GradArgs a, double *__restrict__ fv) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= F) return;
  if (f >= Fi) { // boundary values are not corrected
#pragma unroll
    for (int k = 0; k < NF; ++k) fv[(size_t)k * F + f] = fp[(size_t)k * F + f];
    return;
  }
  const int i = ld_mat<kStreamFaces>(&own[f]), j = ld_mat<kStreamFaces>(&nei[f]);
  const double gf = ld_mat<kStreamFaces>(&g[f]), og = 1. - gf;
  const double rx = ld_mat<kStreamFaces>(&kx[f]), ry = ld_mat<kStreamFaces>(&ky[f]),
               rz = ld_mat<kStreamFaces>(&kz[f]);
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const double *gr = a.grad[k];
    double cx = gr[i] * gf + gr[j] * og;
    cx = cx * rx;
    double cy = gr[N + i] * gf + gr[N + j] * og;
    cy = cy * ry;
    double cz = gr[2 * N + i] * gf + gr[2 * N + j] * og;
    cz = cz * rz;
    fv[(size_t)k * F + f] = ld_mat<kStreamFaces>(&fp[(size_t)k * F + f]) + cx + cy + cz;
  }
}

// Cell sums of phi_f S_f over the cell's faces in slot (= face) order, then
// division by the volume (:35-64, :148-172).  Interior faces add +phi_f S_f
// at the owner and -phi_f S_f at the neighbour; boundary faces add the
// boundary value times S_b.
// One face's term of the cell sums: +phi_f S_f (owner, boundary) or
// -phi_f S_f (neighbour).
template <int NF>
__device__ __forceinline__ void gg_term(int role, double sx, double sy, double sz, const double *v, double *gx,
                                        double *gy, double *gz) {
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const double tx = v[k] * sx, ty = v[k] * sy, tz = v[k] * sz;
    if (role == ROLE_NEIGH) {
      gx[k] = gx[k] - tx;
      gy[k] = gy[k] - ty;
      gz[k] = gz[k] - tz;
    } else {
      gx[k] = gx[k] + tx;
      gy[k] = gy[k] + ty;
      gz[k] = gz[k] + tz;
    }
  }
}

// WC > 0: a mesh of WC slots per cell, all slots' loads issued before the
// sums (codes first, then the faces' S and values); WC = 0: any width.
template <int NF, int WC>
__global__ void k_gg_cell(int N, int W, int F, const int *__restrict__ code,
                          const double *__restrict__ Sx, const double *__restrict__ Sy,
                          const double *__restrict__ Sz, const double *__restrict__ vol,
                          const double *__restrict__ fv, GradArgs a) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c >= N) return;
  double gx[NF], gy[NF], gz[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) gx[k] = gy[k] = gz[k] = 0.0;
  if constexpr (WC > 0) {
    int cd[WC];
    double sx[WC], sy[WC], sz[WC], v[WC][NF];
#pragma unroll
    for (int s = 0; s < WC; ++s) cd[s] = code[(size_t)s * N + c];
#pragma unroll
    for (int s = 0; s < WC; ++s) {
      const int f = cd[s] >> 2;
      sx[s] = Sx[f];
      sy[s] = Sy[f];
      sz[s] = Sz[f];
#pragma unroll
      for (int k = 0; k < NF; ++k) v[s][k] = fv[(size_t)k * F + f];
    }
#pragma unroll
    for (int s = 0; s < WC; ++s)
      if ((cd[s] & 3) != ROLE_NONE) gg_term<NF>(cd[s] & 3, sx[s], sy[s], sz[s], v[s], gx, gy, gz);
  } else {
    for (int s = 0; s < W; ++s) {
      const int cd = code[(size_t)s * N + c];
      const int role = cd & 3, f = cd >> 2;
      if (role == ROLE_NONE) continue;
      double v[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) v[k] = fv[(size_t)k * F + f];
      gg_term<NF>(role, Sx[f], Sy[f], Sz[f], v, gx, gy, gz);
    }
  }
  const double v = vol[c];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    a.grad[k][c] = gx[k] / v;
    a.grad[k][N + c] = gy[k] / v;
    a.grad[k][2 * N + c] = gz[k] / v;
  }
}

// Tangential projection of the boundary pressure gradient: at no-slip walls
// CorrectPresureGradientAtWalls (CorrectGradients.for:3-60), at symmetry
// faces CorrectPressureGradientAtSymmetryBoundary (CorrectGradients.for
// :244-289) with its own operand order and the z-normal it reads from face
// (i,i) of the patch (bzq).  btype == nullptr: every face is a wall.
__global__ void k_wall_pgrad(int Fb, int Fi, int N, const int *__restrict__ own,
                             const double *__restrict__ nx, const double *__restrict__ ny,
                             const double *__restrict__ nz, const int *__restrict__ btype,
                             const int *__restrict__ bzq, const double *__restrict__ gp,
                             double *__restrict__ bgp) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= Fb) return;
  if (btype && btype[b] == FS_BC_PERIODIC) return; // the periodic correction's faces
  const int f = Fi + b, i1 = own[f];
  const double x = nx[f], y = ny[f], z = nz[f];
  const double gx = gp[i1], gy = gp[N + i1], gz = gp[2 * N + i1];
  if (btype && btype[b] == FS_BC_SYMMETRY) {
    const double zq = nz[Fi + bzq[b]];
    bgp[b] = gx * (1. - x * x) - gy * x * y - gz * x * zq;
    bgp[Fb + b] = gy * (1. - y * y) - gx * y * x - gz * y * zq;
    bgp[2 * Fb + b] = gz * (1. - zq * zq) - gx * zq * x - gy * zq * y;
    return;
  }
  bgp[b] = (1 - x * x) * gx - x * y * gy - x * z * gz;
  bgp[Fb + b] = (1 - y * y) * gy - y * x * gx - y * z * gz;
  bgp[2 * Fb + b] = (1 - z * z) * gz - z * x * gx - z * y * gy;
}

// Gradient limiting (LimitGradients.for:3): fields k of the preceding
// gradient launch whose method is 1 or 2; their boundary values are that
// launch's boundary face values fp[k F + f].
struct LimArgs {
  const double *phi[4];
  double *grad[4];
  int method[4];      // 0: leave the field, 1 or 2
  const double *eps2; // per field, (1e-3 (max - min))^2 of the method-2 fields
};

// Range of each method-2 field over its cells and boundary values
// (PhiFieldMinMax, LimitGradients.for:440-476): per-block min and max into
// part[(2k) nb + b] and part[(2k+1) nb + b].
template <int NF>
__global__ void k_lim_range(int N, int Fi, int F, const double *__restrict__ fp, LimArgs a,
                            double *__restrict__ part) {
  double lo[NF], hi[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    lo[k] = 1.e30;
    hi[k] = -1.e30;
  }
  const int Fb = F - Fi;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < N + Fb; t += gridDim.x * kBlock)
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      if (a.method[k] != 2) continue;
      const double v = t < N ? a.phi[k][t] : fp[(size_t)k * F + Fi + (t - N)];
      lo[k] = min2(lo[k], v);
      hi[k] = max2(hi[k], v);
    }
  __shared__ double s_lo[NF][kBlock], s_hi[NF][kBlock];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    s_lo[k][threadIdx.x] = lo[k];
    s_hi[k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        s_lo[k][threadIdx.x] = min2(s_lo[k][threadIdx.x], s_lo[k][threadIdx.x + h]);
        s_hi[k][threadIdx.x] = max2(s_hi[k][threadIdx.x], s_hi[k][threadIdx.x + h]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      part[(size_t)(2 * k) * gridDim.x + blockIdx.x] = s_lo[k][0];
      part[(size_t)(2 * k + 1) * gridDim.x + blockIdx.x] = s_hi[k][0];
    }
}

// eps2 = (epsilon (phiMax - phiMin))^2, epsilon = 1e-3 (Constants1.F90,
// LimitGradients.for:307), from nb per-block ranges; one block.
template <int NF>
__global__ void k_lim_eps(int nb, const double *__restrict__ part, double *__restrict__ eps2) {
  __shared__ double s_lo[NF][kBlock], s_hi[NF][kBlock];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    double lo = 1.e30, hi = -1.e30;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
      lo = min2(lo, part[(size_t)(2 * k) * nb + b]);
      hi = max2(hi, part[(size_t)(2 * k + 1) * nb + b]);
    }
    s_lo[k][threadIdx.x] = lo;
    s_hi[k][threadIdx.x] = hi;
  }
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        s_lo[k][threadIdx.x] = min2(s_lo[k][threadIdx.x], s_lo[k][threadIdx.x + h]);
        s_hi[k][threadIdx.x] = max2(s_hi[k][threadIdx.x], s_hi[k][threadIdx.x + h]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const double e = 1.e-3 * (s_hi[k][0] - s_lo[k][0]);
      eps2[k] = e * e;
    }
}

// One factor per cell scales its gradient (LimitGradientMethod1,
// LimitGradients.for:68-262 with rk = 1, and LimitGradientMethod2, :263-437):
// the min/max over the cell, its neighbours and its boundary values, then
// each face's extrapolation d = grad . (r_f - r_C), interior faces in slot
// order before boundary faces as the reference's two face loops visit them.
// Every cell reads only its own gradient, so the update is in place; the
// boundary gradients are the owners' (read from the cell arrays).
template <int NF>
__global__ void k_limit(int N, int W, int F, const int *__restrict__ code, const int *__restrict__ nbr,
                        const double *__restrict__ fcx, const double *__restrict__ fcy,
                        const double *__restrict__ fcz, const double *__restrict__ xc,
                        const double *__restrict__ yc, const double *__restrict__ zc,
                        const double *__restrict__ fp, LimArgs a) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c >= N) return;
  const double tiny = 1.e-15;
  const double cx = xc[c], cy = yc[c], cz = zc[c];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const int meth = a.method[k];
    if (meth == 0) continue;
    const double *phi = a.phi[k];
    const double pc = phi[c];
    double lo = min2(1.e30, pc), hi = max2(-1.e30, pc);
    for (int pass = 0; pass < 2; ++pass)
      for (int s = 0; s < W; ++s) {
        const int cd = code[(size_t)s * N + c], role = cd & 3;
        if (pass == 0 ? (role != ROLE_OWNER && role != ROLE_NEIGH) : role != ROLE_BOUND) continue;
        const double v = pass == 0 ? phi[nbr[(size_t)s * N + c]] : fp[(size_t)k * F + (cd >> 2)];
        lo = min2(lo, v);
        hi = max2(hi, v);
      }
    double *gr = a.grad[k];
    const double gx = gr[c], gy = gr[N + c], gz = gr[2 * N + c];
    const double eps2 = meth == 2 ? a.eps2[k] : 0.;
    const double dmax = hi - pc, dmin = lo - pc;
    double lim = meth == 2 ? (double)1.e30f : 1.;
    for (int pass = 0; pass < 2; ++pass)
      for (int s = 0; s < W; ++s) {
        const int cd = code[(size_t)s * N + c], role = cd & 3, f = cd >> 2;
        if (pass == 0 ? (role != ROLE_OWNER && role != ROLE_NEIGH) : role != ROLE_BOUND) continue;
        const double d = gx * (fcx[f] - cx) + gy * (fcy[f] - cy) + gz * (fcz[f] - cz);
        if (meth == 1) {
          if (d > dmax + tiny) lim = min2(lim, dmax / (d + tiny));
          else if (d < dmin - tiny) lim = min2(lim, dmin / (d + tiny));
        } else {
          const double base = pass == 0 ? pc : fp[(size_t)k * F + f];
          const double dp = d >= 0. ? hi - base : lo - base;
          const double dp2 = dp * dp, dm2 = d * d;
          const double part1 = (dp2 + eps2) * d + 2. * dm2 * dp;
          const double part2 = dp2 + 2. * dm2 + dp * d + eps2;
          const double r = part1 / (copysign(1., d) * (fabs(d) + tiny) * part2);
          lim = min2(min2(1., r), lim);
        }
      }
    gr[c] = lim * gx;
    gr[N + c] = lim * gy;
    gr[2 * N + c] = lim * gz;
  }
}

struct MomArgs {
  int comp;
  double mu, c23;
  const double *bgam;                      // BeDiffCoefficient per boundary face, nullptr = mu
  const double *phi, *u, *v, *w;          // current velocity components
  const double *bu, *bv, *bw;              // boundary velocities
  const int *btype;                        // per boundary face, nullptr = all walls
  const double *gu, *gv, *gw;              // cell gradients (boundary copies)
  const double *mdot;
};

// One field's face gradient by GradientInterpolationScheme
// (InterpolateGradientFromElementToFace.for:16-108): the owner/neighbour
// weight is g ('average', 'averagecorrected', 'harmonic') or follows the
// mass flux ('upwind': 1 for mdot > 0, 'downwind' the opposite);
// 'averagecorrected' adds the centroid-line correction, 'harmonic' weighs by
// the (constant) density.
__device__ __forceinline__ void face_grad(int scheme, double gf, double md, double rho, double gix, double giy,
                                          double giz, double gjx, double gjy, double gjz, double pi, double pj,
                                          double dcf, double ex, double ey, double ez, double &fx, double &fy,
                                          double &fz) {
  if (scheme == FS_GRADINT_UPWIND) gf = md > 0. ? 1. : 0.;
  else if (scheme == FS_GRADINT_DOWNWIND) gf = md > 0. ? 0. : 1.;
  if (scheme == FS_GRADINT_HARMONIC) {
    fx = rho * (gf * gix / rho + (1. - gf) * gjx / rho);
    fy = rho * (gf * giy / rho + (1. - gf) * gjy / rho);
    fz = rho * (gf * giz / rho + (1. - gf) * gjz / rho);
  } else {
    fx = gf * gix + (1. - gf) * gjx;
    fy = gf * giy + (1. - gf) * gjy;
    fz = gf * giz + (1. - gf) * gjz;
  }
  if (scheme == FS_GRADINT_AVERAGECORRECTED) {
    double corr = (pj - pi) / dcf;
    corr = corr - (fx * ex + fy * ey + fz * ez);
    fx = fx + corr * ex;
    fy = fy + corr * ey;
    fz = fz + corr * ez;
  }
}

// Momentum face fluxes of the interior faces for u, v and w in one pass, each
// component's in the reference's order: diffusion
// (AssembleIsotropicDiffusion.for:99-117), explicit diffusion
// (AssembleExplicitDiffusion.for), bulk viscosity (AssembleBulkViscosity.for),
// upwind convection (AssembleConvection.for:72-88) and, for HRF = FS_HR_TVD
// or FS_HR_NVF, the deferred high-resolution correction
// (AssembleHRSchemeCorrection.for:39-133) from the upwind cell's gradient.
// FluxCf and FluxFf are the same for the three components (cf, ff); FluxTf
// of component c goes to tf3[c F + f].  The velocity face gradients,
// 'average' interpolation (InterpolateGradientFromElementToFace.for:16-93),
// are formed here from the cell gradients.  Interior fluxes depend on u, v,
// w only through the component's own field and gradient, so those of v and w
// are the same whether formed before or after the u solve.
template <int HRF>
__global__ void k_mom_face3(int Fi, int F, int N, const int *__restrict__ own, const int *__restrict__ nei,
                            const double *__restrict__ g, const double *__restrict__ gdiff,
                            const double *__restrict__ Tx, const double *__restrict__ Ty,
                            const double *__restrict__ Tz, const double *__restrict__ Sx,
                            const double *__restrict__ Sy, const double *__restrict__ Sz, double mu,
                            const double *__restrict__ gam, double c23,
                            const double *__restrict__ u, const double *__restrict__ v,
                            const double *__restrict__ w, const double *__restrict__ gu,
                            const double *__restrict__ gv, const double *__restrict__ gw,
                            const double *__restrict__ mdot, const double *__restrict__ dcx,
                            const double *__restrict__ dcy, const double *__restrict__ dcz, int scheme,
                            double bleed, int gint, double rho, const double *__restrict__ ex,
                            const double *__restrict__ ey, const double *__restrict__ ez,
                            const double *__restrict__ dcf, double *__restrict__ cf, double *__restrict__ ff,
                            double *__restrict__ tf3, double *__restrict__ fg_out) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= Fi) return;
  const int i = ld_mat<kStreamFaces>(&own[f]), j = ld_mat<kStreamFaces>(&nei[f]);
  const double gf = ld_mat<kStreamFaces>(&g[f]), og = 1. - gf;
  // cell gradients of component k in direction d at the owner and neighbour
  const double *G[3] = {gu, gv, gw};
  double gi[3][3], gj[3][3], fgv[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      gi[k][d] = G[k][(size_t)d * N + i];
      gj[k][d] = G[k][(size_t)d * N + j];
    }
  const double *P[3] = {u, v, w};
  if (gint == FS_GRADINT_AVERAGE) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) fgv[k][d] = gf * gi[k][d] + og * gj[k][d];
  } else {
    const double mdf = mdot[f];
    const bool corr = gint == FS_GRADINT_AVERAGECORRECTED;
    const double e0 = corr ? ex[f] : 0., e1 = corr ? ey[f] : 0., e2 = corr ? ez[f] : 0., dd = corr ? dcf[f] : 1.;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      face_grad(gint, gf, mdf, rho, gi[k][0], gi[k][1], gi[k][2], gj[k][0], gj[k][1], gj[k][2], P[k][i], P[k][j],
                dd, e0, e1, e2, fgv[k][0], fgv[k][1], fgv[k][2]);
  }
  if (fg_out) // the face gradients of the restart file, component-major
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) fg_out[(size_t)(3 * k + d) * Fi + f] = fgv[k][d];
  // GamaFace: the viscosity, or the eddy-viscosity model's eDiffCoefficient
  const double gamf = gam ? gf * gam[i] + (1. - gf) * gam[j] : gf * mu + (1. - gf) * mu;
  const double gd = ld_mat<kStreamFaces>(&gdiff[f]);
  const double cfl = gamf * gd, ffl = -(gamf * gd);
  const double md = ld_mat<kStreamFaces>(&mdot[f]);
  const double ccl = md > 0. ? md : 0.;
  const double fcl = -(-md > 0. ? -md : 0.);
  // cf, ff: diffusion, explicit (+0), bulk (+0), convection
  double C = 0.0 + cfl, Fv = 0.0 + ffl;
  C = C + 0.;
  Fv = Fv + 0.;
  C = C + 0.;
  Fv = Fv + 0.;
  C = C + ccl;
  Fv = Fv + fcl;
  cf[f] = C;
  ff[f] = Fv;
  const double tx = ld_mat<kStreamFaces>(&Tx[f]), ty = ld_mat<kStreamFaces>(&Ty[f]),
               tz = ld_mat<kStreamFaces>(&Tz[f]), sx = ld_mat<kStreamFaces>(&Sx[f]),
               sy = ld_mat<kStreamFaces>(&Sy[f]), sz = ld_mat<kStreamFaces>(&Sz[f]);
  const double dv = fgv[0][0] + fgv[1][1] + fgv[2][2];
  // the upwind side: the owner for mdot > 0, else the neighbour
  const bool up = md > 0.;
  double ddx = 0., ddy = 0., ddz = 0.;
  if constexpr (HRF != FS_HR_NONE) {
    ddx = up ? dcx[f] : -dcx[f];
    ddy = up ? dcy[f] : -dcy[f];
    ddz = up ? dcz[f] : -dcz[f];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double pi = P[c][i], pj = P[c][j];
    const double vfl = -(gamf * (fgv[c][0] * tx + fgv[c][1] * ty + fgv[c][2] * tz));
    double T = 0.0 + cfl * pi + ffl * pj + vfl;
    // explicit diffusion: -Gamma_f (grad u_c)^T . S
    const double v2 = -(gamf * (fgv[0][c] * sx + fgv[1][c] * sy + fgv[2][c] * sz));
    T = T + v2;
    // bulk viscosity
    const double S = c == 0 ? sx : c == 1 ? sy : sz;
    const double v3 = -(c23 * gamf * dv * S);
    T = T + v3;
    // upwind convection
    T = T + ccl * pi + fcl * pj;
    if constexpr (HRF != FS_HR_NONE) {
      const double dphi = hr_increment<HRF>(scheme, bleed, up ? pi : pj, up ? pj : pi, up ? gi[c][0] : gj[c][0],
                                            up ? gi[c][1] : gj[c][1], up ? gi[c][2] : gj[c][2], ddx, ddy, ddz);
      T = T + md * dphi * 1.;
    }
    tf3[(size_t)c * F + f] = T;
  }
}

// No-slip wall faces (AssembleIsotropicDiffusion.for:148-219) or symmetry
// faces (AssembleIsotropicDiffusion.for:562-620), then the explicit and bulk
// terms with the owner's gradient as boundary gradient.
__global__ void k_mom_wall(int Fb, int Fi, int N, const int *__restrict__ own,
                           const double *__restrict__ A, const double *__restrict__ nx,
                           const double *__restrict__ ny, const double *__restrict__ nz,
                           const double *__restrict__ dx, const double *__restrict__ dy,
                           const double *__restrict__ dz, const double *__restrict__ Sx,
                           const double *__restrict__ Sy, const double *__restrict__ Sz, MomArgs a,
                           double *__restrict__ cf, double *__restrict__ ff, double *__restrict__ tf) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= Fb) return;
  if (a.btype && a.btype[b] == FS_BC_PERIODIC) return; // k_per_momentum's faces
  const int f = Fi + b, i1 = own[f], c = a.comp;
  const double gamf = a.bgam ? a.bgam[b] : a.mu;
  const double x = nx[f], y = ny[f], z = nz[f], area = A[f];
  const double dn = dx[f] * x + dy[f] * y + dz[f] * z;
  const double bphi = c == 0 ? a.bu[b] : c == 1 ? a.bv[b] : a.bw[b];
  double cfl, vfl;
  if (a.btype && a.btype[b] == FS_BC_SYMMETRY) {
    if (c == 0) {
      cfl = 2. * gamf * area * (x * x) / dn;
      vfl = -((-(2. * gamf * area / dn)) * (a.v[i1] * y + a.w[i1] * z) * x);
    } else if (c == 1) {
      cfl = 2. * gamf * area * (y * y) / dn;
      vfl = -((-(2. * gamf * area / dn)) * (a.u[i1] * x + a.w[i1] * z) * y);
    } else {
      cfl = 2. * gamf * area * (z * z) / dn;
      vfl = -((-(2. * gamf * area / dn)) * (a.u[i1] * x + a.v[i1] * y) * z);
    }
  } else if (c == 0) {
    cfl = gamf * area * (1. - x * x) / dn;
    vfl = -((gamf * area / dn) * (bphi * (1. - x * x) + (a.v[i1] - a.bv[b]) * x * y +
                                   (a.w[i1] - a.bw[b]) * x * z));
  } else if (c == 1) {
    cfl = gamf * area * (1. - y * y) / dn;
    vfl = -((gamf * area / dn) * ((a.u[i1] - a.bu[b]) * x * y + bphi * (1. - y * y) +
                                   (a.w[i1] - a.bw[b]) * z * y));
  } else {
    cfl = gamf * area * (1. - z * z) / dn;
    vfl = -((gamf * area / dn) * ((a.u[i1] - a.bu[b]) * x * z + (a.v[i1] - a.bv[b]) * y * z +
                                   bphi * (1. - z * z)));
  }
  const double ffl = 0.;
  double C = 0.0 + cfl, T = 0.0 + cfl * a.phi[i1] + ffl * bphi + vfl;
  const double bga = a.gu[(size_t)c * N + i1], bgb = a.gv[(size_t)c * N + i1],
               bgc = a.gw[(size_t)c * N + i1];
  const double v2 = -(gamf * (bga * Sx[f] + bgb * Sy[f] + bgc * Sz[f]));
  C = C + 0.;
  T = T + v2;
  const double dv = a.gu[i1] + a.gv[N + i1] + a.gw[2 * N + i1];
  const double S = c == 0 ? Sx[f] : c == 1 ? Sy[f] : Sz[f];
  const double v3 = -(a.c23 * gamf * dv * S);
  C = C + 0.;
  T = T + v3;
  cf[f] = C;
  ff[f] = 0.0;
  tf[f] = T;
}

// Per-cell gather of the momentum system: face fluxes in slot order
// (AssembleGlobalMatrixFaceFluxes), effective divergence and element fluxes
// (CalculateEffectiveDivergence.for, AssembleConvection.for:169-178), the
// pressure gradient (AssemblePressureGradient.for), under-relaxation
// (UnderRelaxation.for) and D = V/a_P, u* = u (StoreTheDValues.for).
// One slot of the momentum gather (face fluxes cf, ff, tf and mass flux md
// of the slot's face).
// PER: periodic faces are boundary faces with the partner cell in the slot
// (LPeriodicImplicit): the slot takes FluxFf (AssembleGolbalElementAndFace
// Fluxes.for:99-121) and the effective divergence the face's Bmdot; wall and
// symmetry faces carry FluxFf = 0 and Bmdot = 0.
template <bool PER>
__device__ __forceinline__ double mom_slot(int role, double vc, double vf, double vt, double md, double &a,
                                           double &b, double &e) {
  double an = 0.0;
  if (role == ROLE_OWNER) {
    a = a + vc;
    an = an + vf;
    b = b - vt;
    e = e + md;
  } else if (role == ROLE_NEIGH) {
    a = a - vf;
    an = an - vc;
    b = b + vt;
    e = e - md;
  } else if (role == ROLE_BOUND) {
    a = a + vc;
    if constexpr (PER) an = an + vf;
    b = b - vt;
    if constexpr (PER) e = e + md;
    else e = e + 0.;
  }
  return an;
}

template <int WC, bool PER>
__global__ void k_mom_cell(int N, int W, int Fi, const int *__restrict__ code,
                           const double *__restrict__ cf, const double *__restrict__ ff,
                           const double *__restrict__ tf, const double *__restrict__ mdot,
                           const double *__restrict__ bmdot, const double *__restrict__ lin,
                           const double *__restrict__ phi, const double *__restrict__ gpc,
                           const double *__restrict__ vol, double urf, const int *__restrict__ perm,
                           double *__restrict__ ac, double *__restrict__ anb, double *__restrict__ bc,
                           double *__restrict__ D, double *__restrict__ star, double *__restrict__ xq,
                           double *__restrict__ D2, double *__restrict__ eff_out, const double *__restrict__ grk) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c >= N) return;
  // the system row of cell c: its colour-major position when perm is given
  // (the unknown is then copied to that position as well, into xq)
  const int q = perm ? perm[c] : c;
  double a = 0.0, b = 0.0, e = 0.0;
  if constexpr (WC > 0) {
    // all slots' loads first (the mass flux exists for interior faces only)
    int cd[WC];
    double vc[WC], vf[WC], vt[WC], vm[WC];
#pragma unroll
    for (int s = 0; s < WC; ++s) cd[s] = code[(size_t)s * N + c];
#pragma unroll
    for (int s = 0; s < WC; ++s) {
      const int f = cd[s] >> 2, role = cd[s] & 3;
      vc[s] = cf[f];
      vf[s] = ff[f];
      vt[s] = tf[f];
      if (PER && role == ROLE_BOUND) vm[s] = bmdot[f - Fi];
      else vm[s] = mdot[(role == ROLE_OWNER || role == ROLE_NEIGH) ? f : 0];
    }
#pragma unroll
    for (int s = 0; s < WC; ++s)
      anb[(size_t)s * N + q] = mom_slot<PER>(cd[s] & 3, vc[s], vf[s], vt[s], vm[s], a, b, e);
  } else {
    for (int s = 0; s < W; ++s) {
      const int cd = code[(size_t)s * N + c];
      const int role = cd & 3, f = cd >> 2;
      const double md = (role == ROLE_OWNER || role == ROLE_NEIGH) ? mdot[f]
                        : (PER && role == ROLE_BOUND) ? bmdot[f - Fi] : 0.0;
      anb[(size_t)s * N + q] = mom_slot<PER>(role, cf[f], ff[f], tf[f], md, a, b, e);
    }
  }
  if (eff_out) eff_out[c] = e; // effdiv of the restart file
  const double pc = phi[c];
  const double ep = e > 0.0 ? e : 0.0;
  const double fce = 0.0 + (ep - e);
  double fte = 0.0 + ((ep - e) * pc + (-(ep * pc)));
  if constexpr (PER) fte = fte - lin[0] * vol[c]; // AssembleLinearPressureTerm
  fte = fte + gpc[c] * vol[c];
  if (grk) fte = fte + grk[c] * vol[c]; // MinusTwoThirdsRhoK
  a = a + fce;
  b = b - fte;
  a = a / urf;
  ac[q] = a;
  bc[q] = b;
  D[c] = vol[c] / a;
  star[c] = pc;
  if (perm) xq[q] = pc;
  if (D2) { // SIMPLEC: V / (a_P + the row's anb in slot order)
    double t = a;
    for (int s = 0; s < W; ++s) t = t + anb[(size_t)s * N + q];
    D2[c] = vol[c] / t;
  }
}

// y[c] = x[perm[c]]: a colour-major vector back in cell order
__global__ void k_unpermute(int N, const int *__restrict__ perm, const double *__restrict__ x,
                            double *__restrict__ y) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c < N) y[c] = x[perm[c]];
}

struct RcArgs {
  double rho, urf_u;
  const double *u, *v, *w, *us, *vs, *ws, *du, *dv, *dw, *gp, *p;
  const double *du2, *dv2, *dw2; // SIMPLEC's D2 (the p' coefficients)
  int gint;                      // GradientInterpolationSchemeContinuity
  double *fgp_out;               // the pressure face gradient of the restart file, or nullptr
  const double *rhok, *grk;      // 2/3 rho k and its cell gradient (the eddy-viscosity model), or nullptr
};

// geoDiff = |D_E| / d_CF of the face's D S split by MethodDecomposeS DEC
// (AssembleMdotUsingRhieChow.for:575-615, AssemblePressureCorrectionFluxes.for
// :99-141): 1 D_E = (e.DS) e, 2 D_E = |DS| e, 3 D_E = |DS|^2/(e.DS) e (the
// last in the reference's operand order); the E part of D S goes to due..dwe.
template <int DEC>
__device__ __forceinline__ double geo_split(double dusf, double dvsf, double dwsf, double ex, double ey,
                                            double ez, double dcf, double &due, double &dve, double &dwe) {
  if constexpr (DEC == 1) {
    const double dot = ex * dusf + ey * dvsf + ez * dwsf;
    due = dot * ex;
    dve = dot * ey;
    dwe = dot * ez;
  } else if constexpr (DEC == 2) {
    const double mag = sqrt(dusf * dusf + dvsf * dvsf + dwsf * dwsf);
    due = ex * mag;
    dve = ey * mag;
    dwe = ez * mag;
  } else {
    const double mag = sqrt(dusf * dusf + dvsf * dvsf + dwsf * dwsf);
    const double eu = dusf / mag, ev = dvsf / mag, ew = dwsf / mag;
    const double dot = ex * eu + ey * ev + ez * ew;
    due = mag * ex / dot;
    dve = mag * ey / dot;
    dwe = mag * ez / dot;
  }
  return sqrt(due * due + dve * dve + dwe * dwe) / dcf;
}

// Rhie-Chow face mass flux (AssembleMdotUsingRhieChow.for:554-759, steady,
// LRelaxMomentum) followed by the pressure-correction face coefficients
// (AssemblePressureCorrectionFluxes.for:85-153).  The pressure face gradient,
// 'averagecorrected' (InterpolateGradientFromElementToFace.for:16-93), is
// completed here from the averaged cell gradient the flux already forms.
// SIMPLEC: the p' coefficients take geoDiff of D2 instead of D1.
template <bool SIMPLEC, int DEC>
__global__ void k_rhie_chow(int Fi, int N, const int *__restrict__ own,
                            const int *__restrict__ nei, const double *__restrict__ g,
                            const double *__restrict__ Sx, const double *__restrict__ Sy,
                            const double *__restrict__ Sz, const double *__restrict__ ex,
                            const double *__restrict__ ey, const double *__restrict__ ez,
                            const double *__restrict__ dcf, RcArgs a, double *__restrict__ mdot,
                            double *__restrict__ cf, double *__restrict__ ff,
                            double *__restrict__ tf) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= Fi) return;
  const int i = own[f], j = nei[f];
  const double gf = g[f], og = 1. - gf;
  const double du = gf * a.du[i] + og * a.du[j], dv = gf * a.dv[i] + og * a.dv[j],
               dw = gf * a.dw[i] + og * a.dw[j];
  const double ub = gf * a.u[i] + og * a.u[j], vb = gf * a.v[i] + og * a.v[j],
               wb = gf * a.w[i] + og * a.w[j];
  const double usf = gf * a.us[i] + og * a.us[j], vsf = gf * a.vs[i] + og * a.vs[j],
               wsf = gf * a.ws[i] + og * a.ws[j];
  const double pbx = gf * a.gp[i] + og * a.gp[j], pby = gf * a.gp[N + i] + og * a.gp[N + j],
               pbz = gf * a.gp[2 * N + i] + og * a.gp[2 * N + j];
  const double sx = Sx[f], sy = Sy[f], sz = Sz[f];
  const double rhof = a.rho;
  double fv = rhof * (ub * sx + vb * sy + wb * sz);
  const double dusf = du * sx, dvsf = dv * sy, dwsf = dw * sz;
  double due, dve, dwe;
  const double geo = geo_split<DEC>(dusf, dvsf, dwsf, ex[f], ey[f], ez[f], dcf[f], due, dve, dwe);
  const double dut = dusf - due, dvt = dvsf - dve, dwt = dwsf - dwe;
  const double cfl = rhof * geo, ffl = -(rhof * geo);
  // the pressure face gradient by GradientInterpolationSchemeContinuity (the
  // averaged cell gradient above also stands in the D_S term)
  double fgx, fgy, fgz;
  if (a.gint == FS_GRADINT_AVERAGECORRECTED) {
    double corr = (a.p[j] - a.p[i]) / dcf[f];
    corr = corr - (pbx * ex[f] + pby * ey[f] + pbz * ez[f]);
    fgx = pbx + corr * ex[f];
    fgy = pby + corr * ey[f];
    fgz = pbz + corr * ez[f];
  } else {
    face_grad(a.gint, gf, mdot[f], a.rho, a.gp[i], a.gp[N + i], a.gp[2 * N + i], a.gp[j], a.gp[N + j],
              a.gp[2 * N + j], a.p[i], a.p[j], dcf[f], ex[f], ey[f], ez[f], fgx, fgy, fgz);
  }
  fv = fv - rhof * (fgx * dut + fgy * dvt + fgz * dwt);
  fv = fv + rhof * (pbx * dusf + pby * dvsf + pbz * dwsf);
  if (a.rhok) {
    // the -2/3 rho k gradient's difference (AssembleMdotUsingRhieChow.for
    // :627-635), its face gradient 'average' and 'averagecorrected'
    // (:294-306)
    const double kbx = gf * a.grk[i] + og * a.grk[j], kby = gf * a.grk[N + i] + og * a.grk[N + j],
                 kbz = gf * a.grk[2 * N + i] + og * a.grk[2 * N + j];
    double kc = (a.rhok[j] - a.rhok[i]) / dcf[f];
    kc = kc - (kbx * ex[f] + kby * ey[f] + kbz * ez[f]);
    const double kfx = kbx + kc * ex[f], kfy = kby + kc * ey[f], kfz = kbz + kc * ez[f];
    fv = fv - rhof * (kfx * dut + kfy * dvt + kfz * dwt);
    fv = fv + rhof * (kbx * dusf + kby * dvsf + kbz * dwsf);
    fv = fv + cfl * a.rhok[i] + ffl * a.rhok[j];
  }
  if (a.fgp_out) {
    a.fgp_out[f] = fgx;
    a.fgp_out[(size_t)Fi + f] = fgy;
    a.fgp_out[2 * (size_t)Fi + f] = fgz;
  }
  const double mstar = rhof * (usf * sx + vsf * sy + wsf * sz);
  fv = fv + (1. - a.urf_u) * (mdot[f] - mstar);
  const double m = cfl * a.p[i] + ffl * a.p[j] + fv;
  mdot[f] = m;
  if constexpr (SIMPLEC) {
    const double d2u = gf * a.du2[i] + og * a.du2[j], d2v = gf * a.dv2[i] + og * a.dv2[j],
                 d2w = gf * a.dw2[i] + og * a.dw2[j];
    double e2u, e2v, e2w;
    const double geo2 = geo_split<DEC>(d2u * sx, d2v * sy, d2w * sz, ex[f], ey[f], ez[f], dcf[f], e2u, e2v, e2w);
    cf[f] = rhof * geo2;
    ff[f] = -(rhof * geo2);
  } else {
    cf[f] = cfl;
    ff[f] = ffl;
  }
  tf[f] = m;
}

// p' system gather (walls carry no flux: FluxCf = 0, Bmdot = 0; PER: the
// boundary faces' fluxes of k_per_pc_boundary, FluxFf into the slot).
template <bool PER>
__device__ __forceinline__ double pc_slot(int role, double vc, double vf, double vt, double &a, double &b) {
  double an = 0.0;
  if (role == ROLE_OWNER) {
    a = a + vc;
    an = an + vf;
    b = b - vt;
  } else if (role == ROLE_NEIGH) {
    a = a - vf;
    an = an - vc;
    b = b + vt;
  } else if (PER && role == ROLE_BOUND) {
    a = a + vc;
    an = an + vf;
    b = b - vt;
  }
  return an;
}

template <int WC, bool PER>
__global__ void k_pc_cell(int N, int W, const int *__restrict__ code,
                          const double *__restrict__ cf, const double *__restrict__ ff,
                          const double *__restrict__ tf, double *__restrict__ ac,
                          double *__restrict__ anb, double *__restrict__ bc) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c >= N) return;
  double a = 0.0, b = 0.0;
  if constexpr (WC > 0) {
    int cd[WC];
    double vc[WC], vf[WC], vt[WC];
#pragma unroll
    for (int s = 0; s < WC; ++s) cd[s] = code[(size_t)s * N + c];
#pragma unroll
    for (int s = 0; s < WC; ++s) {
      const int f = cd[s] >> 2;
      vc[s] = cf[f];
      vf[s] = ff[f];
      vt[s] = tf[f];
    }
#pragma unroll
    for (int s = 0; s < WC; ++s) anb[(size_t)s * N + c] = pc_slot<PER>(cd[s] & 3, vc[s], vf[s], vt[s], a, b);
  } else {
    for (int s = 0; s < W; ++s) {
      const int cd = code[(size_t)s * N + c];
      const int f = cd >> 2;
      anb[(size_t)s * N + c] = pc_slot<PER>(cd & 3, cf[f], ff[f], tf[f], a, b);
    }
  }
  ac[c] = a;
  bc[c] = b;
}

// CorrectPressure (CorrectPressure.for:21-42)
__global__ void k_correct_p(int N, double urf_p, int ref, const double *__restrict__ pc,
                            double *__restrict__ p) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N) return;
  const double r = ref >= 0 ? pc[ref] : 0.0;
  p[c] = p[c] + urf_p * (pc[c] - r);
}

// CorrectVelocity (CorrectMdotandVelocity.for:774-778)
__global__ void k_correct_u(int N, const double *__restrict__ gpc, const double *__restrict__ du,
                            const double *__restrict__ dv, const double *__restrict__ dw,
                            double *__restrict__ u, double *__restrict__ v, double *__restrict__ w) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N) return;
  u[c] = u[c] - du[c] * gpc[c];
  v[c] = v[c] - dv[c] * gpc[N + c];
  w[c] = w[c] - dw[c] * gpc[2 * N + c];
}

// Correctmdot (CorrectMdotandVelocity.for:55-63)
__global__ void k_correct_mdot(int Fi, const int *__restrict__ own, const int *__restrict__ nei,
                               const double *__restrict__ cf, const double *__restrict__ ff,
                               const double *__restrict__ pc, double *__restrict__ mdot) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= Fi) return;
  mdot[f] = mdot[f] + cf[f] * pc[own[f]] + ff[f] * pc[nei[f]];
}

// UpdateBoundaryPressure at walls (CorrectPressure.for:49-90)
__global__ void k_wall_p(int Fb, int Fi, const int *__restrict__ own, const double *__restrict__ dx,
                         const double *__restrict__ dy, const double *__restrict__ dz,
                         const double *__restrict__ p, const double *__restrict__ bgp,
                         const int *__restrict__ btype, double *__restrict__ bp) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= Fb) return;
  if (btype && btype[b] == FS_BC_PERIODIC) return; // interpolated between the pair's cells
  const int f = Fi + b;
  bp[b] = p[own[f]] + bgp[b] * dx[f] + bgp[Fb + b] * dy[f] + bgp[2 * Fb + b] * dz[f];
}

// UpdateSymmetry (SymmetryAndAxisBoundary.for:27-44): boundary velocity of a
// symmetry face = the owner's velocity less its normal component.
__global__ void k_update_symmetry(int Fb, int Fi, const int *__restrict__ own,
                                  const int *__restrict__ btype, const double *__restrict__ nx,
                                  const double *__restrict__ ny, const double *__restrict__ nz,
                                  const double *__restrict__ u, const double *__restrict__ v,
                                  const double *__restrict__ w, double *__restrict__ bu,
                                  double *__restrict__ bv, double *__restrict__ bw) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= Fb || btype[b] != FS_BC_SYMMETRY) return;
  const int f = Fi + b, i1 = own[f];
  const double ax = nx[f], ay = ny[f], az = nz[f];
  const double amdot = u[i1] * ax + v[i1] * ay + w[i1] * az;
  bu[b] = u[i1] - amdot * ax;
  bv[b] = v[i1] - amdot * ay;
  bw[b] = w[i1] - amdot * az;
}

// L1 of the cell mass imbalance: per-cell sum of outgoing mdot in slot order
__global__ void k_mass_l1(int N, int W, int Fi, const int *__restrict__ code,
                          const double *__restrict__ mdot, const double *__restrict__ bmdot,
                          double *__restrict__ part) {
  double s = 0.0;
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < N; c += gridDim.x * kBlock) {
    double r = 0.0;
    for (int t = 0; t < W; ++t) {
      const int cd = code[(size_t)t * N + c];
      const int role = cd & 3, f = cd >> 2;
      if (role == ROLE_OWNER) r = r + mdot[f];
      else if (role == ROLE_NEIGH) r = r - mdot[f];
      else if (role == ROLE_BOUND && bmdot) r = r + bmdot[f - Fi];
    }
    s += fabs(r);
  }
  s = block_sum<kBlock>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

} // namespace

struct SimpleCase {
  HostGeom geom;
  fs_mesh *mesh = nullptr;
  hipStream_t s = nullptr;
  int N = 0, W = 0, F = 0, Fi = 0, Fb = 0;
  DevBuf<int> own, nei, code;
  DevBuf<double> Sx, Sy, Sz, A, nx, ny, nz, dx, dy, dz, ex, ey, ez, d, g, gdiff, Tx, Ty, Tz, kx,
      ky, kz, vol;
  DevBuf<double> bu, bv, bw;
  DevBuf<double> u, v, w, p, pc, mdot, gu, gv, gw, gp, gpc, bp, bgp, du, dv, dw, us, vs, ws;
  DevBuf<double> fp, fv, cf, ff, tf, ac, anb, bc, dphi, part;
  DevBuf<double> tf3; // momentum FluxTf of u, v, w (F each)
  DevBuf<double> pcprev; // the last converged p' (input of the warm-start network)
  bool have_pcprev = false;
  // the last ring_m converged p' fields (FS_OPT_WARM_PREVIOUS >= 2), the
  // latest at ring[(ring_head + ring_m - 1) % ring_m]
  DevBuf<double> ring[kProjMax];
  int ring_m = 0, ring_n = 0, ring_head = 0;
  DevBuf<int> btype, bzq; // boundary condition per boundary face (empty: all walls)
  // high-resolution convection and gradient limiting; iter = nIter, the
  // number of the outer iteration in progress (0 before the first)
  int hr_fw = FS_HR_NONE, hr_scheme = FS_CONV_UPWIND, hr_start = 0, lim_mom = 0, lim_cont = 0, iter = 0;
  double bleed = 0.0;
  DevBuf<double> fcx, fcy, fcz, xc, yc, zc; // centroids, uploaded with the first limiter
  DevBuf<double> lim_part, lim_eps;
  int ngrad_mom = 2, ngrad_cont = 2; // correction sweeps (nIterGradient) of u,v,w and of p, p'
  int decomp = 3;                  // MethodDecomposeS of the E/T split and of Rhie-Chow
  int gi_mom = FS_GRADINT_AVERAGE, gi_cont = FS_GRADINT_AVERAGECORRECTED; // face-gradient schemes
  std::vector<int> h_own;          // owner of every face (host copy, 0-based)
  std::vector<int> h_nei;          // neighbour of every interior face
  bool simplec = false;            // Algorithm 'simplec': D2 in du2/dv2/dw2
  // restart-file state the iteration forms but does not keep: the face
  // gradients of u, v, w and p and the effective divergence
  bool capture = false;
  int captured_iter = -1;          // iteration whose state the buffers hold
  DevBuf<double> cap_fgm, cap_fgp, cap_eff;
  DevBuf<double> du2, dv2, dw2;
  // momentum systems in colour-major cell order (the colour classes of the
  // multicolour Gauss-Seidel contiguous): their solver context on the
  // renumbered mesh (sharing the stream), the renumbering, the permuted
  // unknown, and whether ac/anb/bc currently hold such a system
  fs_mesh *mom_mesh = nullptr;
  DevBuf<int> mperm;
  std::vector<int> h_mperm;
  DevBuf<double> mx;
  bool sys_perm = false;
  bool sym_dirty = false;  // symmetry boundary velocities not yet derived
  double rho = 1.0, mu = 0.01, urf_u = 0.7, urf_p = 0.3;
  int msolver = FS_SOR, mlmg = 0, mitmax = 20, psolver = FS_BICGS, plmg = FS_MG_PRECOND,
      pitmax = 500, refcell = 0;
  double mrrf = 0.1, prrf = 1e-8;
  fs_simple_stats last{};
  fs_mpnn *warm = nullptr;  // p' initial guess from the network
  fs_mpnn *sink = nullptr;  // p' snapshots into the network's dataset
  // translational periodicity: the pairs, the boundary mass fluxes (Bmdot,
  // zero off the periodic faces), the periodic faces' boundary gradients of
  // u, v, w, the boundary p', and per boundary face its periodic list index
  Periodic per;
  DevBuf<double> bmdot, bgu, bgv, bgw, bpc;
  DevBuf<int> pk;
  bool periodic() const { return per.np > 0; }
  PerArgs pargs() const { return periodic_args(per, N, Fi, Fb, own.p); }
  // the k-omega SST model (fs_case_set_sst): its settings, fields and the
  // iteration counts of the last k and omega solves
  struct Turb {
    bool on = false;
    int ksolver = FS_ILU, kitmax = 20, wsolver = FS_ILU, witmax = 20, kfw = FS_HR_NONE, kscheme = FS_CONV_UPWIND,
        wfw = FS_HR_NONE, wscheme = FS_CONV_UPWIND, kngrad = 2, wngrad = 2, klim = 0, wlim = 0;
    double urf_k = 0.5, urf_w = 0.5, urf_mut = 0.5, krrf = 0.1, wrrf = 0.1, max_ratio = 1.0e6;
    DevBuf<double> k, om, mut, wd, bk, bom, bmut, gk, gom, bgk, bgom, rhok, brhok, grk, bgrk, F1, F2, bF1, bF2,
        prod, sr, bsr, med, bmed;
    int it_k = 0, it_w = 0;
  } tb;
  SstArgs sst() const {
    SstArgs a{};
    a.N = N; a.Fi = Fi; a.Fb = Fb;
    a.own = own.p; a.nei = nei.p; a.btype = btype.p; a.pk = pk.p;
    a.rho = rho; a.mu = mu; a.urf_mut = tb.urf_mut; a.max_ratio = tb.max_ratio;
    a.gu = gu.p; a.gv = gv.p; a.gw = gw.p; a.bgu = bgu.p; a.bgv = bgv.p; a.bgw = bgw.p;
    a.mdot = mdot.p; a.bmdot = bmdot.p;
    a.k = tb.k.p; a.om = tb.om.p; a.mut = tb.mut.p; a.wd = tb.wd.p; a.bk = tb.bk.p; a.bom = tb.bom.p;
    a.bmut = tb.bmut.p; a.gk = tb.gk.p; a.gom = tb.gom.p; a.bgk = tb.bgk.p; a.bgom = tb.bgom.p;
    a.rhok = tb.rhok.p; a.brhok = tb.brhok.p; a.F1 = tb.F1.p; a.F2 = tb.F2.p; a.bF1 = tb.bF1.p; a.bF2 = tb.bF2.p;
    a.prod = tb.prod.p; a.sr = tb.sr.p; a.bsr = tb.bsr.p; a.med = tb.med.p; a.bmed = tb.bmed.p;
    return a;
  }
};

static void upload_geometry(SimpleCase &C) {
  hipStream_t s = C.s;
  const FaceGeom &F = C.geom.face;
  C.Sx.upload(F.Sx, s); C.Sy.upload(F.Sy, s); C.Sz.upload(F.Sz, s); C.A.upload(F.A, s);
  C.nx.upload(F.nx, s); C.ny.upload(F.ny, s); C.nz.upload(F.nz, s);
  C.dx.upload(F.dx, s); C.dy.upload(F.dy, s); C.dz.upload(F.dz, s);
  C.ex.upload(F.ex, s); C.ey.upload(F.ey, s); C.ez.upload(F.ez, s); C.d.upload(F.d, s);
  C.g.upload(F.g, s); C.gdiff.upload(F.gdiff, s);
  C.Tx.upload(F.Tx, s); C.Ty.upload(F.Ty, s); C.Tz.upload(F.Tz, s);
  C.kx.upload(F.kx, s); C.ky.upload(F.ky, s); C.kz.upload(F.kz, s);
  C.vol.upload(C.geom.vol, s);
}

// Gradient of up to four fields: Green-Gauss plus nIterGradient = 2
// face-centroid corrections (CalculateGradients.for:212-215).
// Calls f(std::integral_constant<int, WC>) with WC = the case's slots per cell
// when the cell kernels have a fixed-width form for it, else WC = 0.
template <class Fn> static void by_width(int W, Fn &&fn) {
  switch (W) {
  case 4: fn(std::integral_constant<int, 4>()); break;
  case 5: fn(std::integral_constant<int, 5>()); break;
  case 6: fn(std::integral_constant<int, 6>()); break;
  case 8: fn(std::integral_constant<int, 8>()); break;
  default: fn(std::integral_constant<int, 0>());
  }
}

template <int NF> static void gradients(SimpleCase &C, GradArgs a, int ncorr) {
  const int Fi = C.Fi, F = C.F, N = C.N;
  hipLaunchKernelGGL(k_gg_face0<NF>, dim3(blocks(F)), dim3(kBlock), 0, C.s, Fi, F, C.own.p, C.nei.p,
                     C.g.p, a, C.fp.p);
  auto cell = [&](const double *fv) {
    by_width(C.W, [&](auto wc) {
      hipLaunchKernelGGL((k_gg_cell<NF, decltype(wc)::value>), dim3(blocks(N)), dim3(kBlock), 0, C.s, N, C.W, F,
                         C.code.p, C.Sx.p, C.Sy.p, C.Sz.p, C.vol.p, fv, a);
    });
  };
  cell(C.fp.p);
  for (int it = 0; it < ncorr; ++it) {
    hipLaunchKernelGGL(k_gg_facecorr<NF>, dim3(blocks(F)), dim3(kBlock), 0, C.s, Fi, F, N, C.own.p,
                       C.nei.p, C.g.p, C.kx.p, C.ky.p, C.kz.p, C.fp.p, a, C.fv.p);
    cell(C.fv.p);
  }
}

// The face and cell centroids the limiters read, uploaded once.
static void upload_centroids(SimpleCase &C) {
  if (C.xc.p) return;
  const auto &F = C.geom.face;
  C.fcx.upload(F.cx, C.s);
  C.fcy.upload(F.cy, C.s);
  C.fcz.upload(F.cz, C.s);
  C.xc.upload(C.geom.xc, C.s);
  C.yc.upload(C.geom.yc, C.s);
  C.zc.upload(C.geom.zc, C.s);
  FS_HIP_THROW(hipStreamSynchronize(C.s));
}

// The limiters of the fields of the gradient launch just made (its boundary
// face values are still in fp), methods[k] per field.
template <int NF> static void limit_gradients(SimpleCase &C, const GradArgs &g, const int *methods) {
  LimArgs a{};
  bool any = false, range = false;
  for (int k = 0; k < NF; ++k) {
    a.phi[k] = g.phi[k];
    a.grad[k] = g.grad[k];
    a.method[k] = methods[k];
    any = any || methods[k] != 0;
    range = range || methods[k] == 2;
  }
  if (!any) return;
  if (!C.xc.p || !C.fcx.p) throw std::runtime_error("gradient limiter without the centroids uploaded");
  if (range) {
    const int nb = std::min(256, blocks(C.N + C.Fb));
    C.lim_part.alloc((size_t)2 * NF * nb);
    C.lim_eps.alloc(4);
    hipLaunchKernelGGL(k_lim_range<NF>, dim3(nb), dim3(kBlock), 0, C.s, C.N, C.Fi, C.F, C.fp.p, a,
                       C.lim_part.p);
    hipLaunchKernelGGL(k_lim_eps<NF>, dim3(1), dim3(kBlock), 0, C.s, nb, C.lim_part.p, C.lim_eps.p);
    a.eps2 = C.lim_eps.p;
  }
  hipLaunchKernelGGL(k_limit<NF>, dim3(blocks(C.N)), dim3(kBlock), 0, C.s, C.N, C.W, C.F, C.code.p,
                     C.mesh->d_nbr.p, C.fcx.p, C.fcy.p, C.fcz.p, C.xc.p, C.yc.p, C.zc.p, C.fp.p, a);
}

static void phase_momentum_gradients(SimpleCase &C) {
  GradArgs a{};
  a.phi[0] = C.u.p; a.phi[1] = C.v.p; a.phi[2] = C.w.p; a.phi[3] = C.p.p;
  a.bval[0] = C.bu.p; a.bval[1] = C.bv.p; a.bval[2] = C.bw.p; a.bval[3] = C.bp.p;
  a.grad[0] = C.gu.p; a.grad[1] = C.gv.p; a.grad[2] = C.gw.p; a.grad[3] = C.gp.p;
  // LimitGradientVariable after each gradient, before the boundary
  // corrections (CalculateGradients.for:252-272); u, v, w and p in one pass
  // when they take the same number of correction sweeps
  if (C.ngrad_mom == C.ngrad_cont) {
    gradients<4>(C, a, C.ngrad_mom);
    const int methods[4] = {C.lim_mom, C.lim_mom, C.lim_mom, C.lim_cont};
    limit_gradients<4>(C, a, methods);
  } else {
    gradients<3>(C, a, C.ngrad_mom);
    const int mm[3] = {C.lim_mom, C.lim_mom, C.lim_mom};
    limit_gradients<3>(C, a, mm);
    GradArgs b{};
    b.phi[0] = C.p.p; b.bval[0] = C.bp.p; b.grad[0] = C.gp.p;
    gradients<1>(C, b, C.ngrad_cont);
    const int mc[1] = {C.lim_cont};
    limit_gradients<1>(C, b, mc);
  }
  if (C.periodic()) {
    // CorrectGradientAtPeriodicBoundary after the limiters, before the
    // pressure's symmetry and wall projections (CalculateGradients.for:263)
    const double *cg[4] = {C.gu.p, C.gv.p, C.gw.p, C.gp.p};
    double *bg[4] = {C.bgu.p, C.bgv.p, C.bgw.p, C.bgp.p};
    periodic_bgrad(C.pargs(), C.s, 4, cg, bg);
  }
  hipLaunchKernelGGL(k_wall_pgrad, dim3(blocks(C.Fb)), dim3(kBlock), 0, C.s, C.Fb, C.Fi, C.N,
                     C.own.p, C.nx.p, C.ny.p, C.nz.p, C.btype.p, C.bzq.p, C.gp.p, C.bgp.p);
  if (C.tb.on) {
    // CalculateEffectiveDiffusionCoefficient of the velocities (Viscosity +
    // TurbulentViscosity) and MinusTwoThirdsRhoK('velx'): 2/3 rho k and its
    // gradient with the momentum's gradient settings (its wall and
    // symmetry projections feed nothing here)
    const SstArgs sa = C.sst();
    sst_ediff(sa, C.s, 2);
    sst_rhok(sa, C.s);
    GradArgs r{};
    r.phi[0] = C.tb.rhok.p; r.bval[0] = C.tb.brhok.p; r.grad[0] = C.tb.grk.p;
    gradients<1>(C, r, C.ngrad_mom);
    const int mr[1] = {C.lim_mom};
    limit_gradients<1>(C, r, mr);
    if (C.periodic()) {
      const double *cg[1] = {C.tb.grk.p};
      double *bg[1] = {C.tb.bgrk.p};
      periodic_bgrad(C.pargs(), C.s, 1, cg, bg);
    }
  }
  // the interior momentum face fluxes of all three components, with the
  // high-resolution correction from iteration nIterStartApplyingHR on
  // (AssembleConvection.for:156-165)
  const int hrf = (C.hr_fw != FS_HR_NONE && C.hr_scheme != FS_CONV_UPWIND && C.iter >= C.hr_start)
                      ? C.hr_fw : FS_HR_NONE;
  auto face3 = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(blocks(C.Fi)), dim3(kBlock), 0, C.s, C.Fi, C.F, C.N, C.own.p, C.nei.p,
                       C.g.p, C.gdiff.p, C.Tx.p, C.Ty.p, C.Tz.p, C.Sx.p, C.Sy.p, C.Sz.p, C.mu,
                       C.tb.on ? (const double *)C.tb.med.p : nullptr, (double)(-2.0f / 3.0f), C.u.p, C.v.p, C.w.p, C.gu.p, C.gv.p, C.gw.p, C.mdot.p, C.dx.p,
                       C.dy.p, C.dz.p, C.hr_scheme, C.bleed, C.gi_mom, C.rho, C.ex.p, C.ey.p, C.ez.p, C.d.p,
                       C.cf.p, C.ff.p, C.tf3.p, C.capture ? C.cap_fgm.p : nullptr);
  };
  if (hrf == FS_HR_TVD) face3(k_mom_face3<FS_HR_TVD>);
  else if (hrf == FS_HR_NVF) face3(k_mom_face3<FS_HR_NVF>);
  else face3(k_mom_face3<FS_HR_NONE>);
}

// The momentum systems go to colour-major order only for the multicolour
// Gauss-Seidel ('sor') solve, whose result depends on the colour classes
// alone; every other solver sees the reference's numbering (ILU factors and
// Krylov sums depend on it).
static bool mom_permuted(const SimpleCase &C) { return C.mom_mesh && C.msolver == FS_SOR && !C.mlmg; }

static void phase_momentum_assemble(SimpleCase &C, int comp) {
  MomArgs a{};
  const bool perm = mom_permuted(C);
  C.sys_perm = perm;
  a.comp = comp;
  a.mu = C.mu;
  a.c23 = (double)(-2.0f / 3.0f); // REAL(4) literal -2./3. (AssembleBulkViscosity.for:49)
  a.bgam = C.tb.on ? C.tb.bmed.p : nullptr;
  a.phi = comp == 0 ? C.u.p : comp == 1 ? C.v.p : C.w.p;
  a.u = C.u.p; a.v = C.v.p; a.w = C.w.p;
  a.bu = C.bu.p; a.bv = C.bv.p; a.bw = C.bw.p;
  a.btype = C.btype.p;
  a.gu = C.gu.p; a.gv = C.gv.p; a.gw = C.gw.p;
  a.mdot = C.mdot.p;
  // interior faces: k_mom_face3 (momentum gradient phase); boundary faces of
  // this component, then the cell gather
  double *tf = C.tf3.p + (size_t)comp * C.F;
  hipLaunchKernelGGL(k_mom_wall, dim3(blocks(C.Fb)), dim3(kBlock), 0, C.s, C.Fb, C.Fi, C.N, C.own.p,
                     C.A.p, C.nx.p, C.ny.p, C.nz.p, C.dx.p, C.dy.p, C.dz.p, C.Sx.p, C.Sy.p, C.Sz.p, a,
                     C.cf.p, C.ff.p, tf);
  const bool per = C.periodic();
  if (per)
    periodic_momentum(C.pargs(), C.s, comp, C.mu, a.c23, C.tb.on ? C.tb.med.p : nullptr,
                      C.tb.on ? C.tb.bmed.p : nullptr, a.phi, C.u.p, C.v.p, C.w.p, C.bgu.p, C.bgv.p, C.bgw.p,
                      C.Sx.p, C.Sy.p, C.Sz.p, C.bmdot.p, C.cf.p, C.ff.p, tf);
  double *D = comp == 0 ? C.du.p : comp == 1 ? C.dv.p : C.dw.p;
  double *st = comp == 0 ? C.us.p : comp == 1 ? C.vs.p : C.ws.p;
  const double *lin = per ? C.per.scal.p + PSCAL_LIN + comp : nullptr;
  by_width(C.W, [&](auto wc) {
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(blocks(C.N)), dim3(kBlock), 0, C.s, C.N, C.W, C.Fi, C.code.p, C.cf.p,
                         C.ff.p, tf, C.mdot.p, C.bmdot.p, lin, a.phi, C.gp.p + (size_t)comp * C.N, C.vol.p,
                         C.urf_u, perm ? C.mperm.p : nullptr, C.ac.p, C.anb.p, C.bc.p, D, st, C.mx.p,
                         C.simplec ? (comp == 0 ? C.du2.p : comp == 1 ? C.dv2.p : C.dw2.p) : nullptr,
                         C.capture ? C.cap_eff.p : nullptr,
                         C.tb.on ? (const double *)C.tb.grk.p + (size_t)comp * C.N : nullptr);
    };
    if (per) launch(k_mom_cell<decltype(wc)::value, true>);
    else launch(k_mom_cell<decltype(wc)::value, false>);
  });
}

// The momentum context follows the solver options set on the case's context.
static void mirror_options(fs_mesh *d, const fs_mesh *s) {
  const bool structural = d->mg_maxlevels != s->mg_maxlevels || d->mg_minparents != s->mg_minparents ||
                          d->mg_tail_cells != s->mg_tail_cells || d->mg_sell_rows != s->mg_sell_rows;
  d->mg_pre = s->mg_pre;
  d->mg_post = s->mg_post;
  d->mg_maxcycles = s->mg_maxcycles;
  d->mg_maxlevels = s->mg_maxlevels;
  d->mg_minparents = s->mg_minparents;
  d->mg_reagg = s->mg_reagg;
  d->mg_fp32 = s->mg_fp32;
  d->mg_smoother = s->mg_smoother;
  d->mg_omega = s->mg_omega;
  d->mg_tail_cells = s->mg_tail_cells;
  d->mg_coarse_sweeps = s->mg_coarse_sweeps;
  d->mg_cpre = s->mg_cpre;
  d->mg_cpost = s->mg_cpost;
  d->mg_sell_rows = s->mg_sell_rows;
  d->mg_alpha = s->mg_alpha;
  d->gmres_m = s->gmres_m;
  d->check_every = s->check_every;
  if (structural) d->amg.reset();
}

static int phase_momentum_solve(SimpleCase &C, int comp, fs_simple_stats *st) {
  double *phi = comp == 0 ? C.u.p : comp == 1 ? C.v.p : C.w.p;
  fs_solve_stats ss{};
  double res4[4];
  int rc;
  if (mom_permuted(C)) {
    // the system is in colour-major order, and so is the unknown the
    // assembly copied into mx: solve for it, then bring it back
    mirror_options(C.mom_mesh, C.mesh);
    rc = solve_device(C.mom_mesh, comp + 1, C.mx.p, C.mrrf, C.mitmax, C.msolver, C.mlmg, C.ac.p, C.anb.p, C.bc.p,
                      C.dphi.p, res4, &ss);
    if (rc < 0) return rc;
    hipLaunchKernelGGL(k_unpermute, dim3(blocks(C.N)), dim3(kBlock), 0, C.s, C.N, C.mperm.p, C.mx.p, phi);
  } else {
    rc = solve_device(C.mesh, comp + 1, phi, C.mrrf, C.mitmax, C.msolver, C.mlmg, C.ac.p, C.anb.p, C.bc.p,
                      C.dphi.p, res4, &ss);
    if (rc < 0) return rc;
  }
  if (C.periodic()) { // UpdatePeriodic (SolveXMomentumEquation.for:299)
    const double *ph[3] = {C.u.p, C.v.p, C.w.p};
    double *bph[3] = {C.bu.p, C.bv.p, C.bw.p};
    periodic_values(C.pargs(), C.s, 3, ph, bph);
  }
  if (st) {
    st->momentum_iterations += ss.iterations;
    st->momentum_seconds += ss.seconds;
    st->res_momentum[comp] = res4[0];
  }
  return 0;
}

static void phase_pc_assemble(SimpleCase &C) {
  RcArgs a{};
  C.sys_perm = false;
  a.rho = C.rho;
  a.urf_u = C.urf_u;
  a.u = C.u.p; a.v = C.v.p; a.w = C.w.p;
  a.us = C.us.p; a.vs = C.vs.p; a.ws = C.ws.p;
  a.du = C.du.p; a.dv = C.dv.p; a.dw = C.dw.p;
  a.gp = C.gp.p;
  a.p = C.p.p;
  a.du2 = C.du2.p; a.dv2 = C.dv2.p; a.dw2 = C.dw2.p;
  a.gint = C.gi_cont;
  a.fgp_out = C.capture ? C.cap_fgp.p : nullptr;
  a.rhok = C.tb.on ? C.tb.rhok.p : nullptr;
  a.grk = C.tb.on ? C.tb.grk.p : nullptr;
  auto rc = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(blocks(C.Fi)), dim3(kBlock), 0, C.s, C.Fi, C.N, C.own.p, C.nei.p, C.g.p,
                       C.Sx.p, C.Sy.p, C.Sz.p, C.ex.p, C.ey.p, C.ez.p, C.d.p, a, C.mdot.p, C.cf.p, C.ff.p,
                       C.tf.p);
  };
  switch (C.decomp * 2 + (C.simplec ? 1 : 0)) {
  case 2: rc(k_rhie_chow<false, 1>); break;
  case 3: rc(k_rhie_chow<true, 1>); break;
  case 4: rc(k_rhie_chow<false, 2>); break;
  case 5: rc(k_rhie_chow<true, 2>); break;
  case 7: rc(k_rhie_chow<true, 3>); break;
  default: rc(k_rhie_chow<false, 3>);
  }
  const bool per = C.periodic();
  if (per) {
    const PerArgs pa = C.pargs();
    periodic_rhie_chow(pa, C.s, C.decomp, C.rho, C.urf_u, C.u.p, C.v.p, C.w.p, C.us.p, C.vs.p, C.ws.p, C.du.p,
                       C.dv.p, C.dw.p, C.p.p, C.bgp.p, C.Sx.p, C.Sy.p, C.Sz.p,
                       C.tb.on ? C.tb.rhok.p : nullptr, C.tb.on ? C.tb.bgrk.p : nullptr, C.bmdot.p);
    periodic_pc_boundary(pa, C.s, C.decomp, C.rho, C.pk.p, C.simplec ? C.du2.p : C.du.p,
                         C.simplec ? C.dv2.p : C.dv.p, C.simplec ? C.dw2.p : C.dw.p, C.Sx.p, C.Sy.p, C.Sz.p,
                         C.bmdot.p, C.cf.p, C.ff.p, C.tf.p);
  }
  by_width(C.W, [&](auto wc) {
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(blocks(C.N)), dim3(kBlock), 0, C.s, C.N, C.W, C.code.p, C.cf.p, C.ff.p,
                         C.tf.p, C.ac.p, C.anb.p, C.bc.p);
    };
    if (per) launch(k_pc_cell<decltype(wc)::value, true>);
    else launch(k_pc_cell<decltype(wc)::value, false>);
  });
}

static void update_symmetry(SimpleCase &C) {
  C.sym_dirty = false;
  if (!C.btype.p) return;
  hipLaunchKernelGGL(k_update_symmetry, dim3(blocks(C.Fb)), dim3(kBlock), 0, C.s, C.Fb, C.Fi, C.own.p,
                     C.btype.p, C.nx.p, C.ny.p, C.nz.p, C.u.p, C.v.p, C.w.p, C.bu.p, C.bv.p, C.bw.p);
}

static int phase_pc_solve_correct(SimpleCase &C, fs_simple_stats *st) {
  FS_HIP(hipMemsetAsync(C.pc.p, 0, sizeof(double) * C.N, C.s)); // InitializePressureCorrection
  fs_solve_stats ss{};
  double res4[4];
  const bool warm = C.warm != nullptr;
  if (C.pcprev.n != (size_t)C.N) {
    C.pcprev.alloc(C.N);
    FS_HIP(hipMemsetAsync(C.pcprev.p, 0, sizeof(double) * C.N, C.s));
  }
  if (warm) {
    const int rp = mpnn_predict_device(C.warm, C.s, C.bc.p, C.ac.p, C.anb.p, C.vol.p, C.pcprev.p, C.dphi.p);
    if (rp < 0) return rp;
  }
  // FS_OPT_WARM_PREVIOUS: without a network, the previous converged p' as the
  // initial guess, or (2..4) the combination of the last m p' fields with the
  // smallest residual for this system; either is then scaled and accepted or
  // dropped like any guess
  const int pm = C.mesh->warm_prev;
  if (pm >= 2 && pm != C.ring_m) {
    C.ring_m = pm;
    C.ring_n = C.ring_head = 0;
  }
  const bool proj = !warm && pm >= 2 && C.ring_n > 0;
  const bool prev = !warm && (proj || (pm == 1 && C.have_pcprev));
  if (proj) {
    const double *X[kProjMax];
    for (int a = 0; a < C.ring_n; ++a) X[a] = C.ring[(C.ring_head + 2 * C.ring_m - 1 - a) % C.ring_m].p;
    try {
      project_guess(C.mesh, C.ac.p, C.anb.p, C.bc.p, X, C.ring_n, C.dphi.p);
    } catch (const std::exception &e) {
      set_error(e.what());
      return -2;
    }
  } else if (prev) {
    FS_HIP(hipMemcpyAsync(C.dphi.p, C.pcprev.p, sizeof(double) * C.N, hipMemcpyDeviceToDevice, C.s));
  }
  const int rc = solve_device(C.mesh, 4, C.pc.p, C.prrf, C.pitmax, C.psolver, C.plmg, C.ac.p,
                              C.anb.p, C.bc.p, C.dphi.p, res4, &ss, warm || prev);
  if (rc < 0) return rc;
  if (C.sink) mpnn_capture(C.sink, C.s, C.bc.p, C.dphi.p, C.ac.p, C.anb.p, C.vol.p, C.pcprev.p);
  FS_HIP(hipMemcpyAsync(C.pcprev.p, C.dphi.p, sizeof(double) * C.N, hipMemcpyDeviceToDevice, C.s));
  C.have_pcprev = true;
  if (pm >= 2) {
    DevBuf<double> &r = C.ring[C.ring_head];
    if (r.n != (size_t)C.N) r.alloc(C.N);
    FS_HIP(hipMemcpyAsync(r.p, C.dphi.p, sizeof(double) * C.N, hipMemcpyDeviceToDevice, C.s));
    C.ring_head = (C.ring_head + 1) % C.ring_m;
    if (C.ring_n < C.ring_m) ++C.ring_n;
  }
  if (st) {
    st->pc_iterations += ss.iterations;
    st->pc_seconds += ss.seconds;
    st->pc_initial_l1 = ss.initial_l1;
    st->pc_final_l1 = ss.final_l1;
    st->pc_levels = ss.levels;
    if (rc == 1) st->pc_unconverged++;
  }
  const int N = C.N;
  hipLaunchKernelGGL(k_correct_p, dim3(blocks(N)), dim3(kBlock), 0, C.s, N, C.urf_p, C.refcell,
                     C.pc.p, C.p.p);
  GradArgs a{};
  a.phi[0] = C.pc.p;
  a.bval[0] = nullptr; // walls: BPressureC = PressureC(owner)
  a.grad[0] = C.gpc.p;
  if (C.periodic()) { // UpdateBoundaryPressureCorrection (:263-290)
    const PerArgs pa = C.pargs();
    owner_values(pa, C.s, C.pc.p, C.bpc.p);
    const double *ph[1] = {C.pc.p};
    double *bph[1] = {C.bpc.p};
    periodic_values(pa, C.s, 1, ph, bph);
    a.bval[0] = C.bpc.p;
  }
  gradients<1>(C, a, C.ngrad_cont);
  const int methods[1] = {C.lim_cont};
  limit_gradients<1>(C, a, methods);
  // CorrectVelocity with D2 (= D1 for SIMPLE)
  hipLaunchKernelGGL(k_correct_u, dim3(blocks(N)), dim3(kBlock), 0, C.s, N, C.gpc.p,
                     C.simplec ? C.du2.p : C.du.p, C.simplec ? C.dv2.p : C.dv.p, C.simplec ? C.dw2.p : C.dw.p,
                     C.u.p, C.v.p, C.w.p);
  hipLaunchKernelGGL(k_correct_mdot, dim3(blocks(C.Fi)), dim3(kBlock), 0, C.s, C.Fi, C.own.p, C.nei.p,
                     C.cf.p, C.ff.p, C.pc.p, C.mdot.p);
  hipLaunchKernelGGL(k_wall_p, dim3(blocks(C.Fb)), dim3(kBlock), 0, C.s, C.Fb, C.Fi, C.own.p, C.dx.p,
                     C.dy.p, C.dz.p, C.p.p, C.bgp.p, C.btype.p, C.bp.p);
  if (C.periodic()) {
    const PerArgs pa = C.pargs();
    const double *ph[1] = {C.p.p};
    double *bph[1] = {C.bp.p};
    periodic_values(pa, C.s, 1, ph, bph);
    const double *d2u = C.simplec ? C.du2.p : C.du.p, *d2v = C.simplec ? C.dv2.p : C.dv.p,
                 *d2w = C.simplec ? C.dw2.p : C.dw.p;
    periodic_correct_bmdot(pa, C.s, d2u, d2v, d2w, C.gpc.p, C.pc.p, C.cf.p, C.ff.p, C.bu.p, C.bv.p, C.bw.p,
                           C.bmdot.p);
    if (C.per.calc_beta && C.iter % C.per.every == 0)
      periodic_beta(pa, C.s, C.decomp, C.rho, C.per.mdot, C.per.relax, C.per.iters, C.per.tuse, d2u, d2v, d2w,
                    C.Sx.p, C.Sy.p, C.Sz.p, C.own.p, C.nei.p, C.g.p, C.ex.p, C.ey.p, C.ez.p, C.d.p, C.u.p, C.v.p,
                    C.w.p, C.bu.p, C.bv.p, C.bw.p, C.bmdot.p, C.mdot.p);
  }
  update_symmetry(C); // SolveContinuityEquation.for:96
  return 0;
}

// One k or omega equation of the SST model (SolveTKE.for, SolveTOmega.for):
// gradient (Green-Gauss, limiter, periodic correction), effective diffusion
// coefficient, sources (production, or the blending functions), face fluxes,
// the cell gather, SolveEquation, bounding and boundary values.
static int phase_sst_equation(SimpleCase &C, int var, fs_simple_stats *st) {
  auto &T = C.tb;
  const SstArgs a = C.sst();
  const bool per = C.periodic();
  const PerArgs pa = C.pargs();
  GradArgs g{};
  g.phi[0] = var ? T.om.p : T.k.p;
  g.bval[0] = var ? T.bom.p : T.bk.p;
  g.grad[0] = var ? T.gom.p : T.gk.p;
  gradients<1>(C, g, var ? T.wngrad : T.kngrad);
  const int m[1] = {var ? T.wlim : T.klim};
  limit_gradients<1>(C, g, m);
  if (per) {
    const double *cg[1] = {g.grad[0]};
    double *bg[1] = {var ? T.bgom.p : T.bgk.p};
    periodic_bgrad(pa, C.s, 1, cg, bg);
  }
  sst_ediff(a, C.s, var);
  if (var == 0) sst_production(a, C.s);
  else sst_blending(a, C.s);
  const int fw = C.iter >= C.hr_start ? (var ? T.wfw : T.kfw) : FS_HR_NONE;
  sst_faces(a, per ? &pa : nullptr, C.s, var, fw, var ? T.wscheme : T.kscheme, C.g.p, C.gdiff.p, C.Tx.p, C.Ty.p,
            C.Tz.p, C.dx.p, C.dy.p, C.dz.p, C.cf.p, C.ff.p, C.tf.p);
  C.sys_perm = false;
  sst_cells(a, per, C.s, var, C.W, C.code.p, C.cf.p, C.ff.p, C.tf.p, C.vol.p, var ? T.urf_w : T.urf_k, C.ac.p,
            C.anb.p, C.bc.p);
  fs_solve_stats ss{};
  double res4[4];
  const int rc = solve_device(C.mesh, var ? 10 : 8, var ? T.om.p : T.k.p, var ? T.wrrf : T.krrf,
                              var ? T.witmax : T.kitmax, var ? T.wsolver : T.ksolver, 0, C.ac.p, C.anb.p, C.bc.p,
                              C.dphi.p, res4, &ss);
  if (rc < 0) return rc;
  (var ? T.it_w : T.it_k) = ss.iterations;
  if (st) st->turbulence_iterations += ss.iterations;
  sst_bound(a, pa, C.s, var);
  return 0;
}

// CalculateNormalDistanceToWall, method 2 (CalculateDistanceToWall.for
// :154-470): passes of assembly (Green-Gauss gradient, method 2 with two
// sweeps, 'average' face gradients) and solve until a pass starts from a
// residual |bc|_1 at most tol times the first pass's, then the distance from
// the last pass's gradient, the cells at the walls their faces' normal
// distance.  The reference stops on the inner solve's final residual (1e-10)
// instead; the fixed point is the same.  The solves are CG preconditioned by
// the AMG V-cycle; the hierarchy is dropped afterwards so that the p' solves
// aggregate their own operator.
static int wall_distance(SimpleCase &C, double tol, int maxouter, int *passes) {
  const int N = C.N, Fb = C.Fb;
  auto &T = C.tb;
  std::vector<int> bt(Fb, FS_BC_WALL);
  if (C.btype.p) FS_HIP(hipMemcpy(bt.data(), C.btype.p, sizeof(int) * Fb, hipMemcpyDeviceToHost));
  std::vector<double> hb(Fb);
  for (int b = 0; b < Fb; ++b) hb[b] = bt[b] == FS_BC_WALL ? 0. : 1.;
  DevBuf<double> phi, bphi, gr;
  phi.upload(std::vector<double>(N, 1.0), C.s);
  bphi.upload(hb, C.s);
  gr.alloc(3 * (size_t)N);
  T.wd.alloc(N);
  double first = -1.0;
  int pass = 0;
  bool done = false;
  while (pass < maxouter && !done) {
    GradArgs g{};
    g.phi[0] = phi.p; g.bval[0] = bphi.p; g.grad[0] = gr.p;
    gradients<1>(C, g, 2);
    wall_distance_faces(C.s, C.Fi, Fb, N, C.own.p, C.nei.p, C.btype.p, C.g.p, C.gdiff.p, C.Tx.p, C.Ty.p, C.Tz.p,
                        phi.p, gr.p, bphi.p, C.cf.p, C.ff.p, C.tf.p);
    wall_distance_cells(C.s, N, C.W, C.code.p, C.cf.p, C.ff.p, C.tf.p, C.vol.p, phi.p, C.ac.p, C.anb.p, C.bc.p);
    C.sys_perm = false;
    fs_solve_stats ss{};
    double res4[4];
    const int rc = solve_device(C.mesh, 5, phi.p, 1e-8, 500, FS_CG, FS_MG_PRECOND, C.ac.p, C.anb.p, C.bc.p,
                                C.dphi.p, res4, &ss);
    if (rc < 0) return rc;
    if (first < 0.0) first = ss.initial_l1;
    done = !(ss.initial_l1 > tol * first);
    ++pass;
  }
  wall_distance_final(C.s, N, C.W, C.Fi, C.code.p, C.btype.p, phi.p, gr.p, C.dx.p, C.dy.p, C.dz.p, C.nx.p, C.ny.p,
                      C.nz.p, T.wd.p);
  FS_HIP(hipStreamSynchronize(C.s));
  C.mesh->amg.reset();
  if (passes) *passes = pass;
  if (!done) {
    set_error("fs_case_wall_distance: the passes did not reach the tolerance");
    return 1;
  }
  return 0;
}

// CalculateEffectsOfTurbulence for the SST model
// (SolveSelectedTurbulenceModelEquations.for:4-13,150-160): strain rate, the
// k and omega equations, the eddy viscosity.
static int phase_turbulence(SimpleCase &C, fs_simple_stats *st) {
  sst_strain(C.sst(), C.s);
  int rc = phase_sst_equation(C, 0, st);
  if (rc < 0) return rc;
  rc = phase_sst_equation(C, 1, st);
  if (rc < 0) return rc;
  sst_viscosity(C.sst(), C.s);
  return 0;
}

static double mass_l1(SimpleCase &C) {
  const int np = std::min(kPartials, blocks(C.N));
  hipLaunchKernelGGL(k_mass_l1, dim3(np), dim3(kBlock), 0, C.s, C.N, C.W, C.Fi, C.code.p, C.mdot.p,
                     C.periodic() ? C.bmdot.p : nullptr, C.part.p);
  std::vector<double> h(np);
  FS_HIP_THROW(hipMemcpyAsync(h.data(), C.part.p, sizeof(double) * np, hipMemcpyDeviceToHost, C.s));
  FS_HIP_THROW(hipStreamSynchronize(C.s));
  double t = 0.0;
  for (double x : h) t += x;
  return t;
}

// The geometry every periodic branch of the reference forms for face k from
// the partner cell's centroid carried by the partner patch's translation
// (e.g. PeriodicBoundary.for:54-67; the E/T split of
// AssembleIsotropicDiffusion.for:663-700 by MethodDecomposeS), in the
// reference's operation order, uploaded per periodic face.
void periodic_upload_geometry(SimpleCase &C) {
  Periodic &P = C.per;
  const FaceGeom &G = C.geom.face;
  std::vector<double> g(P.np), ux(P.np), uy(P.np), uz(P.np), dcf(P.np), gd(P.np), tx(P.np), ty(P.np), tz(P.np);
  for (int k = 0; k < P.np; ++k) {
    const int f = C.Fi + P.h_b[k], i1 = C.h_own[f], j1 = P.h_j1[k];
    const int q2 = P.h_pair[P.h_patch[k]];
    const double xf1 = C.geom.xc[j1] + P.h_trans[3 * q2], yf1 = C.geom.yc[j1] + P.h_trans[3 * q2 + 1],
                 zf1 = C.geom.zc[j1] + P.h_trans[3 * q2 + 2];
    const double ax = G.cx[f] - C.geom.xc[i1], ay = G.cy[f] - C.geom.yc[i1], az = G.cz[f] - C.geom.zc[i1];
    const double bx = G.cx[f] - xf1, by = G.cy[f] - yf1, bz = G.cz[f] - zf1;
    const double d1 = sqrt(ax * ax + ay * ay + az * az), d2 = sqrt(bx * bx + by * by + bz * bz);
    g[k] = d2 / (d1 + d2);
    const double dx = xf1 - C.geom.xc[i1], dy = yf1 - C.geom.yc[i1], dz = zf1 - C.geom.zc[i1];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    dcf[k] = d;
    const double ex = dx / d, ey = dy / d, ez = dz / d;
    ux[k] = ex;
    uy[k] = ey;
    uz[k] = ez;
    const double sx = G.Sx[f], sy = G.Sy[f], sz = G.Sz[f], A = G.A[f];
    double fex, fey, fez, fe;
    if (C.decomp == 1) {
      const double dot = ex * sx + ey * sy + ez * sz;
      fex = dot * ex;
      fey = dot * ey;
      fez = dot * ez;
      fe = fabs(dot);
    } else if (C.decomp == 2) {
      fex = A * ex;
      fey = A * ey;
      fez = A * ez;
      fe = A;
    } else {
      const double dot = ex * sx + ey * sy + ez * sz;
      fex = (A * A / dot) * ex;
      fey = (A * A / dot) * ey;
      fez = (A * A / dot) * ez;
      fe = sqrt(fex * fex + fey * fey + fez * fez);
    }
    gd[k] = fe / d;
    tx[k] = sx - fex;
    ty[k] = sy - fey;
    tz[k] = sz - fez;
  }
  P.g.upload(g, C.s);
  P.ux.upload(ux, C.s);
  P.uy.upload(uy, C.s);
  P.uz.upload(uz, C.s);
  P.dcf.upload(dcf, C.s);
  P.gd.upload(gd, C.s);
  P.tx.upload(tx, C.s);
  P.ty.upload(ty, C.s);
  P.tz.upload(tz, C.s);
}

// The solver contexts of the case from its topology (C.geom.enbr, with the
// periodic partners once the pairs are set): the mesh context with its face
// codes, and the momentum context on the colour-major renumbering (the greedy
// colouring of the mesh, the Gauss-Seidel colours, colour by colour in
// increasing cell order; greedy colouring of the renumbered mesh gives the
// same classes, now contiguous), sharing the case's stream.  Replaces any
// contexts the case held.
int build_meshes(SimpleCase &C) {
  const int nif = C.Fi, nfaces = C.F;
  std::vector<int> own1(nif), nei1(nif);
  for (int f = 0; f < nif; ++f) {
    own1[f] = C.h_own[f] + 1;
    nei1[f] = C.h_nei[f] + 1;
  }
  if (C.s) FS_HIP(hipStreamSynchronize(C.s));
  fs_mesh *mesh = nullptr, *mom = nullptr;
  int rc = fs_mesh_create(C.N, C.W, C.geom.enbr.data(), C.geom.nef.data(), nif, own1.data(), nei1.data(),
                          C.geom.iown.data(), C.geom.inbr.data(), &mesh);
  if (rc == 0) rc = fs_mesh_set_cell_faces(mesh, nfaces, C.geom.gef.data(), C.geom.nef.data());
  if (rc != 0) {
    if (mesh) fs_mesh_destroy(mesh);
    return rc;
  }
  std::vector<int> ptr, cells;
  greedy_colors(C.N, C.W, mesh->h_nbr_dev.data(), ptr, cells);
  std::vector<int> perm(C.N, 0);
  for (int q = 0; q < C.N; ++q) perm[cells[q]] = q;
  std::vector<int> enbr_p((size_t)C.N * C.W, 0), nef_p(C.N, 0);
  for (int c = 0; c < C.N; ++c) {
    const int q = perm[c];
    nef_p[q] = C.geom.nef[c];
    for (int j = 0; j < C.W; ++j) {
      const int k = C.geom.enbr[(size_t)j * C.N + c];
      enbr_p[(size_t)j * C.N + q] = k > 0 ? perm[k - 1] + 1 : 0;
    }
  }
  rc = fs_mesh_create(C.N, C.W, enbr_p.data(), nef_p.data(), 0, nullptr, nullptr, nullptr, nullptr, &mom);
  if (rc != 0) {
    fs_mesh_destroy(mesh);
    return rc;
  }
  // the old contexts go (their solver options stay), the momentum one on the
  // new case stream
  if (C.mesh) mirror_options(mesh, C.mesh);
  if (C.mom_mesh) fs_mesh_destroy(C.mom_mesh);
  if (C.mesh) fs_mesh_destroy(C.mesh);
  C.mesh = mesh;
  C.mom_mesh = mom;
  C.s = mesh->stream;
  FS_HIP(hipStreamSynchronize(mom->stream));
  FS_HIP(hipStreamDestroy(mom->stream));
  mom->stream = C.s;
  mom->own_stream = false;
  C.h_mperm = perm;
  C.mperm.upload(C.h_mperm, C.s);
  C.mx.alloc(C.N);
  // the per-slot face code of the mesh context
  C.code.alloc((size_t)C.N * C.W);
  FS_HIP(hipMemcpyAsync(C.code.p, mesh->d_face_code.p, sizeof(int) * C.N * C.W, hipMemcpyDeviceToDevice, C.s));
  FS_HIP(hipStreamSynchronize(C.s));
  return 0;
}

int simple_iterate(SimpleCase &C, int niter, fs_simple_stats *st) {
  auto t0 = std::chrono::steady_clock::now();
  fs_simple_stats loc{};
  if (C.sym_dirty) update_symmetry(C); // the start-up UpdateSymmetry (MainProgram.for:128)
  for (int it = 0; it < niter; ++it) {
    C.iter = C.iter + 1; // nIter (MainProgram.for:257)
    if (C.capture) C.captured_iter = C.iter;
    if (C.periodic()) periodic_linear(C.pargs(), C.s, C.per.tuse);
    phase_momentum_gradients(C);
    for (int comp = 0; comp < 3; ++comp) {
      phase_momentum_assemble(C, comp);
      const int rc = phase_momentum_solve(C, comp, &loc);
      if (rc < 0) return rc;
    }
    phase_pc_assemble(C);
    int rc = phase_pc_solve_correct(C, &loc);
    if (rc < 0) return rc;
    if (C.tb.on) {
      rc = phase_turbulence(C, &loc);
      if (rc < 0) return rc;
    }
  }
  FS_HIP(hipStreamSynchronize(C.s));
  loc.iterations = niter;
  loc.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  loc.mass_l1 = mass_l1(C);
  C.last = loc;
  if (st) *st = loc;
  return 0;
}

} // namespace fs

using fs::set_error;

struct fs_case {
  fs::SimpleCase c;
};

#define CASE_GUARD_BEGIN try {
#define CASE_GUARD_END                                                        \
  }                                                                           \
  catch (const std::exception &e) {                                           \
    set_error(e.what());                                                      \
    return -2;                                                                \
  }

extern "C" {

int fs_case_create(int npoints, const double *x, const double *y, const double *z, int nfaces,
                   const int *face_ptr, const int *face_nodes, const int *owner, int nif,
                   const int *neighbour, const double *wall_u, const double *wall_v,
                   const double *wall_w, double rho, double mu, fs_case **out) {
  CASE_GUARD_BEGIN
  if (!out || !x || !y || !z || !face_ptr || !face_nodes || !owner || !neighbour || !wall_u ||
      !wall_v || !wall_w || npoints <= 0 || nfaces <= 0 || nif <= 0 || nif > nfaces) {
    set_error("fs_case_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  if (face_ptr[0] != 0) {
    set_error("fs_case_create: face_ptr[0] must be 0");
    return -1;
  }
  for (int f = 0; f < nfaces; ++f) {
    if (face_ptr[f + 1] - face_ptr[f] < 3 || owner[f] < 0 || (f < nif && neighbour[f] < 0)) {
      set_error("fs_case_create: bad face " + std::to_string(f));
      return -1;
    }
    for (int t = face_ptr[f]; t < face_ptr[f + 1]; ++t)
      if (face_nodes[t] < 0 || face_nodes[t] >= npoints) {
        set_error("fs_case_create: face node out of range at face " + std::to_string(f));
        return -1;
      }
  }
  auto *cs = new fs_case();
  fs::SimpleCase &C = cs->c;
  try {
    fs::PolyInput in;
    in.npoints = npoints;
    in.nfaces = nfaces;
    in.nif = nif;
    in.x = x;
    in.y = y;
    in.z = z;
    in.fptr = face_ptr;
    in.fnodes = face_nodes;
    in.owner = owner;
    in.neigh = neighbour;
    fs::build_topology(in, C.geom);
    fs::build_geometry(in, C.geom);
    C.N = C.geom.N;
    C.W = C.geom.W;
    C.F = C.geom.F;
    C.Fi = C.geom.Fi;
    C.Fb = C.F - C.Fi;
    C.h_own.assign(owner, owner + nfaces);
    C.h_nei.assign(neighbour, neighbour + nif);
    const int rc = fs::build_meshes(C);
    if (rc != 0) {
      std::string msg = fs_last_error();
      delete cs;
      set_error("fs_case_create: " + msg);
      return rc;
    }
    C.own.upload(owner, nfaces, C.s);
    C.nei.upload(neighbour, nif, C.s);
    fs::upload_geometry(C);
    C.bu.upload(wall_u, C.Fb, C.s);
    C.bv.upload(wall_v, C.Fb, C.s);
    C.bw.upload(wall_w, C.Fb, C.s);
    C.rho = rho;
    C.mu = mu;
    const size_t N = C.N, Fi = C.Fi, Fb = C.Fb, F = C.F;
    for (auto *b : {&C.u, &C.v, &C.w, &C.p, &C.pc, &C.du, &C.dv, &C.dw, &C.us, &C.vs, &C.ws, &C.ac,
                    &C.bc, &C.dphi}) {
      b->alloc(N);
      FS_HIP_THROW(hipMemsetAsync(b->p, 0, sizeof(double) * N, C.s));
    }
    for (auto *b : {&C.gu, &C.gv, &C.gw, &C.gp, &C.gpc}) {
      b->alloc(3 * N);
      FS_HIP_THROW(hipMemsetAsync(b->p, 0, sizeof(double) * 3 * N, C.s));
    }
    C.mdot.alloc(Fi);
    FS_HIP_THROW(hipMemsetAsync(C.mdot.p, 0, sizeof(double) * Fi, C.s));
    C.bp.alloc(Fb);
    FS_HIP_THROW(hipMemsetAsync(C.bp.p, 0, sizeof(double) * Fb, C.s));
    C.bgp.alloc(3 * Fb);
    FS_HIP_THROW(hipMemsetAsync(C.bgp.p, 0, sizeof(double) * 3 * Fb, C.s));
    C.fp.alloc(4 * F);
    C.fv.alloc(4 * F);
    C.cf.alloc(F);
    C.ff.alloc(F);
    C.tf.alloc(F);
    C.tf3.alloc(3 * F);
    FS_HIP_THROW(hipMemsetAsync(C.ff.p, 0, sizeof(double) * F, C.s));
    C.anb.alloc(N * C.W);
    C.part.alloc(fs::kPartials);
    FS_HIP_THROW(hipStreamSynchronize(C.s));
  } catch (...) {
    if (C.mom_mesh) fs_mesh_destroy(C.mom_mesh);
    if (C.mesh) fs_mesh_destroy(C.mesh);
    delete cs;
    throw;
  }
  *out = cs;
  return 0;
  CASE_GUARD_END
}

int fs_case_destroy(fs_case *cs) {
  if (!cs) return 0;
  fs_mesh *m = cs->c.mesh, *mm = cs->c.mom_mesh;
  if (m) (void)hipStreamSynchronize(m->stream);
  delete cs; // device buffers first, the meshes last, the stream's owner after the other
  if (mm) fs_mesh_destroy(mm);
  if (m) fs_mesh_destroy(m);
  return 0;
}

fs_mesh *fs_case_mesh(fs_case *cs) { return cs ? cs->c.mesh : nullptr; }

int fs_case_configure(fs_case *cs, double urf_u, double urf_p, int msolver, double mrrf,
                      int mitmax, int mlmg, int psolver, double prrf, int pitmax, int plmg) {
  if (!cs) return -1;
  auto &C = cs->c;
  if (!(urf_u > 0.0 && urf_u <= 1.0) || !(urf_p > 0.0 && urf_p <= 1.0) || mitmax < 0 || pitmax < 0) {
    set_error("fs_case_configure: invalid relaxation factors or iteration limits");
    return -1;
  }
  C.urf_u = urf_u;
  C.urf_p = urf_p;
  C.msolver = msolver;
  C.mrrf = mrrf;
  C.mitmax = mitmax;
  C.mlmg = mlmg;
  C.psolver = psolver;
  C.prrf = prrf;
  C.pitmax = pitmax;
  C.plmg = plmg;
  return 0;
}

int fs_case_iterate(fs_case *cs, int niter, fs_simple_stats *st) {
  if (!cs || niter < 0) return -1;
  CASE_GUARD_BEGIN
  return fs::simple_iterate(cs->c, niter, st);
  CASE_GUARD_END
}

int fs_case_phase(fs_case *cs, int phase) {
  if (!cs) return -1;
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  switch (phase) {
  case FS_PHASE_MOMENTUM_GRADIENTS: fs::phase_momentum_gradients(C); break;
  case FS_PHASE_ASSEMBLE_U: fs::phase_momentum_assemble(C, 0); break;
  case FS_PHASE_ASSEMBLE_V: fs::phase_momentum_assemble(C, 1); break;
  case FS_PHASE_ASSEMBLE_W: fs::phase_momentum_assemble(C, 2); break;
  case FS_PHASE_SOLVE_U: { int rc = fs::phase_momentum_solve(C, 0, nullptr); if (rc < 0) return rc; break; }
  case FS_PHASE_SOLVE_V: { int rc = fs::phase_momentum_solve(C, 1, nullptr); if (rc < 0) return rc; break; }
  case FS_PHASE_SOLVE_W: { int rc = fs::phase_momentum_solve(C, 2, nullptr); if (rc < 0) return rc; break; }
  case FS_PHASE_PC_ASSEMBLE: fs::phase_pc_assemble(C); break;
  case FS_PHASE_PC_SOLVE_CORRECT: { int rc = fs::phase_pc_solve_correct(C, nullptr); if (rc < 0) return rc; break; }
  case FS_PHASE_UPDATE_SYMMETRY: fs::update_symmetry(C); break;
  case FS_PHASE_TURBULENCE: {
    if (!C.tb.on) {
      set_error("fs_case_phase: the case has no turbulence model");
      return -1;
    }
    int rc = fs::phase_turbulence(C, nullptr);
    if (rc < 0) return rc;
    break;
  }
  default:
    set_error("fs_case_phase: unknown phase");
    return -1;
  }
  FS_HIP(hipStreamSynchronize(C.s));
  return 0;
  CASE_GUARD_END
}

// name -> (buffer, length)
static bool case_field(fs::SimpleCase &C, const char *name, double **p, size_t *n) {
  const size_t N = C.N, Fi = C.Fi, Fb = C.Fb, F = C.F;
  const std::map<std::string, std::pair<fs::DevBuf<double> *, size_t>> tab = {
      {"u", {&C.u, N}}, {"v", {&C.v, N}}, {"w", {&C.w, N}}, {"p", {&C.p, N}}, {"pc", {&C.pc, N}},
      {"mdot", {&C.mdot, Fi}}, {"gu", {&C.gu, 3 * N}}, {"gv", {&C.gv, 3 * N}},
      {"gw", {&C.gw, 3 * N}}, {"gp", {&C.gp, 3 * N}}, {"gpc", {&C.gpc, 3 * N}},
      {"bu", {&C.bu, Fb}}, {"bv", {&C.bv, Fb}}, {"bw", {&C.bw, Fb}}, {"bp", {&C.bp, Fb}},
      {"bgp", {&C.bgp, 3 * Fb}}, {"du", {&C.du, N}}, {"dv", {&C.dv, N}},
      {"dw", {&C.dw, N}}, {"du2", {&C.du2, C.simplec ? N : 0}}, {"dv2", {&C.dv2, C.simplec ? N : 0}},
      {"dw2", {&C.dw2, C.simplec ? N : 0}}, {"us", {&C.us, N}}, {"vs", {&C.vs, N}}, {"ws", {&C.ws, N}},
      {"ac", {&C.ac, N}}, {"anb", {&C.anb, N * C.W}}, {"bc", {&C.bc, N}}, {"cf", {&C.cf, F}},
      {"ff", {&C.ff, F}}, {"tf", {&C.tf, F}}, {"vol", {&C.vol, N}}, {"Sx", {&C.Sx, F}},
      {"Sy", {&C.Sy, F}}, {"Sz", {&C.Sz, F}}, {"g", {&C.g, F}}, {"gdiff", {&C.gdiff, F}},
      {"Tx", {&C.Tx, F}}, {"Ty", {&C.Ty, F}}, {"Tz", {&C.Tz, F}}, {"dcf", {&C.d, F}},
      {"A", {&C.A, F}}, {"kx", {&C.kx, Fi}}, {"ky", {&C.ky, Fi}}, {"kz", {&C.kz, Fi}},
      {"bmdot", {&C.bmdot, C.periodic() ? Fb : 0}}, {"beta", {&C.per.scal, C.periodic() ? 1 : 0}},
      {"k", {&C.tb.k, C.tb.on ? N : 0}}, {"omega", {&C.tb.om, C.tb.on ? N : 0}},
      {"mut", {&C.tb.mut, C.tb.on ? N : 0}}, {"wd", {&C.tb.wd, C.tb.on ? N : 0}},
      {"bk", {&C.tb.bk, C.tb.on ? Fb : 0}}, {"bomega", {&C.tb.bom, C.tb.on ? Fb : 0}},
      {"bmut", {&C.tb.bmut, C.tb.on ? Fb : 0}}, {"F1", {&C.tb.F1, C.tb.on ? N : 0}},
      {"F2", {&C.tb.F2, C.tb.on ? N : 0}}, {"grk", {&C.tb.grk, C.tb.on ? 3 * N : 0}}};
  auto it = tab.find(name);
  if (it == tab.end()) return false;
  *p = it->second.first->p;
  *n = it->second.second;
  return true;
}

int fs_case_get(fs_case *cs, const char *name, double *out, long long count) {
  if (!cs || !name || !out) return -1;
  CASE_GUARD_BEGIN
  double *p;
  size_t n;
  if (!case_field(cs->c, name, &p, &n)) {
    set_error(std::string("fs_case_get: unknown field ") + name);
    return -1;
  }
  if ((size_t)count != n) {
    set_error(std::string("fs_case_get: field ") + name + " has " + std::to_string(n) + " entries");
    return -1;
  }
  FS_HIP(hipMemcpyAsync(out, p, sizeof(double) * n, hipMemcpyDeviceToHost, cs->c.s));
  FS_HIP(hipStreamSynchronize(cs->c.s));
  const std::string nm(name);
  auto &C = cs->c;
  if (C.sys_perm && (nm == "ac" || nm == "bc" || nm == "anb")) {
    // a momentum system is held in colour-major order: return cell order
    std::vector<double> tmp(out, out + n);
    const size_t N = C.N, blocks_ = n / N;
    for (size_t j = 0; j < blocks_; ++j)
      for (size_t c = 0; c < N; ++c) out[j * N + c] = tmp[j * N + C.h_mperm[c]];
  }
  return 0;
  CASE_GUARD_END
}

int fs_case_set(fs_case *cs, const char *name, const double *in, long long count) {
  if (!cs || !name || !in) return -1;
  CASE_GUARD_BEGIN
  double *p;
  size_t n;
  if (!case_field(cs->c, name, &p, &n)) {
    set_error(std::string("fs_case_set: unknown field ") + name);
    return -1;
  }
  if ((size_t)count != n) {
    set_error(std::string("fs_case_set: field ") + name + " has " + std::to_string(n) + " entries");
    return -1;
  }
  FS_HIP(hipMemcpyAsync(p, in, sizeof(double) * n, hipMemcpyHostToDevice, cs->c.s));
  FS_HIP(hipStreamSynchronize(cs->c.s));
  const std::string nm(name);
  if (nm == "u" || nm == "v" || nm == "w") cs->c.sym_dirty = cs->c.btype.p != nullptr;
  return 0;
  CASE_GUARD_END
}

int fs_case_set_convection(fs_case *cs, int framework, int scheme, double bleed, int start_iter) {
  if (!cs) {
    set_error("fs_case_set_convection: null case");
    return -1;
  }
  if (!fs::hr_scheme_defined(framework, scheme)) {
    set_error("fs_case_set_convection: scheme " + std::to_string(scheme) + " is not defined in framework " +
              std::to_string(framework));
    return -1;
  }
  auto &C = cs->c;
  C.hr_fw = framework;
  C.hr_scheme = scheme;
  C.bleed = bleed;
  C.hr_start = start_iter;
  return 0;
}

int fs_case_set_gradient_method(fs_case *cs, int momentum_method, int momentum_iters, int continuity_method,
                                int continuity_iters) {
  if (!cs || momentum_method < 1 || momentum_method > 2 || continuity_method < 1 || continuity_method > 2 ||
      momentum_iters < 0 || continuity_iters < 0) {
    set_error("fs_case_set_gradient_method: methods are 1 or 2, iterations >= 0");
    return -1;
  }
  auto &C = cs->c;
  C.ngrad_mom = momentum_method == 1 ? 0 : momentum_iters;
  C.ngrad_cont = continuity_method == 1 ? 0 : continuity_iters;
  return 0;
}

int fs_case_set_restart_capture(fs_case *cs, int on) {
  if (!cs) {
    set_error("fs_case_set_restart_capture: null case");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  C.capture = on != 0;
  C.captured_iter = -1;
  if (C.capture) {
    C.cap_fgm.alloc((size_t)9 * C.Fi);
    C.cap_fgp.alloc((size_t)3 * C.Fi);
    C.cap_eff.alloc(C.N);
  }
  return 0;
  CASE_GUARD_END
}

namespace {
// A Fortran unformatted sequential file: each record framed by its byte
// length as a 4-byte integer before and after.
struct RecordWriter {
  std::vector<char> out;
  void rec(const std::vector<const std::vector<double> *> &parts) {
    size_t n = 0;
    for (auto *p : parts) n += p->size() * sizeof(double);
    if (n > 0x7fffffffULL) throw std::runtime_error("restart record above 2 GiB");
    const int32_t len = (int32_t)n;
    const char *lp = reinterpret_cast<const char *>(&len);
    out.insert(out.end(), lp, lp + 4);
    for (auto *p : parts) {
      const char *d = reinterpret_cast<const char *>(p->data());
      out.insert(out.end(), d, d + p->size() * sizeof(double));
    }
    out.insert(out.end(), lp, lp + 4);
  }
};
} // namespace

int fs_case_write_solution(fs_case *cs, const char *path, int npatches, const int *patch_nfaces) {
  if (!cs || !path || npatches <= 0 || !patch_nfaces) {
    set_error("fs_case_write_solution: invalid arguments");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  if (!C.capture || C.captured_iter != C.iter || C.iter < 1) {
    set_error("fs_case_write_solution: the last iteration must run with fs_case_set_restart_capture on");
    return -1;
  }
  const size_t N = C.N, Fi = C.Fi, Fb = C.Fb;
  long long tot = 0;
  int nbfmax = 0;
  for (int q = 0; q < npatches; ++q) {
    if (patch_nfaces[q] < 0) {
      set_error("fs_case_write_solution: negative patch size");
      return -1;
    }
    tot += patch_nfaces[q];
    nbfmax = std::max(nbfmax, patch_nfaces[q]);
  }
  if (tot != (long long)Fb) {
    set_error("fs_case_write_solution: patches hold " + std::to_string(tot) + " faces, the boundary " +
              std::to_string(Fb));
    return -1;
  }
  auto get = [&](const fs::DevBuf<double> &b, size_t n) {
    std::vector<double> h(n);
    if (n) FS_HIP_THROW(hipMemcpyAsync(h.data(), b.p, sizeof(double) * n, hipMemcpyDeviceToHost, C.s));
    return h;
  };
  // the constant-density gradient of the start-up's CalculateFaceDensity
  // (ComputeFaceDensity.for: Gradient with MethodCalcGradientDensity = 2,
  // nIterGradientDensity = 2, no limiter)
  fs::DevBuf<double> rc, rb, rg;
  rc.upload(std::vector<double>(N, C.rho), C.s);
  rb.upload(std::vector<double>(Fb, C.rho), C.s);
  rg.alloc(3 * N);
  {
    fs::GradArgs a{};
    a.phi[0] = rc.p;
    a.bval[0] = rb.p;
    a.grad[0] = rg.p;
    fs::gradients<1>(C, a, 2);
  }
  const auto u = get(C.u, N), v = get(C.v, N), w = get(C.w, N), p = get(C.p, N), mdot = get(C.mdot, Fi);
  const auto gu = get(C.gu, 3 * N), gv = get(C.gv, 3 * N), gw = get(C.gw, 3 * N), gp = get(C.gp, 3 * N);
  const auto bu = get(C.bu, Fb), bv = get(C.bv, Fb), bw = get(C.bw, Fb), bp = get(C.bp, Fb), bgp = get(C.bgp, 3 * Fb);
  const auto du = get(C.du, N), dv = get(C.dv, N), dw = get(C.dw, N);
  const auto us = get(C.us, N), vs = get(C.vs, N), ws = get(C.ws, N);
  const auto du2 = C.simplec ? get(C.du2, N) : du, dv2 = C.simplec ? get(C.dv2, N) : dv,
             dw2 = C.simplec ? get(C.dw2, N) : dw;
  const auto fgm = get(C.cap_fgm, 9 * Fi), fgp = get(C.cap_fgp, 3 * Fi), eff = get(C.cap_eff, N);
  const auto dg = get(rg, 3 * N);
  FS_HIP(hipStreamSynchronize(C.s));
  const size_t nbc = npatches, nbm = nbfmax, P = nbc * nbm;
  // boundary face b of patch q at record j: (q, j) of a (NumberOfBCSets,
  // NBFacesMax) array, column-major; unused entries 0
  std::vector<int> bq(Fb), bj(Fb);
  {
    size_t b = 0;
    for (int q = 0; q < npatches; ++q)
      for (int j = 0; j < patch_nfaces[q]; ++j, ++b) {
        bq[b] = q;
        bj[b] = j;
      }
  }
  auto pad = [&](auto value) {
    std::vector<double> a(P, 0.0);
    for (size_t b = 0; b < Fb; ++b) a[bq[b] + (size_t)bj[b] * nbc] = value(b);
    return a;
  };
  auto slice = [](const std::vector<double> &a, size_t off, size_t n) {
    return std::vector<double>(a.begin() + off, a.begin() + off + n);
  };
  const std::vector<int> &own = C.h_own;
  auto owner_copy = [&](const std::vector<double> &g, size_t d) {
    return pad([&](size_t b) { return g[d * N + own[Fi + b]]; });
  };
  const std::vector<double> zN(N, 0.0), zP(P, 0.0), z3nbc(3 * nbc, 0.0), muN(N, C.mu), rhoN(N, C.rho),
      rhoFi(Fi, C.rho);
  // BViscosity holds mu on the boundary faces only; BDensity is
  // ConstantDensity over the whole array, padding included (it is filled
  // whole at allocation, AllocateStorageRoutine.for:1929-1930)
  const auto muP = pad([&](size_t) { return C.mu; });
  const std::vector<double> rhoP(P, C.rho);
  RecordWriter R;
  // LSolveMomentum (WriteSolutionToContinue.for:24-46)
  R.rec({&u, &v, &w});
  R.rec({&gu});
  R.rec({&gv});
  R.rec({&gw});
  const auto pu = pad([&](size_t b) { return bu[b]; }), pv = pad([&](size_t b) { return bv[b]; }),
             pw = pad([&](size_t b) { return bw[b]; });
  R.rec({&pu, &pv, &pw});
  for (const auto *g : {&gu, &gv, &gw}) {
    const auto x = owner_copy(*g, 0), y = owner_copy(*g, 1), z = owner_copy(*g, 2);
    R.rec({&x, &y, &z});
  }
  for (int k = 0; k < 3; ++k) {
    const auto f = slice(fgm, (size_t)3 * k * Fi, 3 * Fi);
    R.rec({&f});
  }
  R.rec({&zP, &zP, &zP});           // x/y/zVeldirection
  R.rec({&muN});                    // Viscosity
  R.rec({&muP});                    // BViscosity
  for (int k = 0; k < 3; ++k) R.rec({&zN, &zN}); // ScMomentum, SbMomentum
  R.rec({&mdot, &zP, &eff});        // mdot, Bmdot (walls and symmetry planes carry none), effdiv
  R.rec({&du, &du2, &us});
  R.rec({&dv, &dv2, &vs});
  R.rec({&dw, &dw2, &ws});
  R.rec({&zN, &zP});                // MachNumber, BMachNumber
  // LSolveContinuity (:61-66)
  R.rec({&p});
  const auto pp = pad([&](size_t b) { return bp[b]; });
  R.rec({&pp});
  R.rec({&gp});
  {
    const auto x = pad([&](size_t b) { return bgp[b]; }), y = pad([&](size_t b) { return bgp[Fb + b]; }),
               z = pad([&](size_t b) { return bgp[2 * Fb + b]; });
    R.rec({&x, &y, &z});
  }
  R.rec({&fgp});
  // density (:173-175), effective diffusion coefficient (:207-208)
  {
    const auto gx = slice(dg, 0, N), gy = slice(dg, N, N), gz = slice(dg, 2 * N, N);
    R.rec({&rhoN, &gx, &gy, &gz});
    const auto bx = owner_copy(dg, 0), by = owner_copy(dg, 1), bz = owner_copy(dg, 2);
    R.rec({&rhoP, &bx, &by, &bz});
  }
  R.rec({&rhoFi});                  // Densityf
  R.rec({&muN});                    // eDiffCoefficient
  R.rec({&muP});                    // BeDiffCoefficient
  R.rec({&z3nbc});                  // MassFlowRate, Old, OldOld
  R.rec({&z3nbc});                  // OutletPressure, Old, OldOld
  FILE *fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("fs_case_write_solution: cannot open ") + path);
    return -1;
  }
  const size_t wr = std::fwrite(R.out.data(), 1, R.out.size(), fp);
  const int cl = std::fclose(fp);
  if (wr != R.out.size() || cl != 0) {
    set_error(std::string("fs_case_write_solution: write failed for ") + path);
    return -1;
  }
  return 0;
  CASE_GUARD_END
}

int fs_case_read_solution(fs_case *cs, const char *path, int npatches, const int *patch_nfaces) {
  if (!cs || !path || npatches <= 0 || !patch_nfaces) {
    set_error("fs_case_read_solution: invalid arguments");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  const size_t N = C.N, Fi = C.Fi, Fb = C.Fb;
  long long tot = 0;
  int nbfmax = 0;
  for (int q = 0; q < npatches; ++q) {
    if (patch_nfaces[q] < 0) {
      set_error("fs_case_read_solution: negative patch size");
      return -1;
    }
    tot += patch_nfaces[q];
    nbfmax = std::max(nbfmax, patch_nfaces[q]);
  }
  if (tot != (long long)Fb) {
    set_error("fs_case_read_solution: patches do not cover the boundary");
    return -1;
  }
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    set_error(std::string("fs_case_read_solution: cannot open ") + path);
    return -1;
  }
  // the whole file at once (about 40 N + 13 Fi doubles for a steady laminar flow)
  std::vector<char> blob;
  {
    long len = -1;
    if (std::fseek(fp, 0, SEEK_END) == 0) len = std::ftell(fp);
    if (len < 0 || std::fseek(fp, 0, SEEK_SET) != 0) {
      std::fclose(fp);
      set_error(std::string("fs_case_read_solution: cannot size ") + path);
      return -1;
    }
    blob.resize((size_t)len);
    const size_t got = len ? std::fread(blob.data(), 1, (size_t)len, fp) : 0;
    std::fclose(fp);
    if (got != (size_t)len) {
      set_error(std::string("fs_case_read_solution: short read of ") + path);
      return -1;
    }
  }
  // the records, each framed by its 4-byte length
  std::vector<std::pair<size_t, size_t>> recs; // offset, bytes
  for (size_t o = 0; o < blob.size();) {
    int32_t n0, n1;
    if (o + 4 > blob.size()) {
      set_error("fs_case_read_solution: truncated record marker");
      return -1;
    }
    std::memcpy(&n0, &blob[o], 4);
    if (n0 < 0 || o + 8 + (size_t)n0 > blob.size()) {
      set_error("fs_case_read_solution: truncated record");
      return -1;
    }
    std::memcpy(&n1, &blob[o + 4 + n0], 4);
    if (n1 != n0) {
      set_error("fs_case_read_solution: record markers disagree");
      return -1;
    }
    recs.push_back({o + 4, (size_t)n0});
    o += 8 + (size_t)n0;
  }
  const size_t P = (size_t)npatches * nbfmax;
  // the laminar steady layout of fs_case_write_solution
  const size_t expect[] = {3 * N, 3 * N, 3 * N, 3 * N, 3 * P, 3 * P, 3 * P, 3 * P, 3 * Fi, 3 * Fi, 3 * Fi, 3 * P,
                           N, P, 2 * N, 2 * N, 2 * N, Fi + P + N, 3 * N, 3 * N, 3 * N, N + P, N, P, 3 * N,
                           3 * P, 3 * Fi, 4 * N, 4 * P, Fi, N, P, 3 * (size_t)npatches, 3 * (size_t)npatches};
  const size_t nrec = sizeof(expect) / sizeof(expect[0]);
  if (recs.size() != nrec) {
    set_error("fs_case_read_solution: " + std::to_string(recs.size()) + " records, a steady laminar flow has " +
              std::to_string(nrec));
    return -1;
  }
  for (size_t k = 0; k < nrec; ++k)
    if (recs[k].second != expect[k] * sizeof(double)) {
      set_error("fs_case_read_solution: record " + std::to_string(k) + " does not fit this mesh and patches");
      return -1;
    }
  auto at = [&](size_t k, size_t off) { return reinterpret_cast<const double *>(&blob[recs[k].first]) + off; };
  std::vector<int> bq(Fb), bj(Fb);
  {
    size_t b = 0;
    for (int q = 0; q < npatches; ++q)
      for (int j = 0; j < patch_nfaces[q]; ++j, ++b) {
        bq[b] = q;
        bj[b] = j;
      }
  }
  auto unpad = [&](const double *a) {
    std::vector<double> v(Fb);
    for (size_t b = 0; b < Fb; ++b) v[b] = a[bq[b] + (size_t)bj[b] * npatches];
    return v;
  };
  auto put = [&](fs::DevBuf<double> &d, const double *h, size_t n) {
    FS_HIP_THROW(hipMemcpyAsync(d.p, h, sizeof(double) * n, hipMemcpyHostToDevice, C.s));
  };
  put(C.u, at(0, 0), N);
  put(C.v, at(0, N), N);
  put(C.w, at(0, 2 * N), N);
  const auto bu = unpad(at(4, 0)), bv = unpad(at(4, P)), bw = unpad(at(4, 2 * P));
  put(C.bu, bu.data(), Fb);
  put(C.bv, bv.data(), Fb);
  put(C.bw, bw.data(), Fb);
  put(C.mdot, at(17, 0), Fi);
  put(C.p, at(22, 0), N);
  const auto bp = unpad(at(23, 0));
  put(C.bp, bp.data(), Fb);
  FS_HIP(hipStreamSynchronize(C.s));
  C.sym_dirty = false; // the boundary velocities come from the file
  return 0;
  CASE_GUARD_END
}

int fs_case_set_gradient_interpolation(fs_case *cs, int momentum, int continuity) {
  if (!cs || momentum < FS_GRADINT_AVERAGE || momentum > FS_GRADINT_HARMONIC || continuity < FS_GRADINT_AVERAGE ||
      continuity > FS_GRADINT_HARMONIC) {
    set_error("fs_case_set_gradient_interpolation: schemes are FS_GRADINT_*");
    return -1;
  }
  cs->c.gi_mom = momentum;
  cs->c.gi_cont = continuity;
  return 0;
}

int fs_case_set_decompose(fs_case *cs, int method) {
  if (!cs || method < 1 || method > 3) {
    set_error("fs_case_set_decompose: MethodDecomposeS is 1, 2 or 3");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  fs::decompose_faces(C.geom, C.h_own.data(), method);
  const auto &F = C.geom.face;
  C.gdiff.upload(F.gdiff, C.s);
  C.Tx.upload(F.Tx, C.s);
  C.Ty.upload(F.Ty, C.s);
  C.Tz.upload(F.Tz, C.s);
  C.kx.upload(F.kx, C.s);
  C.ky.upload(F.ky, C.s);
  C.kz.upload(F.kz, C.s);
  C.decomp = method;
  if (C.periodic()) fs::periodic_upload_geometry(C); // the periodic faces' E/T split follows
  FS_HIP(hipStreamSynchronize(C.s));
  return 0;
  CASE_GUARD_END
}

int fs_case_set_algorithm(fs_case *cs, int algorithm) {
  if (!cs || (algorithm != FS_ALGORITHM_SIMPLE && algorithm != FS_ALGORITHM_SIMPLEC)) {
    set_error("fs_case_set_algorithm: algorithm is FS_ALGORITHM_SIMPLE or FS_ALGORITHM_SIMPLEC");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  C.simplec = algorithm == FS_ALGORITHM_SIMPLEC;
  if (C.simplec && !C.du2.p) {
    for (auto *b : {&C.du2, &C.dv2, &C.dw2}) {
      b->alloc(C.N);
      FS_HIP(hipMemsetAsync(b->p, 0, sizeof(double) * C.N, C.s));
    }
    FS_HIP(hipStreamSynchronize(C.s));
  }
  return 0;
  CASE_GUARD_END
}

int fs_case_set_gradient_limiters(fs_case *cs, int momentum_method, int continuity_method) {
  if (!cs || momentum_method < 0 || momentum_method > 2 || continuity_method < 0 || continuity_method > 2) {
    set_error("fs_case_set_gradient_limiters: methods are 0 (off), 1 or 2");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  if (momentum_method || continuity_method) fs::upload_centroids(C);
  C.lim_mom = momentum_method;
  C.lim_cont = continuity_method;
  return 0;
  CASE_GUARD_END
}

int fs_case_set_patch_types(fs_case *cs, int npatches, const int *patch_nfaces, const int *patch_type) {
  if (!cs || npatches <= 0 || !patch_nfaces || !patch_type) {
    set_error("fs_case_set_patch_types: invalid arguments");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  std::vector<int> type(C.Fb), zq(C.Fb);
  long long start = 0;
  bool any = false;
  for (int q = 0; q < npatches; ++q) {
    const int n = patch_nfaces[q], t = patch_type[q];
    if (n < 0 || (t != FS_BC_WALL && t != FS_BC_SYMMETRY && t != FS_BC_PERIODIC) || start + n > C.Fb) {
      set_error("fs_case_set_patch_types: patch " + std::to_string(q) + " has a bad size or type");
      return -1;
    }
    if (t == FS_BC_SYMMETRY && n < q + 1) {
      set_error("fs_case_set_patch_types: symmetry patch " + std::to_string(q + 1) + " has fewer than " +
                std::to_string(q + 1) + " faces");
      return -1;
    }
    for (int j = 0; j < n; ++j) {
      type[start + j] = t;
      zq[start + j] = (int)start + q; // face i of patch i
    }
    any = any || t != FS_BC_WALL;
    start += n;
  }
  if (start != C.Fb) {
    set_error("fs_case_set_patch_types: patches hold " + std::to_string(start) + " faces, the boundary " +
              std::to_string(C.Fb));
    return -1;
  }
  if (!any) {
    C.btype.release();
    C.bzq.release();
    C.sym_dirty = false;
    return 0;
  }
  C.btype.upload(type, C.s);
  C.bzq.upload(zq, C.s);
  FS_HIP(hipStreamSynchronize(C.s));
  C.sym_dirty = true;
  return 0;
  CASE_GUARD_END
}

int fs_case_set_periodic(fs_case *cs, int npatches, const int *patch_nfaces, const int *pair,
                         const double *translation, double beta, int calc_beta, double mdot_target, double relax,
                         int iters, int every) {
  if (!cs || npatches <= 0 || !patch_nfaces || !pair || !translation || iters < 1) {
    set_error("fs_case_set_periodic: invalid arguments");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  if (C.periodic()) {
    set_error("fs_case_set_periodic: the case already has its periodic pairs");
    return -1;
  }
  if (!C.btype.p) {
    set_error("fs_case_set_periodic: mark the periodic patches with fs_case_set_patch_types first");
    return -1;
  }
  std::vector<int> type(C.Fb);
  FS_HIP(hipMemcpy(type.data(), C.btype.p, sizeof(int) * C.Fb, hipMemcpyDeviceToHost));
  std::vector<int> start(npatches + 1, 0);
  for (int q = 0; q < npatches; ++q) start[q + 1] = start[q] + patch_nfaces[q];
  if (start[npatches] != C.Fb) {
    set_error("fs_case_set_periodic: patches do not cover the boundary");
    return -1;
  }
  for (int q = 0; q < npatches; ++q) {
    const int n = patch_nfaces[q];
    if (n == 0) continue;
    const bool per = type[start[q]] == FS_BC_PERIODIC;
    for (int j = 0; j < n; ++j)
      if ((type[start[q] + j] == FS_BC_PERIODIC) != per) {
        set_error("fs_case_set_periodic: patch " + std::to_string(q) + " mixes periodic and other faces");
        return -1;
      }
    if (per != (pair[q] >= 0) ||
        (pair[q] >= 0 && (pair[q] >= npatches || pair[q] == q || pair[pair[q]] != q ||
                          patch_nfaces[pair[q]] != n))) {
      set_error("fs_case_set_periodic: patch " + std::to_string(q) + " has no mutual periodic partner of its size");
      return -1;
    }
  }
  const fs::FaceGeom &G = C.geom.face;
  const int Fi = C.Fi;
  // RotationMatrixCoefficients, translational (CalculateRotationMatrixCoefficients
  // .for:171-242): the partner face whose centroid offset is closest to the
  // translation (largest per-axis miss, first minimum)
  std::vector<int> corr(C.Fb, -1), act(C.Fb, 0);
  for (int q1 = 0; q1 < npatches; ++q1) {
    const int q2 = pair[q1];
    if (q2 <= q1) continue;
    for (int f1 = Fi + start[q1]; f1 < Fi + start[q1 + 1]; ++f1) {
      double best = 1.e30;
      int fd = -1;
      for (int f2 = Fi + start[q2]; f2 < Fi + start[q2 + 1]; ++f2) {
        const double d1 = fabs(fabs(G.cx[f2] - G.cx[f1]) - fabs(translation[3 * q1]));
        const double d2 = fabs(fabs(G.cy[f2] - G.cy[f1]) - fabs(translation[3 * q1 + 1]));
        const double d3 = fabs(fabs(G.cz[f2] - G.cz[f1]) - fabs(translation[3 * q1 + 2]));
        double dif = d1;
        if (d2 > dif) dif = d2;
        if (d3 > dif) dif = d3;
        if (dif < best) {
          best = dif;
          fd = f2;
        }
      }
      if (fd < 0 || corr[fd - Fi] >= 0) {
        set_error("fs_case_set_periodic: a periodic face finds no partner of its own");
        return -1;
      }
      corr[f1 - Fi] = fd - Fi;
      corr[fd - Fi] = f1 - Fi;
      act[f1 - Fi] = 1;
    }
  }
  fs::Periodic &P = C.per;
  P.h_pair.assign(pair, pair + npatches);
  P.h_trans.assign(translation, translation + 3 * npatches);
  P.h_b.clear();
  P.h_patch.clear();
  std::vector<int> kof(C.Fb, -1);
  for (int q = 0; q < npatches; ++q)
    if (pair[q] >= 0)
      for (int b = start[q]; b < start[q + 1]; ++b) {
        kof[b] = (int)P.h_b.size();
        P.h_b.push_back(b);
        P.h_patch.push_back(q);
      }
  P.np = (int)P.h_b.size();
  P.h_part.assign(P.np, 0);
  P.h_j1.assign(P.np, 0);
  P.h_act.clear();
  for (int k = 0; k < P.np; ++k) {
    const int b = P.h_b[k];
    P.h_part[k] = kof[corr[b]];
    P.h_j1[k] = C.h_own[Fi + corr[b]];
    if (act[b]) P.h_act.push_back(k);
  }
  P.na = (int)P.h_act.size();
  for (int q = 0; q < npatches; ++q)
    if (pair[q] >= 0) {
      for (int d = 0; d < 3; ++d) P.tuse[d] = fabs(translation[3 * q + d]);
      break;
    }
  // LPeriodicImplicit: the partner cell in the face's slot of ElementNeighbor,
  // the slot searched over NumbOfElementNodes entries (:249-289)
  std::vector<int> enbr = C.geom.enbr;
  for (int k = 0; k < P.np; ++k) {
    const int f = Fi + P.h_b[k], i1 = C.h_own[f];
    const int nen = C.geom.cell_node_ptr[i1 + 1] - C.geom.cell_node_ptr[i1];
    bool found = false;
    for (int sl = 0; sl < nen && sl < C.W; ++sl)
      if (C.geom.gef[(size_t)sl * C.N + i1] == f + 1) {
        enbr[(size_t)sl * C.N + i1] = P.h_j1[k] + 1;
        found = true;
        break;
      }
    if (!found) {
      set_error("fs_case_set_periodic: a periodic face is in no slot of its cell");
      return -1;
    }
  }
  C.geom.enbr.swap(enbr);
  int rc = fs::build_meshes(C);
  if (rc != 0) {
    C.geom.enbr.swap(enbr);
    P = fs::Periodic();
    return rc;
  }
  P.b.upload(P.h_b, C.s);
  P.part.upload(P.h_part, C.s);
  P.j1.upload(P.h_j1, C.s);
  P.act.upload(P.h_act.empty() ? std::vector<int>{0} : P.h_act, C.s);
  fs::periodic_upload_geometry(C);
  std::vector<double> scal(8, 0.0);
  scal[fs::PSCAL_BETA] = beta;
  P.scal.upload(scal, C.s);
  P.calc_beta = calc_beta != 0;
  P.mdot = mdot_target;
  P.relax = relax;
  P.iters = iters;
  P.every = every > 0 ? every : 1;
  std::vector<int> pk(C.Fb, -1);
  for (int k = 0; k < P.np; ++k) pk[P.h_b[k]] = k;
  C.pk.upload(pk, C.s);
  const std::vector<double> zb(3 * (size_t)C.Fb, 0.0);
  C.bmdot.upload(zb.data(), C.Fb, C.s);
  C.bpc.upload(zb.data(), C.Fb, C.s);
  C.bgu.upload(zb, C.s);
  C.bgv.upload(zb, C.s);
  C.bgw.upload(zb, C.s);
  FS_HIP(hipStreamSynchronize(C.s));
  return 0;
  CASE_GUARD_END
}

int fs_case_wall_distance(fs_case *cs, double tol, int maxouter, double *wd, int *passes) {
  if (!cs || !(tol > 0.0) || maxouter < 1) {
    set_error("fs_case_wall_distance: NULL case, non-positive tolerance or no passes");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  int n = 0;
  const int rc = fs::wall_distance(C, tol, maxouter, &n);
  if (passes) *passes = n;
  if (rc < 0) return rc;
  if (wd) {
    FS_HIP(hipMemcpyAsync(wd, C.tb.wd.p, sizeof(double) * C.N, hipMemcpyDeviceToHost, C.s));
    FS_HIP(hipStreamSynchronize(C.s));
  }
  return rc;
  CASE_GUARD_END
}

int fs_case_set_sst(fs_case *cs, const int *io, const double *d, const double *k, const double *omega,
                    const double *mut, const double *wd, const double *bk, const double *bomega,
                    const double *bmut) {
  if (!cs || !io || !d || !k || !omega || !mut || !bk || !bomega || !bmut) {
    set_error("fs_case_set_sst: NULL argument");
    return -1;
  }
  if (!wd && cs->c.tb.wd.n != (size_t)cs->c.N) {
    set_error("fs_case_set_sst: no wall distance given and none computed (fs_case_wall_distance)");
    return -1;
  }
  for (int q = 0; q < 2; ++q) {
    const int sv = io[2 * q], fw = io[4 + 2 * q], sc = io[5 + 2 * q];
    if (sv < FS_SOR || sv > FS_CG || sv == FS_DIRECT) {
      set_error("fs_case_set_sst: k and omega take SOR, ILU, PBCG, GMRES, BICGS or CG");
      return -3;
    }
    if (io[1 + 2 * q] < 0 || fw < FS_HR_NONE || fw > FS_HR_NVF || sc < FS_CONV_UPWIND || sc > FS_CONV_GAMMA ||
        io[8 + 2 * q] < 1 || io[8 + 2 * q] > 2 || io[9 + 2 * q] < 0 || io[12 + q] < 0 || io[12 + q] > 2) {
      set_error("fs_case_set_sst: invalid option");
      return -1;
    }
  }
  if (!(d[0] > 0.0 && d[0] <= 1.0) || !(d[1] > 0.0 && d[1] <= 1.0) || !(d[2] > 0.0 && d[2] <= 1.0) ||
      !(d[5] > 0.0)) {
    set_error("fs_case_set_sst: relaxation factors must lie in (0, 1] and the viscosity cap be positive");
    return -1;
  }
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  auto &T = C.tb;
  const size_t N = C.N, Fb = C.Fb, Fi = C.Fi;
  for (size_t i = 0; i < N; ++i)
    if ((wd && !(wd[i] > 0.0)) || !std::isfinite(k[i]) || !std::isfinite(omega[i]) || !std::isfinite(mut[i])) {
      set_error("fs_case_set_sst: the wall distance must be positive and k, omega, mu_t finite");
      return -1;
    }
  T.ksolver = io[0]; T.kitmax = io[1]; T.wsolver = io[2]; T.witmax = io[3];
  T.kfw = io[4]; T.kscheme = io[5]; T.wfw = io[6]; T.wscheme = io[7];
  T.kngrad = io[8] == 1 ? 0 : io[9];
  T.wngrad = io[10] == 1 ? 0 : io[11];
  T.klim = io[12]; T.wlim = io[13];
  if (T.klim || T.wlim) fs::upload_centroids(C);
  T.urf_k = d[0]; T.urf_w = d[1]; T.urf_mut = d[2]; T.krrf = d[3]; T.wrrf = d[4]; T.max_ratio = d[5];
  T.k.upload(k, N, C.s);
  T.om.upload(omega, N, C.s);
  T.mut.upload(mut, N, C.s);
  if (wd) T.wd.upload(wd, N, C.s);
  T.bk.upload(bk, Fb, C.s);
  T.bom.upload(bomega, Fb, C.s);
  T.bmut.upload(bmut, Fb, C.s);
  // F1 and F2 start at 0 (AllocateStorageRoutine.for:2618-2631)
  const std::vector<double> zn(3 * N, 0.0), zb(3 * Fb, 0.0);
  T.F1.upload(zn.data(), N, C.s);
  T.F2.upload(zn.data(), N, C.s);
  T.bF1.upload(zb.data(), Fb, C.s);
  T.bF2.upload(zb.data(), Fb, C.s);
  for (auto *b : {&T.gk, &T.gom, &T.grk}) b->upload(zn, C.s);
  for (auto *b : {&T.bgk, &T.bgom, &T.bgrk}) b->upload(zb, C.s);
  for (auto *b : {&T.rhok, &T.prod, &T.sr, &T.med}) b->upload(zn.data(), N, C.s);
  for (auto *b : {&T.brhok, &T.bsr, &T.bmed}) b->upload(zb.data(), Fb, C.s);
  if (C.bmdot.n != Fb) C.bmdot.upload(zb.data(), Fb, C.s); // zero off the periodic faces
  (void)Fi;
  T.on = true;
  FS_HIP(hipStreamSynchronize(C.s));
  return 0;
  CASE_GUARD_END
}

int fs_case_time_spmv(fs_case *cs, int reps, double *avg_ms) {
  if (!cs || reps <= 0 || !avg_ms) return -1;
  CASE_GUARD_BEGIN
  auto &C = cs->c;
  // the p' operator of the last iteration applied to the current pressure,
  // dotted with its right-hand side (three distinct N-vectors, as in BiCGSTAB)
  return fs::time_spmv(C.mesh, C.ac.p, C.anb.p, C.p.p, C.bc.p, C.dphi.p, reps, avg_ms);
  CASE_GUARD_END
}

int fs_case_set_warm_start(fs_case *cs, fs_mpnn *net) {
  if (!cs) {
    set_error("fs_case_set_warm_start: case is NULL");
    return -1;
  }
  if (net && !fs::mpnn_matches(net, cs->c.N, cs->c.W, cs->c.mesh->h_nbr_dev)) {
    set_error("fs_case_set_warm_start: the network was built on another mesh topology");
    return -1;
  }
  cs->c.warm = net;
  return 0;
}

int fs_case_set_snapshot_sink(fs_case *cs, fs_mpnn *net) {
  if (!cs) {
    set_error("fs_case_set_snapshot_sink: case is NULL");
    return -1;
  }
  if (net && !fs::mpnn_matches(net, cs->c.N, cs->c.W, cs->c.mesh->h_nbr_dev)) {
    set_error("fs_case_set_snapshot_sink: the network was built on another mesh topology");
    return -1;
  }
  cs->c.sink = net;
  return 0;
}

int fs_case_dims(fs_case *cs, int *dims) {
  if (!cs || !dims) return -1;
  dims[0] = cs->c.N;
  dims[1] = cs->c.W;
  dims[2] = cs->c.F;
  dims[3] = cs->c.Fi;
  return 0;
}

// Host copies of the topology arrays the reference builds in Readit
// (reference numbering).  which: 0 NumbOfElementFaces(N), 1 NGlobalEFaces(N,W),
// 2 ElementNeighbor(N,W), 3 iOwnerNeighbor(Fi), 4 iNeighborOwner(Fi).
int fs_case_topology(fs_case *cs, int which, int *out, long long count) {
  if (!cs || !out) return -1;
  const auto &g = cs->c.geom;
  const std::vector<int> *v = which == 0 ? &g.nef : which == 1 ? &g.gef : which == 2 ? &g.enbr
                              : which == 3 ? &g.iown : which == 4 ? &g.inbr : nullptr;
  if (!v || (size_t)count != v->size()) {
    set_error("fs_case_topology: bad selector or size");
    return -1;
  }
  std::copy(v->begin(), v->end(), out);
  return 0;
}

} // extern "C"
