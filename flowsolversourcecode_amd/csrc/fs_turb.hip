// The k-omega SST model of the steady SIMPLE iteration on the device
// (TurbulenceModel 'komegasst', ModelNumber 14, low-Reynolds-number walls
// with WallBCModel 'menter'; SolveSelectedTurbulenceModelEquations.for:150):
// strain rate, production, the F1/F2 blending functions, the k and omega
// transport equations and the eddy viscosity.  Face loops are face-ordered
// and form each face's FluxCf/FluxFf/FluxTf in the reference's operand
// order; cells gather their faces in slot order, as the momentum gather
// does, so the assembled systems follow the reference's sums.  The blending
// functions call the device tanh, whose last bit can differ from the host
// library's: the model is compared with the oracle within a tolerance.
#include <hip/hip_runtime.h>

#include "fs_common.h"
#include "fs_hr.h"
#include "fs_kernels.h"
#include "fs_simple.h"

namespace fs {

namespace {

inline int blocks_of(int n) { return std::max(1, (n + kBlock - 1) / kBlock); }

enum { ROLE_OWNER = 1, ROLE_NEIGH = 2, ROLE_BOUND = 3 };

// SetTurbulenceModelCoefficients, 'komegasst' (SetModelCoefficients.for
// :266-300): REAL(4) literals widened to double on assignment.
constexpr double kSigK1 = (double)0.85f, kSigK2 = 1.0, kSigW1 = (double)0.5f, kSigW2 = (double)0.856f;
constexpr double kAlpha1 = (double)(5.0f / 9.0f), kAlpha2 = (double)0.44f;
constexpr double kBeta1 = (double)0.075f, kBeta2 = (double)0.0828f, kBetta = (double)0.075f;
constexpr double kA1 = (double)0.31f, kCmu = (double)0.09f;
constexpr double kC1Lim = 10.0;                               // c1limiter
constexpr double kTiny = 1.e-15;                              // Constants1 tiny
constexpr double kTwoThird = (double)0.666666666666666666666667f; // Constants1 TwoThird (a REAL(4) literal)

// MAX and MIN as the reference build evaluates them: left to right, a NaN
// operand after the first passed over
__device__ __forceinline__ double mx(double a, double b) { return (a > b || b != b) ? a : b; }
__device__ __forceinline__ double mn(double a, double b) { return (a < b || b != b) ? a : b; }
__device__ __forceinline__ double mix(double g, double pi, double pj) { return g * pi + (1. - g) * pj; }

// the boundary gradient of a field at boundary face b: the periodic faces'
// corrected copy (bg, when given), else the owner's cell gradient
__device__ __forceinline__ void bgrad(const double *gr, const double *bg, bool per, int N, int Fb, int b, int i1,
                                      double &gx, double &gy, double &gz) {
  if (per && bg) {
    gx = bg[b];
    gy = bg[Fb + b];
    gz = bg[2 * Fb + b];
  } else {
    gx = gr[i1];
    gy = gr[N + i1];
    gz = gr[2 * N + i1];
  }
}

// CalculateStrainRateTensor: |S| = sqrt(2 S:S), the terms in the reference's
// order (cells, then boundary faces with the boundary velocity gradients)
__device__ __forceinline__ double strain_rate(double ux, double uy, double uz, double vx, double vy, double vz,
                                              double wx, double wy, double wz) {
  const double s11 = ux, s12 = 0.5 * (uy + vx), s13 = 0.5 * (uz + wx), s22 = vy, s23 = 0.5 * (vz + wy), s33 = wz;
  return sqrt(2. * (s11 * s11 + s12 * s12 + s13 * s13 + s12 * s12 + s22 * s22 + s23 * s23 + s13 * s13 + s23 * s23 +
                    s33 * s33));
}

__global__ void k_sst_strain(SstArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int N = a.N, Fb = a.Fb;
  if (t < N) {
    a.sr[t] = strain_rate(a.gu[t], a.gu[N + t], a.gu[2 * N + t], a.gv[t], a.gv[N + t], a.gv[2 * N + t], a.gw[t],
                          a.gw[N + t], a.gw[2 * N + t]);
    return;
  }
  const int b = t - N;
  if (b >= Fb) return;
  const int i1 = a.own[a.Fi + b];
  const bool per = a.btype && a.btype[b] == FS_BC_PERIODIC;
  double g[9];
  bgrad(a.gu, a.bgu, per, N, Fb, b, i1, g[0], g[1], g[2]);
  bgrad(a.gv, a.bgv, per, N, Fb, b, i1, g[3], g[4], g[5]);
  bgrad(a.gw, a.bgw, per, N, Fb, b, i1, g[6], g[7], g[8]);
  a.bsr[b] = strain_rate(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8]);
}

// 2/3 rho k of MinusTwoThirdsRhoK (cells, boundary faces)
__global__ void k_sst_rhok(SstArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < a.N) a.rhok[t] = kTwoThird * a.rho * a.k[t];
  else if (t - a.N < a.Fb) a.brhok[t - a.N] = kTwoThird * a.rho * a.bk[t - a.N];
}

// CalculateEffectiveDiffusionCoefficient: var 0 (k) mu + mu_t sigma_k(F1),
// 1 (omega) mu + mu_t sigma_w(F1), 2 (velocities) mu + mu_t
__global__ void k_sst_ediff(SstArgs a, int var) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const double s1 = var == 1 ? kSigW1 : kSigK1, s2 = var == 1 ? kSigW2 : kSigK2;
  if (t < a.N) {
    if (var == 2) {
      a.med[t] = a.mu + a.mut[t];
    } else {
      const double sl = a.F1[t] * s1 + (1. - a.F1[t]) * s2;
      a.med[t] = a.mu + a.mut[t] * sl;
    }
  } else if (t - a.N < a.Fb) {
    const int b = t - a.N;
    if (var == 2) {
      a.bmed[b] = a.mu + a.bmut[b];
    } else {
      const double sl = a.bF1[b] * s1 + (1. - a.bF1[b]) * s2;
      a.bmed[b] = a.mu + a.bmut[b] * sl;
    }
  }
}

// CalculateTurbulenceProduction, case 14 (CalculateProductionOfTurbulence.for
// :297-316, CalculateShearStressValues.for): P_k = max(tau : grad u, 0),
// tau = 2 mu_t S - 2/3 rho k I, limited to c1limiter cmu rho max(k omega, 0)
__global__ void k_sst_prod(SstArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int N = a.N;
  if (i >= N) return;
  const double ux = a.gu[i], uy = a.gu[N + i], uz = a.gu[2 * N + i];
  const double vx = a.gv[i], vy = a.gv[N + i], vz = a.gv[2 * N + i];
  const double wx = a.gw[i], wy = a.gw[N + i], wz = a.gw[2 * N + i];
  const double s11 = ux, s12 = 0.5 * (uy + vx), s13 = 0.5 * (uz + wx), s22 = vy, s23 = 0.5 * (vz + wy), s33 = wz;
  const double mt = a.mut[i], k = a.k[i];
  double t11 = mt * 2. * s11;
  const double t12 = mt * 2. * s12, t13 = mt * 2. * s13;
  double t22 = mt * 2. * s22;
  const double t23 = mt * 2. * s23;
  double t33 = mt * 2. * s33;
  const double k1 = mx(k, 0.);
  t11 = t11 - kTwoThird * a.rho * k1;
  t22 = t22 - kTwoThird * a.rho * k1;
  t33 = t33 - kTwoThird * a.rho * k1;
  const double p = t11 * ux + t12 * uy + t13 * uz + t12 * vx + t22 * vy + t23 * vz + t13 * wx + t23 * wy + t33 * wz;
  const double lim = kC1Lim * kCmu * a.rho * mx(k * a.om[i], 0.);
  a.prod[i] = mn(mx(p, 0.), lim);
}

// CalculateTurbulenceInterpolationFactors (CalculateTurbulenceModelCoefficients
// .for:3063-3290), no roughness, no rotation corrections: F1, F2 of cells and
// boundary faces (boundary faces with the owner's wall distance)
__device__ __forceinline__ double f1_of(double k, double w, double d, double mu, double rho, double kx, double ky,
                                        double kz, double wx, double wy, double wz) {
  const double tke1 = mx(k, 0.), tom1 = mx(w, kTiny);
  const double f1 = sqrt(tke1) / (kCmu * tom1 * d);
  const double f2 = 500. * mu / (rho * (d * d) * tom1);
  const double f4 = 2. * rho * kSigW2 / tom1;
  const double f5 = kx * wx + ky * wy + kz * wz;
  const double cd = mx(f4 * f5, 1.e-10);
  const double f3 = 4. * rho * kSigW2 * tke1 / (cd * (d * d));
  const double g1 = mn(mn(mx(f1, f2), f3), 10.);
  return tanh(g1 * g1 * g1 * g1);
}
__device__ __forceinline__ double f2_of(double k, double w, double d, double mu, double rho) {
  const double tke1 = mx(k, 0.), tom1 = mx(w, kTiny);
  const double f1 = 2. * sqrt(tke1) / (kCmu * tom1 * d);
  const double f2 = 500. * mu / (rho * (d * d) * tom1);
  const double g2 = mn(mx(f1, f2), 100.);
  return tanh(g2 * g2);
}

__global__ void k_sst_blend(SstArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int N = a.N, Fb = a.Fb;
  if (t < N) {
    a.F1[t] = f1_of(a.k[t], a.om[t], a.wd[t], a.mu, a.rho, a.gk[t], a.gk[N + t], a.gk[2 * N + t], a.gom[t],
                    a.gom[N + t], a.gom[2 * N + t]);
    a.F2[t] = f2_of(a.k[t], a.om[t], a.wd[t], a.mu, a.rho);
    return;
  }
  const int b = t - N;
  if (b >= Fb) return;
  const int i1 = a.own[a.Fi + b];
  const bool per = a.btype && a.btype[b] == FS_BC_PERIODIC;
  double kx, ky, kz, wx, wy, wz;
  bgrad(a.gk, a.bgk, per, N, Fb, b, i1, kx, ky, kz);
  bgrad(a.gom, a.bgom, per, N, Fb, b, i1, wx, wy, wz);
  const double d = a.wd[i1];
  a.bF1[b] = f1_of(a.bk[b], a.bom[b], d, a.mu, a.rho, kx, ky, kz, wx, wy, wz);
  a.bF2[b] = f2_of(a.bk[b], a.bom[b], d, a.mu, a.rho);
}

// Interior faces of the k or omega equation: isotropic diffusion with the
// interpolated effective diffusion coefficient and the 'average' face
// gradient (AssembleIsotropicDiffusion.for:99-117), upwind convection
// (AssembleConvection.for:90-110) and, for HRF, the deferred correction
// (AssembleHRSchemeCorrection.for).
template <int HRF>
__global__ void k_sst_face(SstArgs a, int var, int scheme, const double *__restrict__ g,
                           const double *__restrict__ gdiff, const double *__restrict__ Tx,
                           const double *__restrict__ Ty, const double *__restrict__ Tz,
                           const double *__restrict__ dcx, const double *__restrict__ dcy,
                           const double *__restrict__ dcz, double *__restrict__ cf, double *__restrict__ ff,
                           double *__restrict__ tf) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= a.Fi) return;
  const int N = a.N;
  const double *phi = var ? a.om : a.k, *gr = var ? a.gom : a.gk;
  const int i = a.own[f], j = a.nei[f];
  const double gf = g[f], og = 1. - gf;
  const double gix = gr[i], giy = gr[N + i], giz = gr[2 * N + i];
  const double gjx = gr[j], gjy = gr[N + j], gjz = gr[2 * N + j];
  const double fx = gf * gix + og * gjx, fy = gf * giy + og * gjy, fz = gf * giz + og * gjz;
  const double gamf = gf * a.med[i] + (1. - gf) * a.med[j];
  const double gd = gdiff[f];
  const double cfl = gamf * gd, ffl = -(gamf * gd);
  const double vfl = -(gamf * (fx * Tx[f] + fy * Ty[f] + fz * Tz[f]));
  const double pi = phi[i], pj = phi[j];
  double T = 0.0 + cfl * pi + ffl * pj + vfl;
  const double md = a.mdot[f];
  const double ccl = md > 0. ? md : 0.;
  const double fcl = -(-md > 0. ? -md : 0.);
  T = T + ccl * pi + fcl * pj;
  if constexpr (HRF != FS_HR_NONE) {
    const bool up = md > 0.;
    const double ddx = up ? dcx[f] : -dcx[f], ddy = up ? dcy[f] : -dcy[f], ddz = up ? dcz[f] : -dcz[f];
    const double dphi = hr_increment<HRF>(scheme, 0.0, up ? pi : pj, up ? pj : pi, up ? gix : gjx,
                                          up ? giy : gjy, up ? giz : gjz, ddx, ddy, ddz);
    T = T + md * dphi * 1.;
  }
  cf[f] = (0.0 + cfl) + ccl;
  ff[f] = (0.0 + ffl) + fcl;
  tf[f] = T;
}

// Wall and symmetry faces of the k or omega equation.  Walls: the wall value
// (SetTKEWallBC: k = 0; SetTOmegaWallBC 'menter': omega = 60 mu / (rho beta
// d^2), d the owner's wall distance) and the wall diffusion the source
// routines assemble (AssembleTKESources.for:1063-1093,
// AssembleTOmegaSources.for:680-712); symmetry: no flux, the boundary value
// the owner's (AssembleIsotropicDiffusion.for:2334-2356).
__global__ void k_sst_walls(SstArgs a, int var, const double *__restrict__ gdiff, const double *__restrict__ Tx,
                            const double *__restrict__ Ty, const double *__restrict__ Tz, double *__restrict__ cf,
                            double *__restrict__ ff, double *__restrict__ tf) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= a.Fb) return;
  const int bt = a.btype ? a.btype[b] : FS_BC_WALL;
  if (bt == FS_BC_PERIODIC) return;
  const int f = a.Fi + b, i1 = a.own[f], N = a.N;
  const double *phi = var ? a.om : a.k, *gr = var ? a.gom : a.gk;
  double *bphi = var ? a.bom : a.bk;
  if (bt == FS_BC_SYMMETRY) {
    bphi[b] = phi[i1];
    cf[f] = 0.0;
    ff[f] = 0.0;
    tf[f] = 0.0;
    return;
  }
  double bv;
  if (var == 0) {
    bv = 0.;
  } else {
    const double d = a.wd[i1];
    bv = 10. * 6. * a.mu / (a.rho * kBetta * (d * d));
  }
  bphi[b] = bv;
  const double gamf = a.bmed[b];
  const double gd = gdiff[f];
  const double cfl = gamf * gd, ffl = -(gamf * gd);
  const double vfl = -(gamf * (gr[i1] * Tx[f] + gr[N + i1] * Ty[f] + gr[2 * N + i1] * Tz[f]));
  cf[f] = 0.0 + cfl;
  ff[f] = 0.0 + ffl;
  tf[f] = 0.0 + cfl * phi[i1] + ffl * bv + vfl;
}

// Periodic faces of the k or omega equation: diffusion with BGamaFace
// interpolated between the two cells and the periodic E/T split
// (AssembleIsotropicDiffusion.for:2358-2481), upwind convection with the
// partner's value (AssembleConvection.for:914-943)
__global__ void k_sst_per(SstArgs a, PerArgs p, int var, double *__restrict__ cf, double *__restrict__ ff,
                          double *__restrict__ tf) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= p.np) return;
  const int b = p.b[k], f = p.Fi + b, Fb = p.Fb;
  const int i1 = p.own[f], j1 = p.j1[k];
  const double *phi = var ? a.om : a.k, *bg = var ? a.bgom : a.bgk;
  const double gamf = mix(p.g[k], a.med[i1], a.med[j1]);
  const double cfl = gamf * p.gd[k], ffl = -(gamf * p.gd[k]);
  const double vfl = -(gamf * (bg[b] * p.tx[k] + bg[Fb + b] * p.ty[k] + bg[2 * Fb + b] * p.tz[k]));
  const double ftl = cfl * phi[i1] + ffl * phi[j1] + vfl;
  const double bm = a.bmdot[b];
  const double ccl = bm > 0. ? bm : 0., fcl = -(-bm > 0. ? -bm : 0.);
  cf[f] = (0.0 + cfl) + ccl;
  ff[f] = (0.0 + ffl) + fcl;
  tf[f] = (0.0 + ftl) + ccl * phi[i1] + fcl * phi[j1];
}

// Per-cell gather of the k or omega system: the slots' face fluxes in slot
// order, the sources (AssembleTKESources.for:985-995, AssembleTOmega
// Sources.for:558-583), the convection's element terms
// (AssembleConvection.for:169-178), under-relaxation.
template <bool PER>
__device__ __forceinline__ double sst_slot(int role, double vc, double vf, double vt, double md, double &a,
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
    e = e + md;
  }
  return an;
}

template <int WC, bool PER>
__global__ void k_sst_cell(SstArgs a, int var, int W, const int *__restrict__ code, const double *__restrict__ cf,
                           const double *__restrict__ ff, const double *__restrict__ tf,
                           const double *__restrict__ vol, double urf, double *__restrict__ ac,
                           double *__restrict__ anb, double *__restrict__ bc) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  const int N = a.N;
  if (c >= N) return;
  double acc = 0.0, b = 0.0, e = 0.0;
  auto slot = [&](int s, int cd) {
    const int f = cd >> 2, role = cd & 3;
    double md = 0.0;
    if (role == ROLE_OWNER || role == ROLE_NEIGH) md = a.mdot[f];
    else if (role == ROLE_BOUND) md = a.bmdot[f - a.Fi];
    anb[(size_t)s * N + c] = sst_slot<PER>(role, cf[f], ff[f], tf[f], md, acc, b, e);
  };
  if constexpr (WC > 0) {
#pragma unroll
    for (int s = 0; s < WC; ++s) slot(s, code[(size_t)s * N + c]);
  } else {
    for (int s = 0; s < W; ++s) slot(s, code[(size_t)s * N + c]);
  }
  const double v = vol[c];
  const double phi = var ? a.om[c] : a.k[c];
  double fce = 0.0, fte = 0.0;
  if (var == 0) {
    const double rted = kCmu * a.rho * mx(a.om[c], 0.);
    const double pt = 1. * a.prod[c]; // fr1Coefficient = 1
    fce = fce + rted * v;
    fte = fte - pt * v + rted * v * mx(a.k[c], 0.);
  } else {
    const double w = a.om[c], f1 = a.F1[c];
    double t1 = a.rho * a.prod[c] * 1. / mx(a.mut[c], 1.e-10);
    const double al = f1 * kAlpha1 + (1. - f1) * kAlpha2;
    fte = fte - v * mx(al * t1, 0.);
    t1 = a.rho * mx(w, 0.) * v;
    double be = f1 * kBeta1 + (1. - f1) * kBeta2;
    be = 1. * be; // F4factor = 1
    fce = fce + 2. * mx(be * t1, 0.);
    fte = fte + mx(be * t1 * w, 0.);
    t1 = 2. * a.rho * kSigW2 / mx(w, kTiny);
    const double t2 =
        a.gk[c] * a.gom[c] + a.gk[N + c] * a.gom[N + c] + a.gk[2 * N + c] * a.gom[2 * N + c];
    t1 = (1. - f1) * t1 * t2 * v;
    if (t1 > 0.) {
      fte = fte - t1;
    } else {
      fce = fce - t1 / mx(w, kTiny);
      fte = fte - t1;
    }
  }
  const double ep = e > 0.0 ? e : 0.0;
  const double fcel = ep - e;
  fce = fce + fcel;
  fte = fte + (fcel * phi + (-(ep * phi)));
  acc = acc + fce;
  b = b - fte;
  ac[c] = acc / urf;
  bc[c] = b;
}

// BoundTurbulentKE / BoundTurbulentOmega: cells at tiny, boundary values at 0
__global__ void k_sst_bound_cells(SstArgs a, int var) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.N) return;
  double *phi = var ? a.om : a.k;
  phi[i] = mx(phi[i], kTiny);
}
// then updateValuesFromBoundaryConditions (UpdateVariablesFromBoundary
// Conditions.for:76-146,431-501): symmetry the owner's value, periodic the
// value interpolated between the two cells; walls keep theirs
__global__ void k_sst_bound_faces(SstArgs a, PerArgs p, int var) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= a.Fb) return;
  const double *phi = var ? a.om : a.k;
  double *bphi = var ? a.bom : a.bk;
  double v = mx(bphi[b], 0.);
  const int bt = a.btype ? a.btype[b] : FS_BC_WALL;
  const int i1 = a.own[a.Fi + b];
  if (bt == FS_BC_SYMMETRY) {
    v = phi[i1];
  } else if (bt == FS_BC_PERIODIC) {
    const int k = a.pk[b];
    v = mix(p.g[k], phi[i1], phi[p.j1[k]]);
  }
  bphi[b] = v;
}

// CalculateTurbulentViscosity, case 14 (CalculateTurbulentViscosity.for
// :483-513,925-938): relaxed rho a1 k / max(a1 omega, |S| F2 F3, tiny), capped
// at MaximumViscosityRatio times the viscosity
__global__ void k_sst_mut(SstArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const double u = a.urf_mut;
  if (t < a.N) {
    const double den = mx(mx(kA1 * a.om[t], a.sr[t] * a.F2[t] * 1.), kTiny);
    const double m = (1. - u) * a.mut[t] + u * a.rho * kA1 * mx(a.k[t], 0.) / den;
    a.mut[t] = mn(m, a.max_ratio * a.mu);
  } else if (t - a.N < a.Fb) {
    const int b = t - a.N;
    const double den = mx(mx(kA1 * a.bom[b], a.bsr[b] * a.bF2[b] * 1.), kTiny);
    const double m = (1. - u) * a.bmut[b] + u * a.rho * kA1 * mx(a.bk[b], 0.) / den;
    a.bmut[b] = mn(m, a.max_ratio * a.mu);
  }
}

// ---- CalculateNormalDistanceToWall, method 2 (CalculateDistanceToWall.for
// :154-470): lap(phi) = -1, phi = 0 on the no-slip walls, no flux elsewhere.
// Interior faces: diffusion with the 'average' face gradient
__global__ void k_wd_faces(int Fi, const int *__restrict__ own, const int *__restrict__ nei,
                           const double *__restrict__ g, const double *__restrict__ gdiff,
                           const double *__restrict__ Tx, const double *__restrict__ Ty,
                           const double *__restrict__ Tz, int N, const double *__restrict__ phi,
                           const double *__restrict__ gr, double *__restrict__ cf, double *__restrict__ ff,
                           double *__restrict__ tf) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= Fi) return;
  const int i = own[f], j = nei[f];
  const double gf = g[f], og = 1. - gf;
  const double fx = gf * gr[i] + og * gr[j], fy = gf * gr[N + i] + og * gr[N + j],
               fz = gf * gr[2 * N + i] + og * gr[2 * N + j];
  const double cfl = gdiff[f], ffl = -gdiff[f];
  const double vfl = -fx * Tx[f] - fy * Ty[f] - fz * Tz[f];
  cf[f] = 0.0 + cfl;
  ff[f] = 0.0 + ffl;
  tf[f] = 0.0 + cfl * phi[i] + ffl * phi[j] + vfl;
}
// Boundary faces: walls phi = 0 with their diffusion flux; the others take
// the owner's value and carry no flux
__global__ void k_wd_bfaces(int Fb, int Fi, int N, const int *__restrict__ own, const int *__restrict__ btype,
                            const double *__restrict__ gdiff, const double *__restrict__ Tx,
                            const double *__restrict__ Ty, const double *__restrict__ Tz,
                            const double *__restrict__ phi, const double *__restrict__ gr, double *__restrict__ bphi,
                            double *__restrict__ cf, double *__restrict__ ff, double *__restrict__ tf) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= Fb) return;
  const int f = Fi + b, i1 = own[f];
  if (btype && btype[b] != FS_BC_WALL) {
    bphi[b] = phi[i1];
    cf[f] = 0.0 + 0.;
    ff[f] = 0.0 + 0.;
    tf[f] = 0.0 + 0.;
    return;
  }
  bphi[b] = 0.;
  const double cfl = gdiff[f], ffl = -gdiff[f];
  const double vfl = -(gr[i1] * Tx[f] + gr[N + i1] * Ty[f] + gr[2 * N + i1] * Tz[f]);
  cf[f] = 0.0 + cfl;
  ff[f] = 0.0 + ffl;
  tf[f] = 0.0 + cfl * phi[i1] + ffl * 0. + vfl;
}
// The cell gather with the unit source (AssembleSourceTerm, ScWall = 0,
// SbWall = 1) and urfWall = 1; boundary slots carry no coupling
__global__ void k_wd_cell(int N, int W, const int *__restrict__ code, const double *__restrict__ cf,
                          const double *__restrict__ ff, const double *__restrict__ tf,
                          const double *__restrict__ vol, const double *__restrict__ phi, double *__restrict__ ac,
                          double *__restrict__ anb, double *__restrict__ bc) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c >= N) return;
  double a = 0.0, b = 0.0;
  for (int s = 0; s < W; ++s) {
    const int cd = code[(size_t)s * N + c];
    const int f = cd >> 2, role = cd & 3;
    double an = 0.0;
    if (role == ROLE_OWNER) {
      a = a + cf[f];
      an = an + ff[f];
      b = b - tf[f];
    } else if (role == ROLE_NEIGH) {
      a = a - ff[f];
      an = an - cf[f];
      b = b + tf[f];
    } else if (role == ROLE_BOUND) {
      a = a + cf[f];
      an = an + ff[f];
      b = b - tf[f];
    }
    anb[(size_t)s * N + c] = an;
  }
  const double fte = 0.0 - (1. + 0. * phi[c]) * vol[c];
  a = a + (0.0 + 0.);
  ac[c] = a / 1.;
  bc[c] = b - fte;
}
// d = -|grad phi| + sqrt(|grad phi|^2 + 2 phi); the cells at a wall take the
// normal distance of their last wall face (in boundary order)
__global__ void k_wd_final(int N, int W, int Fi, const int *__restrict__ code, const int *__restrict__ btype,
                           const double *__restrict__ phi, const double *__restrict__ gr,
                           const double *__restrict__ dx, const double *__restrict__ dy,
                           const double *__restrict__ dz, const double *__restrict__ nx,
                           const double *__restrict__ ny, const double *__restrict__ nz, double *__restrict__ wd) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= N) return;
  const double gx = gr[c], gy = gr[N + c], gz = gr[2 * N + c];
  const double t1 = gx * gx + gy * gy + gz * gz;
  const double t2 = t1 + 2. * phi[c];
  double d = -sqrt(t1) + sqrt(t2);
  int last = -1;
  for (int s = 0; s < W; ++s) {
    const int cd = code[(size_t)s * N + c];
    const int f = cd >> 2;
    if ((cd & 3) == ROLE_BOUND && (!btype || btype[f - Fi] == FS_BC_WALL) && f > last) last = f;
  }
  if (last >= 0) d = dx[last] * nx[last] + dy[last] * ny[last] + dz[last] * nz[last];
  wd[c] = d;
}

} // namespace

void wall_distance_faces(hipStream_t s, int Fi, int Fb, int N, const int *own, const int *nei, const int *btype,
                         const double *g, const double *gdiff, const double *Tx, const double *Ty,
                         const double *Tz, const double *phi, const double *gr, double *bphi, double *cf,
                         double *ff, double *tf) {
  hipLaunchKernelGGL(k_wd_faces, dim3(blocks_of(Fi)), dim3(kBlock), 0, s, Fi, own, nei, g, gdiff, Tx, Ty, Tz, N, phi,
                     gr, cf, ff, tf);
  hipLaunchKernelGGL(k_wd_bfaces, dim3(blocks_of(Fb)), dim3(kBlock), 0, s, Fb, Fi, N, own, btype, gdiff, Tx, Ty, Tz,
                     phi, gr, bphi, cf, ff, tf);
}
void wall_distance_cells(hipStream_t s, int N, int W, const int *code, const double *cf, const double *ff,
                         const double *tf, const double *vol, const double *phi, double *ac, double *anb,
                         double *bc) {
  hipLaunchKernelGGL(k_wd_cell, dim3(blocks_of(N)), dim3(kBlock), 0, s, N, W, code, cf, ff, tf, vol, phi, ac, anb,
                     bc);
}
void wall_distance_final(hipStream_t s, int N, int W, int Fi, const int *code, const int *btype, const double *phi,
                         const double *gr, const double *dx, const double *dy, const double *dz, const double *nx,
                         const double *ny, const double *nz, double *wd) {
  hipLaunchKernelGGL(k_wd_final, dim3(blocks_of(N)), dim3(kBlock), 0, s, N, W, Fi, code, btype, phi, gr, dx, dy, dz,
                     nx, ny, nz, wd);
}

void sst_strain(const SstArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_sst_strain, dim3(blocks_of(a.N + a.Fb)), dim3(kBlock), 0, s, a);
}
void sst_rhok(const SstArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_sst_rhok, dim3(blocks_of(a.N + a.Fb)), dim3(kBlock), 0, s, a);
}
void sst_ediff(const SstArgs &a, hipStream_t s, int var) {
  hipLaunchKernelGGL(k_sst_ediff, dim3(blocks_of(a.N + a.Fb)), dim3(kBlock), 0, s, a, var);
}
void sst_production(const SstArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_sst_prod, dim3(blocks_of(a.N)), dim3(kBlock), 0, s, a);
}
void sst_blending(const SstArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_sst_blend, dim3(blocks_of(a.N + a.Fb)), dim3(kBlock), 0, s, a);
}
void sst_faces(const SstArgs &a, const PerArgs *p, hipStream_t s, int var, int framework, int scheme,
               const double *g, const double *gdiff, const double *Tx, const double *Ty, const double *Tz,
               const double *dcx, const double *dcy, const double *dcz, double *cf, double *ff, double *tf) {
  const int fw = scheme == FS_CONV_UPWIND ? FS_HR_NONE : framework;
  auto face = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(blocks_of(a.Fi)), dim3(kBlock), 0, s, a, var, scheme, g, gdiff, Tx, Ty, Tz, dcx,
                       dcy, dcz, cf, ff, tf);
  };
  if (fw == FS_HR_TVD) face(k_sst_face<FS_HR_TVD>);
  else if (fw == FS_HR_NVF) face(k_sst_face<FS_HR_NVF>);
  else face(k_sst_face<FS_HR_NONE>);
  hipLaunchKernelGGL(k_sst_walls, dim3(blocks_of(a.Fb)), dim3(kBlock), 0, s, a, var, gdiff, Tx, Ty, Tz, cf, ff, tf);
  if (p && p->np > 0)
    hipLaunchKernelGGL(k_sst_per, dim3(blocks_of(p->np)), dim3(kBlock), 0, s, a, *p, var, cf, ff, tf);
}
void sst_cells(const SstArgs &a, bool per, hipStream_t s, int var, int W, const int *code, const double *cf,
               const double *ff, const double *tf, const double *vol, double urf, double *ac, double *anb,
               double *bc) {
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(blocks_of(a.N)), dim3(kBlock), 0, s, a, var, W, code, cf, ff, tf, vol, urf, ac,
                       anb, bc);
  };
  switch (W * 2 + (per ? 1 : 0)) {
  case 8: launch(k_sst_cell<4, false>); break;
  case 9: launch(k_sst_cell<4, true>); break;
  case 12: launch(k_sst_cell<6, false>); break;
  case 13: launch(k_sst_cell<6, true>); break;
  default:
    if (per) launch(k_sst_cell<0, true>);
    else launch(k_sst_cell<0, false>);
  }
}
void sst_bound(const SstArgs &a, const PerArgs &p, hipStream_t s, int var) {
  hipLaunchKernelGGL(k_sst_bound_cells, dim3(blocks_of(a.N)), dim3(kBlock), 0, s, a, var);
  hipLaunchKernelGGL(k_sst_bound_faces, dim3(blocks_of(a.Fb)), dim3(kBlock), 0, s, a, p, var);
}
void sst_viscosity(const SstArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_sst_mut, dim3(blocks_of(a.N + a.Fb)), dim3(kBlock), 0, s, a);
}

} // namespace fs
