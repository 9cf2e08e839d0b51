// Translational periodic pairs with the implicit coupling: the periodic
// branches of the reference's SIMPLE iteration, one thread per periodic face
// (or per lower-patch face where the reference's loop visits only the faces
// of the lower patch of a pair, j2 > i2, and copies to the partner).  Each
// formula keeps the reference's operand order; the rotation of a
// translational pair is the identity, applied as the reference applies it
// (a1r x + b1r y + c1r z with (1, 0, 0), ...).
#include <hip/hip_runtime.h>

#include "fs_common.h"
#include "fs_kernels.h"
#include "fs_simple.h"

namespace fs {

namespace {

inline int blocks_of(int n) { return std::max(1, (n + kBlock - 1) / kBlock); }

__device__ __forceinline__ double rot(int comp, double x, double y, double z) {
  const double a = comp == 0 ? 1. : 0., b = comp == 1 ? 1. : 0., c = comp == 2 ? 1. : 0.;
  return a * x + b * y + c * z;
}

__device__ __forceinline__ double mix(double g, double pi, double pj) { return g * pi + (1. - g) * pj; }

// geoDiff of a D S split by MethodDecomposeS (AssembleMdotUsingRhieChow.for
// :575-615, AssemblePressureCorrectionFluxes.for, PeriodicBeta.for)
__device__ __forceinline__ double geo_split_d(int dec, double dusf, double dvsf, double dwsf, double ex, double ey,
                                              double ez, double dcf, double &due, double &dve, double &dwe) {
  if (dec == 1) {
    const double dot = ex * dusf + ey * dvsf + ez * dwsf;
    due = dot * ex;
    dve = dot * ey;
    dwe = dot * ez;
  } else if (dec == 2) {
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

// CorrectGradientAtPeriodicBoundary, translational (CorrectGradients.for
// :141-167): the lower patch's boundary gradient interpolated between the
// two cells; the partner face takes its x and y components, its z component
// stays the owner's copy.
struct BgradArgs {
  const double *cellg[4];
  double *bg[4];
};
__global__ void k_per_bgrad(PerArgs a, int nf, BgradArgs x) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.na) return;
  const int k = a.act[t], b = a.b[k], k2 = a.part[k], b2 = a.b[k2];
  const int i1 = a.own[a.Fi + b], j1 = a.j1[k], i2 = a.own[a.Fi + b2];
  const double g = a.g[k];
  const int N = a.N, Fb = a.Fb;
  for (int q = 0; q < nf; ++q) {
    const double *gr = x.cellg[q];
    double *bg = x.bg[q];
    const double bx = mix(g, gr[i1], gr[j1]), by = mix(g, gr[N + i1], gr[N + j1]),
                 bz = mix(g, gr[2 * N + i1], gr[2 * N + j1]);
    bg[b] = bx;
    bg[Fb + b] = by;
    bg[2 * Fb + b] = bz;
    bg[b2] = bx;
    bg[Fb + b2] = by;
    bg[2 * Fb + b2] = gr[2 * N + i2];
  }
}

struct ValArgs {
  const double *phi[3];
  double *bphi[3];
};
__global__ void k_per_values(PerArgs a, int nf, ValArgs x) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.na) return;
  const int k = a.act[t], b = a.b[k], b2 = a.b[a.part[k]];
  const int i1 = a.own[a.Fi + b], j1 = a.j1[k];
  for (int q = 0; q < nf; ++q) {
    const double v = mix(a.g[k], x.phi[q][i1], x.phi[q][j1]);
    x.bphi[q][b] = v;
    x.bphi[q][b2] = v;
  }
}

// AssembleIsotropicDiffusion.for:630-700,746-770 (translational, implicit),
// the boundary loops of AssembleExplicitDiffusion.for and
// AssembleBulkViscosity.for (BeDiffCoefficient = the viscosity), and
// AssembleConvection.for:582-627, for one momentum component.
__global__ void k_per_momentum(PerArgs a, int comp, double mu, double c23, const double *__restrict__ visc,
                               const double *__restrict__ bvisc, const double *__restrict__ phi, const double *__restrict__ u,
                               const double *__restrict__ v, const double *__restrict__ w,
                               const double *__restrict__ bgu, const double *__restrict__ bgv,
                               const double *__restrict__ bgw, const double *__restrict__ Sx,
                               const double *__restrict__ Sy, const double *__restrict__ Sz,
                               const double *__restrict__ bmdot, double *__restrict__ cf, double *__restrict__ ff,
                               double *__restrict__ tf) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.np) return;
  const int b = a.b[k], f = a.Fi + b, Fb = a.Fb;
  const int i1 = a.own[f], j1 = a.j1[k];
  const double gamf = visc ? mix(a.g[k], visc[i1], visc[j1]) : mix(a.g[k], mu, mu); // BGamaFace
  const double *bg = comp == 0 ? bgu : comp == 1 ? bgv : bgw;
  const double cfl = gamf * a.gd[k], ffl = -(gamf * a.gd[k]);
  const double vfl = -(gamf * (bg[b] * a.tx[k] + bg[Fb + b] * a.ty[k] + bg[2 * Fb + b] * a.tz[k]));
  double C = 0.0 + cfl, Fv = 0.0 + ffl;
  double T = 0.0 + cfl * phi[i1] + ffl * phi[j1] + vfl;
  // explicit diffusion: -Gamma_b (grad u_comp)_b . S
  const double gb = bvisc ? bvisc[b] : mu; // the face's BeDiffCoefficient
  const double sx = Sx[f], sy = Sy[f], sz = Sz[f];
  const double v2 = -(gb * (bgu[(size_t)comp * Fb + b] * sx + bgv[(size_t)comp * Fb + b] * sy +
                            bgw[(size_t)comp * Fb + b] * sz));
  C = C + 0.;
  Fv = Fv + 0.;
  T = T + v2;
  // bulk viscosity
  const double S = comp == 0 ? sx : comp == 1 ? sy : sz;
  const double v3 = -(c23 * gb * (bgu[b] + bgv[Fb + b] + bgw[2 * Fb + b]) * S);
  C = C + 0.;
  Fv = Fv + 0.;
  T = T + v3;
  // upwind convection with the partner cell's value
  const double phib = rot(comp, u[j1], v[j1], w[j1]);
  const double bm = bmdot[b];
  const double ccl = bm > 0. ? bm : 0., fcl = -(-bm > 0. ? -bm : 0.);
  C = C + ccl;
  Fv = Fv + fcl;
  T = T + (ccl * phi[i1] + fcl * phib);
  cf[f] = C;
  ff[f] = Fv;
  tf[f] = T;
}

// AssembleMdotUsingRhieChow.for:1086-1416 (translational, steady,
// LRelaxMomentum): Bmdot of a lower-patch face, its negative on the partner
struct RcPer {
  double rho, urf_u;
  const double *u, *v, *w, *us, *vs, *ws, *du, *dv, *dw, *p, *bgp, *Sx, *Sy, *Sz;
  const double *rhok, *bgrk; // 2/3 rho k and its boundary gradient, or nullptr
};
__global__ void k_per_rhie_chow(PerArgs a, int dec, RcPer r, double *__restrict__ bmdot) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.na) return;
  const int k = a.act[t], b = a.b[k], b2 = a.b[a.part[k]], f = a.Fi + b, Fb = a.Fb;
  const int i1 = a.own[f], j1 = a.j1[k];
  const double gf = a.g[k], rhof = r.rho;
  const double sx = r.Sx[f], sy = r.Sy[f], sz = r.Sz[f];
  const double ex = a.ux[k], ey = a.uy[k], ez = a.uz[k], cf = a.dcf[k];
  const double uf1 = rot(0, r.u[j1], r.v[j1], r.w[j1]), vf1 = rot(1, r.u[j1], r.v[j1], r.w[j1]),
               wf1 = rot(2, r.u[j1], r.v[j1], r.w[j1]);
  const double ub = gf * r.u[i1] + (1. - gf) * uf1, vb = gf * r.v[i1] + (1. - gf) * vf1,
               wb = gf * r.w[i1] + (1. - gf) * wf1;
  double fv = rhof * (ub * sx + vb * sy + wb * sz);
  double dusf = gf * r.du[i1] + (1. - gf) * r.du[j1];
  dusf = dusf * sx;
  double dvsf = gf * r.dv[i1] + (1. - gf) * r.dv[j1];
  dvsf = dvsf * sy;
  double dwsf = gf * r.dw[i1] + (1. - gf) * r.dw[j1];
  dwsf = dwsf * sz;
  double due, dve, dwe;
  const double geo = geo_split_d(dec, dusf, dvsf, dwsf, ex, ey, ez, cf, due, dve, dwe);
  const double dut = dusf - due, dvt = dvsf - dve, dwt = dwsf - dwe;
  const double pbx = r.bgp[b], pby = r.bgp[Fb + b], pbz = r.bgp[2 * Fb + b];
  double corr = (r.p[j1] - r.p[i1]) / cf;
  corr = corr - (pbx * ex + pby * ey + pbz * ez);
  const double pfx = pbx + corr * ex, pfy = pby + corr * ey, pfz = pbz + corr * ez;
  const double cfl = rhof * geo, ffl = -(rhof * geo);
  fv = fv - rhof * (pfx * dut + pfy * dvt + pfz * dwt);
  fv = fv + rhof * (pbx * dusf + pby * dvsf + pbz * dwsf);
  if (r.rhok) { // the -2/3 rho k gradient's difference (:1222-1242)
    const double kbx = r.bgrk[b], kby = r.bgrk[Fb + b], kbz = r.bgrk[2 * Fb + b];
    double kc = (r.rhok[j1] - r.rhok[i1]) / cf;
    kc = kc - (kbx * ex + kby * ey + kbz * ez);
    const double kfx = kbx + kc * ex, kfy = kby + kc * ey, kfz = kbz + kc * ez;
    fv = fv - rhof * (kfx * dut + kfy * dvt + kfz * dwt);
    fv = fv + rhof * (kbx * dusf + kby * dvsf + kbz * dwsf);
    fv = fv + cfl * r.rhok[i1] + ffl * r.rhok[j1];
  }
  const double us1 = rot(0, r.us[j1], r.vs[j1], r.ws[j1]), vs1 = rot(1, r.us[j1], r.vs[j1], r.ws[j1]),
               ws1 = rot(2, r.us[j1], r.vs[j1], r.ws[j1]);
  const double usf = gf * r.us[i1] + (1. - gf) * us1, vsf = gf * r.vs[i1] + (1. - gf) * vs1,
               wsf = gf * r.ws[i1] + (1. - gf) * ws1;
  const double mstar = rhof * (usf * sx + vsf * sy + wsf * sz);
  fv = fv + (1. - r.urf_u) * (bmdot[b] - mstar);
  const double m = cfl * r.p[i1] + ffl * r.p[j1] + fv;
  bmdot[b] = m;
  bmdot[b2] = -m;
}

// p' fluxes of the boundary faces (AssemblePressureCorrectionFluxes.for
// :155-164 FluxTf = Bmdot; periodic faces :748-830): walls carry none
__global__ void k_per_pc_boundary(PerArgs a, int dec, double rho, const int *__restrict__ pk,
                                  const double *__restrict__ d2u, const double *__restrict__ d2v,
                                  const double *__restrict__ d2w, const double *__restrict__ Sx,
                                  const double *__restrict__ Sy, const double *__restrict__ Sz,
                                  const double *__restrict__ bmdot, double *__restrict__ cf,
                                  double *__restrict__ ff, double *__restrict__ tf) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= a.Fb) return;
  const int f = a.Fi + b, k = pk[b];
  tf[f] = bmdot[b];
  if (k < 0) {
    cf[f] = 0.0;
    ff[f] = 0.0;
    return;
  }
  const int i1 = a.own[f], j1 = a.j1[k];
  const double gf = a.g[k];
  double dusf, dvsf, dwsf;
  if (dec == 3) {
    dusf = gf * d2u[i1] + (1. - gf) * d2u[j1];
    dvsf = gf * d2v[i1] + (1. - gf) * d2v[j1];
    dwsf = gf * d2w[i1] + (1. - gf) * d2w[j1];
  } else {
    const double duf1 = rot(0, d2u[j1], d2v[j1], d2w[j1]), dvf1 = rot(1, d2u[j1], d2v[j1], d2w[j1]),
                 dwf1 = rot(2, d2u[j1], d2v[j1], d2w[j1]);
    dusf = gf * d2u[i1] + (1. - gf) * duf1;
    dvsf = gf * d2v[i1] + (1. - gf) * dvf1;
    dwsf = gf * d2w[i1] + (1. - gf) * dwf1;
  }
  dusf = dusf * Sx[f];
  dvsf = dvsf * Sy[f];
  dwsf = dwsf * Sz[f];
  double due, dve, dwe;
  const double geo = geo_split_d(dec, dusf, dvsf, dwsf, a.ux[k], a.uy[k], a.uz[k], a.dcf[k], due, dve, dwe);
  cf[f] = rho * geo;
  ff[f] = -(rho * geo);
}

// CorrectBoundarymdot, periodic faces (CorrectMdotandVelocity.for:115-145)
__global__ void k_per_correct_bmdot(PerArgs a, const double *__restrict__ d2u, const double *__restrict__ d2v,
                                    const double *__restrict__ d2w, const double *__restrict__ gpc,
                                    const double *__restrict__ pc, const double *__restrict__ cf,
                                    const double *__restrict__ ff, double *__restrict__ bu,
                                    double *__restrict__ bv, double *__restrict__ bw, double *__restrict__ bmdot) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.na) return;
  const int k = a.act[t], b = a.b[k], b2 = a.b[a.part[k]], f = a.Fi + b, N = a.N;
  const int i1 = a.own[f], j1 = a.j1[k];
  const double rhof = 1.;
  const double nu = bu[b] - d2u[i1] * gpc[i1], nv = bv[b] - d2v[i1] * gpc[N + i1],
               nw = bw[b] - d2w[i1] * gpc[2 * N + i1];
  bu[b] = nu;
  bv[b] = nv;
  bw[b] = nw;
  const double m = bmdot[b] + rhof * cf[f] * pc[i1] + rhof * ff[f] * pc[j1];
  bmdot[b] = m;
  bu[b2] = rot(0, nu, nv, nw);
  bv[b2] = rot(1, nu, nv, nw);
  bw[b2] = rot(2, nu, nv, nw);
  bmdot[b2] = -m;
}

// geoDiff of the D2 S split of a lower-patch face as updateperiodicBeta forms
// it (PeriodicBeta.for:75-134)
__device__ double beta_geo(const PerArgs &a, int dec, int k, const double *d2u, const double *d2v,
                           const double *d2w, const double *Sx, const double *Sy, const double *Sz) {
  const int f = a.Fi + a.b[k], i1 = a.own[f], j1 = a.j1[k];
  const double gf = a.g[k];
  double dusf = gf * d2u[i1] + (1. - gf) * d2u[j1];
  dusf = dusf * Sx[f];
  double dvsf = gf * d2v[i1] + (1. - gf) * d2v[j1];
  dvsf = dvsf * Sy[f];
  double dwsf = gf * d2w[i1] + (1. - gf) * d2w[j1];
  dwsf = dwsf * Sz[f];
  double due, dve, dwe;
  return geo_split_d(dec, dusf, dvsf, dwsf, a.ux[k], a.uy[k], a.uz[k], a.dcf[k], due, dve, dwe);
}

// the sums of updateperiodicBeta over the lower-patch faces in the
// reference's sequential order (one thread: the sums must round as the
// reference's loop does), dp'/dx and periodicBeta
__global__ void k_beta_sum(PerArgs a, int dec, double rho, double target, double relax, int iters,
                           const double *__restrict__ d2u, const double *__restrict__ d2v,
                           const double *__restrict__ d2w, const double *__restrict__ Sx,
                           const double *__restrict__ Sy, const double *__restrict__ Sz,
                           const double *__restrict__ bmdot) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double dpdx = 0.;
  for (int it = 0; it < iters; ++it) {
    double mdin = 0., rgd = 0.;
    for (int t = 0; t < a.na; ++t) {
      const int k = a.act[t];
      const double geo = beta_geo(a, dec, k, d2u, d2v, d2w, Sx, Sy, Sz);
      mdin = mdin + bmdot[a.b[k]];
      rgd = rgd + rho * geo;
    }
    dpdx = (target - mdin) / rgd;
    a.scal[PSCAL_BETA] = a.scal[PSCAL_BETA] + relax * dpdx;
  }
  a.scal[PSCAL_DPDX] = dpdx;
}

__global__ void k_beta_faces(PerArgs a, int dec, double rho, const double *__restrict__ d2u,
                             const double *__restrict__ d2v, const double *__restrict__ d2w,
                             const double *__restrict__ Sx, const double *__restrict__ Sy,
                             const double *__restrict__ Sz, double *__restrict__ bu, double *__restrict__ bv,
                             double *__restrict__ bw, double *__restrict__ bmdot) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.na) return;
  const int k = a.act[t], b = a.b[k], b2 = a.b[a.part[k]], i1 = a.own[a.Fi + b];
  const double dpdx = a.scal[PSCAL_DPDX];
  bu[b] = bu[b] + d2u[i1] * dpdx;
  bv[b] = bv[b] + d2v[i1] * dpdx;
  bw[b] = bw[b] + d2w[i1] * dpdx;
  bu[b2] = bu[b];
  bv[b2] = bv[b];
  bw[b2] = bw[b];
  const double geo = beta_geo(a, dec, k, d2u, d2v, d2w, Sx, Sy, Sz);
  const double m = bmdot[b] + rho * geo * dpdx;
  bmdot[b] = m;
  bmdot[b2] = -m;
}

__global__ void k_beta_cells(int N, const double *__restrict__ scal, double bx, double by, double bz,
                             const double *__restrict__ d2u, const double *__restrict__ d2v,
                             const double *__restrict__ d2w, double *__restrict__ u, double *__restrict__ v,
                             double *__restrict__ w) {
  const int c = xcd_block() * kBlock + threadIdx.x;
  if (c >= N) return;
  const double dpdx = scal[PSCAL_DPDX];
  u[c] = u[c] + d2u[c] * dpdx * bx;
  v[c] = v[c] + d2v[c] * dpdx * by;
  w[c] = w[c] + d2w[c] * dpdx * bz;
}

// every interior mass flux += rho geoDiff(D2) dp'/dx (PeriodicBeta.for:230-290)
__global__ void k_beta_interior(int Fi, int dec, double rho, const double *__restrict__ scal,
                                const int *__restrict__ own, const int *__restrict__ nei,
                                const double *__restrict__ g, const double *__restrict__ Sx,
                                const double *__restrict__ Sy, const double *__restrict__ Sz,
                                const double *__restrict__ ex, const double *__restrict__ ey,
                                const double *__restrict__ ez, const double *__restrict__ dcf,
                                const double *__restrict__ d2u, const double *__restrict__ d2v,
                                const double *__restrict__ d2w, double *__restrict__ mdot) {
  const int f = xcd_block() * kBlock + threadIdx.x;
  if (f >= Fi) return;
  const int i = own[f], j = nei[f];
  const double gf = g[f];
  const double d1u = gf * d2u[i] + (1. - gf) * d2u[j], d1v = gf * d2v[i] + (1. - gf) * d2v[j],
               d1w = gf * d2w[i] + (1. - gf) * d2w[j];
  double due, dve, dwe;
  const double geo = geo_split_d(dec, d1u * Sx[f], d1v * Sy[f], d1w * Sz[f], ex[f], ey[f], ez[f], dcf[f], due,
                                 dve, dwe);
  mdot[f] = mdot[f] + rho * geo * scal[PSCAL_DPDX];
}

__global__ void k_owner_values(int Fb, int Fi, const int *__restrict__ own, const double *__restrict__ phi,
                               double *__restrict__ bphi) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b < Fb) bphi[b] = phi[own[Fi + b]];
}

// beta along each axis (AssembleLinearPressure.for:18-33)
__global__ void k_per_linear(double *scal, double tx, double ty, double tz) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double lt = sqrt(tx * tx + ty * ty + tz * tz);
  scal[PSCAL_LIN] = scal[PSCAL_BETA] * tx / lt;
  scal[PSCAL_LIN + 1] = scal[PSCAL_BETA] * ty / lt;
  scal[PSCAL_LIN + 2] = scal[PSCAL_BETA] * tz / lt;
}

} // namespace

PerArgs periodic_args(const Periodic &P, int N, int Fi, int Fb, const int *own) {
  PerArgs a{};
  a.N = N;
  a.Fi = Fi;
  a.Fb = Fb;
  a.np = P.np;
  a.na = P.na;
  a.own = own;
  a.b = P.b.p;
  a.part = P.part.p;
  a.j1 = P.j1.p;
  a.act = P.act.p;
  a.g = P.g.p;
  a.ux = P.ux.p;
  a.uy = P.uy.p;
  a.uz = P.uz.p;
  a.dcf = P.dcf.p;
  a.gd = P.gd.p;
  a.tx = P.tx.p;
  a.ty = P.ty.p;
  a.tz = P.tz.p;
  a.scal = P.scal.p;
  return a;
}

void periodic_bgrad(const PerArgs &a, hipStream_t s, int nf, const double *const *cellg, double *const *bg) {
  BgradArgs x{};
  for (int q = 0; q < nf; ++q) {
    x.cellg[q] = cellg[q];
    x.bg[q] = bg[q];
  }
  hipLaunchKernelGGL(k_per_bgrad, dim3(blocks_of(a.na)), dim3(kBlock), 0, s, a, nf, x);
}

void periodic_values(const PerArgs &a, hipStream_t s, int nf, const double *const *phi, double *const *bphi) {
  ValArgs x{};
  for (int q = 0; q < nf; ++q) {
    x.phi[q] = phi[q];
    x.bphi[q] = bphi[q];
  }
  hipLaunchKernelGGL(k_per_values, dim3(blocks_of(a.na)), dim3(kBlock), 0, s, a, nf, x);
}

void periodic_momentum(const PerArgs &a, hipStream_t s, int comp, double mu, double c23, const double *visc,
                       const double *bvisc, const double *phi, const double *u, const double *v, const double *w,
                       const double *bgu, const double *bgv, const double *bgw, const double *Sx,
                       const double *Sy, const double *Sz, const double *bmdot, double *cf, double *ff,
                       double *tf) {
  hipLaunchKernelGGL(k_per_momentum, dim3(blocks_of(a.np)), dim3(kBlock), 0, s, a, comp, mu, c23, visc, bvisc, phi,
                     u, v, w, bgu, bgv, bgw, Sx, Sy, Sz, bmdot, cf, ff, tf);
}

void periodic_rhie_chow(const PerArgs &a, hipStream_t s, int dec, double rho, double urf_u, const double *u,
                        const double *v, const double *w, const double *us, const double *vs, const double *ws,
                        const double *du, const double *dv, const double *dw, const double *p,
                        const double *bgp, const double *Sx, const double *Sy, const double *Sz,
                        const double *rhok, const double *bgrk, double *bmdot) {
  RcPer r{rho, urf_u, u, v, w, us, vs, ws, du, dv, dw, p, bgp, Sx, Sy, Sz, rhok, bgrk};
  hipLaunchKernelGGL(k_per_rhie_chow, dim3(blocks_of(a.na)), dim3(kBlock), 0, s, a, dec, r, bmdot);
}

void periodic_pc_boundary(const PerArgs &a, hipStream_t s, int dec, double rho, const int *pk, const double *d2u,
                          const double *d2v, const double *d2w, const double *Sx, const double *Sy,
                          const double *Sz, const double *bmdot, double *cf, double *ff, double *tf) {
  hipLaunchKernelGGL(k_per_pc_boundary, dim3(blocks_of(a.Fb)), dim3(kBlock), 0, s, a, dec, rho, pk, d2u, d2v, d2w,
                     Sx, Sy, Sz, bmdot, cf, ff, tf);
}

void periodic_correct_bmdot(const PerArgs &a, hipStream_t s, const double *d2u, const double *d2v,
                            const double *d2w, const double *gpc, const double *pc, const double *cf,
                            const double *ff, double *bu, double *bv, double *bw, double *bmdot) {
  hipLaunchKernelGGL(k_per_correct_bmdot, dim3(blocks_of(a.na)), dim3(kBlock), 0, s, a, d2u, d2v, d2w, gpc, pc,
                     cf, ff, bu, bv, bw, bmdot);
}

void periodic_beta(const PerArgs &a, hipStream_t s, int dec, double rho, double mdot_target, double relax,
                   int iters, const double tuse[3], const double *d2u, const double *d2v, const double *d2w,
                   const double *Sx, const double *Sy, const double *Sz, const int *iown, const int *inei,
                   const double *ig, const double *iex, const double *iey, const double *iez,
                   const double *idcf, double *u, double *v, double *w, double *bu, double *bv, double *bw,
                   double *bmdot, double *mdot) {
  hipLaunchKernelGGL(k_beta_sum, dim3(1), dim3(64), 0, s, a, dec, rho, mdot_target, relax, iters, d2u, d2v, d2w, Sx,
                     Sy, Sz, (const double *)bmdot);
  hipLaunchKernelGGL(k_beta_faces, dim3(blocks_of(a.na)), dim3(kBlock), 0, s, a, dec, rho, d2u, d2v, d2w, Sx, Sy,
                     Sz, bu, bv, bw, bmdot);
  const double lt = sqrt(tuse[0] * tuse[0] + tuse[1] * tuse[1] + tuse[2] * tuse[2]);
  hipLaunchKernelGGL(k_beta_cells, dim3(blocks_of(a.N)), dim3(kBlock), 0, s, a.N, (const double *)a.scal,
                     tuse[0] / lt, tuse[1] / lt, tuse[2] / lt, d2u, d2v, d2w, u, v, w);
  hipLaunchKernelGGL(k_beta_interior, dim3(blocks_of(a.Fi)), dim3(kBlock), 0, s, a.Fi, dec, rho,
                     (const double *)a.scal, iown, inei, ig, Sx, Sy, Sz, iex, iey, iez, idcf, d2u, d2v, d2w, mdot);
}

void owner_values(const PerArgs &a, hipStream_t s, const double *phi, double *bphi) {
  hipLaunchKernelGGL(k_owner_values, dim3(blocks_of(a.Fb)), dim3(kBlock), 0, s, a.Fb, a.Fi, a.own, phi, bphi);
}

void periodic_linear(const PerArgs &a, hipStream_t s, const double tuse[3]) {
  hipLaunchKernelGGL(k_per_linear, dim3(1), dim3(64), 0, s, a.scal, tuse[0], tuse[1], tuse[2]);
}

} // namespace fs
