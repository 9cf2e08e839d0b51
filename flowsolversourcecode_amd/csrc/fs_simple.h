// Internal include of the SIMPLE C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "fs_common.h"
#include "fsamd_simple.h"

namespace fs {

// Translational periodic pairs with the implicit coupling on the device.  The
// periodic boundary faces are listed in boundary order (the reference's
// Iperiodic list, SetBoundaryConditionPatches.for:1280-1286); each carries
// the geometry every periodic branch of the reference recomputes from the
// partner cell's centroid carried by the partner patch's translation
// (GFactCF, DistCF and its unit vector, the E/T split of the face area),
// formed once on the host in the reference's operation order.
struct Periodic {
  int np = 0, na = 0;                 // periodic faces; those of the lower patches
  std::vector<int> h_pair;            // partner patch per patch (-1 none)
  std::vector<double> h_trans;        // x/y/zTranslation per patch
  std::vector<int> h_b, h_part, h_j1, h_act, h_patch;
  double tuse[3] = {0.0, 0.0, 0.0};   // x/y/zTranslationUse
  bool calc_beta = false;             // LcalculateBeta
  int iters = 1, every = 1;           // BetaIterations, nIterCorrectBeta
  double mdot = 0.0, relax = 1.0;     // periodicMdot, relaxBeta
  DevBuf<int> b, part, j1, act;       // per face k: boundary index, partner's k, partner cell; act: the
                                      // lower-patch faces' k
  DevBuf<double> g, ux, uy, uz, dcf, gd, tx, ty, tz;
  DevBuf<double> scal;                // [0] periodicBeta [1] dp'/dx [2..4] beta along x, y, z
};

enum { PSCAL_BETA = 0, PSCAL_DPDX = 1, PSCAL_LIN = 2 };

// the arrays the periodic kernels read and write
struct PerArgs {
  int N, Fi, Fb, np, na;
  const int *own, *b, *part, *j1, *act;
  const double *g, *ux, *uy, *uz, *dcf, *gd, *tx, *ty, *tz;
  double *scal;
};

PerArgs periodic_args(const Periodic &P, int N, int Fi, int Fb, const int *own);
// boundary gradients at the periodic faces of up to 4 fields from the cell
// gradients (CorrectGradientAtPeriodicBoundary, CorrectGradients.for:141-167)
void periodic_bgrad(const PerArgs &a, hipStream_t s, int nf, const double *const *cellg, double *const *bg);
// boundary values at the periodic faces, lower patch interpolated, partner
// copied (UpdatePeriodic for u, v, w; UpdateBoundaryPressure and
// UpdateBoundaryPressureCorrection for p, p')
void periodic_values(const PerArgs &a, hipStream_t s, int nf, const double *const *phi, double *const *bphi);
// the momentum fluxes of component comp at the periodic faces (diffusion,
// explicit, bulk viscosity, upwind convection)
void periodic_momentum(const PerArgs &a, hipStream_t s, int comp, double mu, double c23, const double *visc,
                       const double *phi, const double *u, const double *v, const double *w,
                       const double *bgu, const double *bgv, const double *bgw, const double *Sx,
                       const double *Sy, const double *Sz, const double *bmdot, double *cf, double *ff,
                       double *tf);
// Rhie-Chow mass flux of the lower-patch faces, its negative on the partners
void periodic_rhie_chow(const PerArgs &a, hipStream_t s, int dec, double rho, double urf_u, const double *u,
                        const double *v, const double *w, const double *us, const double *vs, const double *ws,
                        const double *du, const double *dv, const double *dw, const double *p,
                        const double *bgp, const double *Sx, const double *Sy, const double *Sz,
                        double *bmdot);
// the p' face fluxes of every boundary face: walls none, periodic faces the
// D2 coefficients; FluxTf = Bmdot
void periodic_pc_boundary(const PerArgs &a, hipStream_t s, int dec, double rho, const int *pk, const double *d2u,
                          const double *d2v, const double *d2w, const double *Sx, const double *Sy,
                          const double *Sz, const double *bmdot, double *cf, double *ff, double *tf);
// bphi[b] = phi[owner] over every boundary face (the walls' BPressureC)
void owner_values(const PerArgs &a, hipStream_t s, const double *phi, double *bphi);
// CorrectBoundarymdot at the periodic faces
void periodic_correct_bmdot(const PerArgs &a, hipStream_t s, const double *d2u, const double *d2v,
                            const double *d2w, const double *gpc, const double *pc, const double *cf,
                            const double *ff, double *bu, double *bv, double *bw, double *bmdot);
// updateperiodicBeta (PeriodicBeta.for:4-386)
void periodic_beta(const PerArgs &a, hipStream_t s, int dec, double rho, double mdot_target, double relax,
                   int iters, const double tuse[3], const double *d2u, const double *d2v, const double *d2w,
                   const double *Sx, const double *Sy, const double *Sz, const int *iown, const int *inei,
                   const double *ig, const double *iex, const double *iey, const double *iez,
                   const double *idcf, double *u, double *v, double *w, double *bu, double *bv, double *bw,
                   double *bmdot, double *mdot);
// beta along x, y, z for AssembleLinearPressureTerm from periodicBeta
void periodic_linear(const PerArgs &a, hipStream_t s, const double tuse[3]);

} // namespace fs
