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
// visc/bvisc: the effective diffusion coefficient of cells and boundary faces
// (nullptr: the constant viscosity mu)
void periodic_momentum(const PerArgs &a, hipStream_t s, int comp, double mu, double c23, const double *visc,
                       const double *bvisc, const double *phi, const double *u, const double *v, const double *w,
                       const double *bgu, const double *bgv, const double *bgw, const double *Sx,
                       const double *Sy, const double *Sz, const double *bmdot, double *cf, double *ff,
                       double *tf);
// Rhie-Chow mass flux of the lower-patch faces, its negative on the partners
// rhok/bgrk: 2/3 rho k and its periodic boundary gradient (nullptr: laminar)
void periodic_rhie_chow(const PerArgs &a, hipStream_t s, int dec, double rho, double urf_u, const double *u,
                        const double *v, const double *w, const double *us, const double *vs, const double *ws,
                        const double *du, const double *dv, const double *dw, const double *p,
                        const double *bgp, const double *Sx, const double *Sy, const double *Sz,
                        const double *rhok, const double *bgrk, double *bmdot);
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

// The k-omega SST model (ModelNumber 14, low-Reynolds-number walls): the
// arrays its kernels read and write.  Cells: k, omega, mu_t, the wall
// distance, F1/F2, TurbulenceProduction, StrainRate, 2/3 rho k and the
// effective diffusion coefficient of the equation in assembly; boundary
// faces: their boundary counterparts.  Gradients are [x | y | z] blocks of N
// (cells) or Fb (boundary faces; only the periodic faces' entries are read,
// the other faces take the owner's cell gradient).  pk: per boundary face
// its periodic list index.
struct SstArgs {
  int N, Fi, Fb;
  const int *own, *nei, *btype, *pk;
  double rho, mu, urf_mut, max_ratio;
  const double *gu, *gv, *gw, *bgu, *bgv, *bgw, *mdot, *bmdot;
  double *k, *om, *mut, *wd, *bk, *bom, *bmut;
  double *gk, *gom, *bgk, *bgom, *rhok, *brhok;
  double *F1, *F2, *bF1, *bF2, *prod, *sr, *bsr, *med, *bmed;
};

void sst_strain(const SstArgs &a, hipStream_t s);
void sst_rhok(const SstArgs &a, hipStream_t s);
// var 0: k, 1: omega, 2: the velocities
void sst_ediff(const SstArgs &a, hipStream_t s, int var);
void sst_production(const SstArgs &a, hipStream_t s);
void sst_blending(const SstArgs &a, hipStream_t s);
// the face fluxes of the k (var 0) or omega (var 1) equation over every face;
// p the periodic faces (nullptr: none); framework FS_HR_NONE before
// nIterStartApplyingHR
void sst_faces(const SstArgs &a, const PerArgs *p, hipStream_t s, int var, int framework, int scheme,
               const double *g, const double *gdiff, const double *Tx, const double *Ty, const double *Tz,
               const double *dcx, const double *dcy, const double *dcz, double *cf, double *ff, double *tf);
void sst_cells(const SstArgs &a, bool per, hipStream_t s, int var, int W, const int *code, const double *cf,
               const double *ff, const double *tf, const double *vol, double urf, double *ac, double *anb,
               double *bc);
void sst_bound(const SstArgs &a, const PerArgs &p, hipStream_t s, int var);
void sst_viscosity(const SstArgs &a, hipStream_t s);
// the Poisson wall distance: face fluxes (interior faces from the cell
// gradient, wall faces phi = 0, other boundary faces none), the cell gather
// with the unit source, and the distance from the converged phi
void wall_distance_faces(hipStream_t s, int Fi, int Fb, int N, const int *own, const int *nei, const int *btype,
                         const double *g, const double *gdiff, const double *Tx, const double *Ty,
                         const double *Tz, const double *phi, const double *gr, double *bphi, double *cf,
                         double *ff, double *tf);
void wall_distance_cells(hipStream_t s, int N, int W, const int *code, const double *cf, const double *ff,
                         const double *tf, const double *vol, const double *phi, double *ac, double *anb,
                         double *bc);
void wall_distance_final(hipStream_t s, int N, int W, int Fi, const int *code, const int *btype, const double *phi,
                         const double *gr, const double *dx, const double *dy, const double *dz, const double *nx,
                         const double *ny, const double *nz, double *wd);

} // namespace fs
