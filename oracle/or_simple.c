/*
 * oracle/or_simple.c -- CPU restatement of one steady SIMPLE outer iteration
 * of the reference (MainProgram.for:257-348) for a laminar incompressible
 * constant-property flow bounded by no-slip walls and symmetry planes:
 *   SolveMomentumX/Y/Z (Momentum/SolveXMomentumEquation.for:175-299 and the
 *   Y/Z twins), SolveContinuity (ContinuityEquation/SolveContinuityEquation.for
 *   :37-114).
 * TEST INFRASTRUCTURE ONLY (see or_solve.c).
 *
 * Case options fixed by this restatement (the reference's User0 switches):
 *   MethodCalcGradient = 2 with nIterGradient = 2 (CalculateGradients.for),
 *   GradientInterpolationScheme 'average' for velocity and 'averagecorrected'
 *   for pressure (User0.F90:419-420), MethodDecomposeS = 3, Algorithm 'simple',
 *   LRelaxMomentum = .true., convection 'upwind' (no HR correction),
 *   Linviscid = .false., no turbulence/buoyancy/transient terms, constant
 *   density rho with face density rho.
 * Every loop keeps the reference's order; face fluxes are accumulated per face
 * in the order diffusion, explicit diffusion, bulk viscosity, convection, and
 * scattered to cells in face order (AssembleGolbalElementAndFaceFluxes.for).
 */
#include "or_simple.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "or_solve.h"

static double *dz(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* the REAL(4) literal -2./3. of AssembleBulkViscosity.for:49 promoted to
 * double */
static double c_m23(void) { return (double)(-2.0f / 3.0f); }

or_case *or_case_new(or_mesh *m, const double *wall_u, const double *wall_v,
                     const double *wall_w, double rho, double mu, double urf_u,
                     double urf_p, int msolver, double mrrf, int mitmax,
                     int psolver, double prrf, int pitmax) {
  or_case *c = (or_case *)calloc(1, sizeof(or_case));
  c->m = m;
  const int N = m->ncells, Fi = m->nif, Fb = m->nfaces - m->nif, W = m->w;
  c->N = N;
  c->Fi = Fi;
  c->Fb = Fb;
  c->W = W;
  c->bu0 = dz(Fb);
  c->bv0 = dz(Fb);
  c->bw0 = dz(Fb);
  memcpy(c->bu0, wall_u, sizeof(double) * Fb);
  memcpy(c->bv0, wall_v, sizeof(double) * Fb);
  memcpy(c->bw0, wall_w, sizeof(double) * Fb);
  c->rho = rho;
  c->mu = mu;
  c->urf_u = urf_u;
  c->urf_p = urf_p;
  c->msolver = msolver;
  c->mrrf = mrrf;
  c->mitmax = mitmax;
  c->psolver = psolver;
  c->prrf = prrf;
  c->pitmax = pitmax;
  c->ngrad = 2;
  c->ngrad_cont = 2;
  c->decomp = 3;
  c->gi_mom = OR_GI_AVERAGE;          /* User0.F90:419-420 */
  c->gi_cont = OR_GI_AVERAGECORRECTED;
  c->refcell = 0; /* ReferencePressureLocation = 1 (User0.F90:50) */
  c->u = dz(N); c->v = dz(N); c->w = dz(N); c->p = dz(N); c->pc = dz(N);
  c->mdot = dz(Fi);
  c->gu = dz(3 * (size_t)N); c->gv = dz(3 * (size_t)N); c->gw = dz(3 * (size_t)N);
  c->gp = dz(3 * (size_t)N); c->gpc = dz(3 * (size_t)N);
  c->bu = dz(Fb); c->bv = dz(Fb); c->bw = dz(Fb); c->bp = dz(Fb); c->bpc = dz(Fb);
  c->bgu = dz(3 * (size_t)Fb); c->bgv = dz(3 * (size_t)Fb); c->bgw = dz(3 * (size_t)Fb);
  c->bgp = dz(3 * (size_t)Fb); c->bgpc = dz(3 * (size_t)Fb);
  c->fgu = dz(3 * (size_t)Fi); c->fgv = dz(3 * (size_t)Fi); c->fgw = dz(3 * (size_t)Fi);
  c->fgp = dz(3 * (size_t)Fi);
  c->du = dz(N); c->dv = dz(N); c->dw = dz(N);
  c->us = dz(N); c->vs = dz(N); c->ws = dz(N);
  c->visc = dz(N);
  for (int i = 0; i < N; ++i) c->visc[i] = mu;
  c->fcf = dz(m->nfaces); c->fff = dz(m->nfaces); c->fvf = dz(m->nfaces); c->ftf = dz(m->nfaces);
  c->fce = dz(N); c->fte = dz(N); c->effdiv = dz(N);
  c->ac = dz(N); c->anb = dz((size_t)N * W); c->bc = dz(N); c->dphi = dz(N);
  c->bmdot = dz(Fb);
  for (int b = 0; b < Fb; ++b) {
    c->bu[b] = c->bu0[b];
    c->bv[b] = c->bv0[b];
    c->bw[b] = c->bw0[b];
  }
  return c;
}

int or_case_set_patch_types(or_case *c, int npatches, const int *patch_nfaces, const int *patch_type) {
  int total = 0;
  for (int q = 0; q < npatches; ++q) {
    if (patch_nfaces[q] < 0) return -1;
    if (patch_type[q] != OR_BC_WALL && patch_type[q] != OR_BC_SYMMETRY && patch_type[q] != OR_BC_PERIODIC)
      return -1;
    /* CorrectPressureGradientAtSymmetryBoundary reads BFaceAreanz(i,i): face
     * i of patch i */
    if (patch_type[q] == OR_BC_SYMMETRY && patch_nfaces[q] < q + 1) return -1;
    total += patch_nfaces[q];
  }
  if (total != c->Fb) return -1;
  free(c->btype);
  free(c->bzq);
  c->btype = (int *)calloc(c->Fb ? c->Fb : 1, sizeof(int));
  c->bzq = (int *)calloc(c->Fb ? c->Fb : 1, sizeof(int));
  int start = 0;
  for (int q = 0; q < npatches; ++q) {
    for (int j = 0; j < patch_nfaces[q]; ++j) {
      c->btype[start + j] = patch_type[q];
      c->bzq[start + j] = start + q;
    }
    start += patch_nfaces[q];
  }
  c->sym_dirty = 1;
  return 0;
}

/* UpdateSymmetry (SymmetryAndAxisBoundary.for:27-44): the boundary velocity
 * of a symmetry face is the owner's velocity less its normal component. */
void or_case_update_symmetry(or_case *c) {
  c->sym_dirty = 0;
  if (!c->btype) return;
  const or_mesh *m = c->m;
  for (int b = 0; b < c->Fb; ++b) {
    if (c->btype[b] != OR_BC_SYMMETRY) continue;
    const int f = c->Fi + b, i1 = m->owner[f];
    const double ax = m->nx[f], ay = m->ny[f], az = m->nz[f];
    const double amdot = c->u[i1] * ax + c->v[i1] * ay + c->w[i1] * az;
    c->bu[b] = c->u[i1] - amdot * ax;
    c->bv[b] = c->v[i1] - amdot * ay;
    c->bw[b] = c->w[i1] - amdot * az;
  }
}

/* ---------------------------------------------------------------------
 * Translational periodicity with the implicit coupling.
 */

/* The geometry the periodic branches compute for boundary face b from the
 * partner cell j1's centroid carried by the partner patch's translation
 * (e.g. UpdatePeriodic, PeriodicBoundary.for:54-67; the E/T split of
 * AssembleIsotropicDiffusion.for:663-700). */
static void periodic_geometry(or_case *c, int b, int q2, const double *trans) {
  const or_mesh *m = c->m;
  const int f = c->Fi + b, i1 = m->owner[f], j1 = c->pj1[b];
  const double xf1 = m->xc[j1] + trans[3 * q2], yf1 = m->yc[j1] + trans[3 * q2 + 1],
               zf1 = m->zc[j1] + trans[3 * q2 + 2];
  const double ax = m->fcx[f] - m->xc[i1], ay = m->fcy[f] - m->yc[i1], az = m->fcz[f] - m->zc[i1];
  const double bx = m->fcx[f] - xf1, by = m->fcy[f] - yf1, bz = m->fcz[f] - zf1;
  const double d1 = sqrt(ax * ax + ay * ay + az * az), d2 = sqrt(bx * bx + by * by + bz * bz);
  c->pg[b] = d2 / (d1 + d2);
  const double dx = xf1 - m->xc[i1], dy = yf1 - m->yc[i1], dz_ = zf1 - m->zc[i1];
  const double dcf = sqrt(dx * dx + dy * dy + dz_ * dz_);
  c->pdx[b] = dx;
  c->pdy[b] = dy;
  c->pdz[b] = dz_;
  c->pdcf[b] = dcf;
  const double ux = dx / dcf, uy = dy / dcf, uz = dz_ / dcf;
  c->pux[b] = ux;
  c->puy[b] = uy;
  c->puz[b] = uz;
  const double sx = m->Sx[f], sy = m->Sy[f], sz = m->Sz[f], A = m->A[f];
  double fex, fey, fez, fe;
  if (c->decomp == 1) {
    const double dot = ux * sx + uy * sy + uz * sz;
    fex = dot * ux; fey = dot * uy; fez = dot * uz;
    fe = fabs(dot);
  } else if (c->decomp == 2) {
    fex = A * ux; fey = A * uy; fez = A * uz;
    fe = A;
  } else {
    const double dot = ux * sx + uy * sy + uz * sz;
    fex = (A * A / dot) * ux; fey = (A * A / dot) * uy; fez = (A * A / dot) * uz;
    fe = sqrt(fex * fex + fey * fey + fez * fez);
  }
  c->pgd[b] = fe / dcf;
  c->ptx[b] = sx - fex;
  c->pty[b] = sy - fey;
  c->ptz[b] = sz - fez;
}

/* RotationMatrixCoefficients, translational branch
 * (CalculateRotationMatrixCoefficients.for:171-242): each face of the lower
 * patch of a pair takes the face of the partner patch whose centroid offset
 * is closest to the translation (largest per-axis miss, first minimum);
 * LPeriodicImplicit then puts the partner cell in the face's slot of
 * ElementNeighbor, the slot searched over NumbOfElementNodes entries of
 * NGlobalEFaces (:249-289). */
int or_case_set_periodic(or_case *c, int npatches, const int *pair, const double *trans, double beta0,
                         int calc_beta, double mdot_target, double relax, int iters, int every) {
  or_mesh *m = c->m;
  const int Fi = c->Fi, Fb = c->Fb, N = c->N;
  if (!c->btype || npatches != m->npatches) return -1;
  for (int q = 0; q < npatches; ++q) {
    const int typ = c->btype[m->patch_start[q] - Fi];
    if (m->patch_start[q + 1] == m->patch_start[q]) continue;
    if ((typ == OR_BC_PERIODIC) != (pair[q] >= 0)) return -1;
    if (pair[q] >= 0 && (pair[q] >= npatches || pair[pair[q]] != q || pair[q] == q)) return -1;
  }
  free(c->pbf); free(c->pact); free(c->pj1);
  c->pbf = (int *)malloc(sizeof(int) * (Fb ? Fb : 1));
  c->pact = (int *)calloc(Fb ? Fb : 1, sizeof(int));
  c->pj1 = (int *)malloc(sizeof(int) * (Fb ? Fb : 1));
  for (int b = 0; b < Fb; ++b) c->pbf[b] = c->pj1[b] = -1;
  double **arrs[] = {&c->pg, &c->pdx, &c->pdy, &c->pdz, &c->pdcf, &c->pux, &c->puy, &c->puz, &c->pgd,
                     &c->ptx, &c->pty, &c->ptz};
  for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); ++i) {
    free(*arrs[i]);
    *arrs[i] = dz(Fb);
  }
  c->nperiodic = 0;
  for (int q1 = 0; q1 < npatches; ++q1) {
    const int q2 = pair[q1];
    if (q2 <= q1) continue;
    for (int f1 = m->patch_start[q1]; f1 < m->patch_start[q1 + 1]; ++f1) {
      const double xf = m->fcx[f1], yf = m->fcy[f1], zf = m->fcz[f1];
      double best = 1.e30; /* big (Constants1.F90) */
      int fd = -1;
      for (int f2 = m->patch_start[q2]; f2 < m->patch_start[q2 + 1]; ++f2) {
        const double d1 = fabs(fabs(m->fcx[f2] - xf) - fabs(trans[3 * q1]));
        const double d2 = fabs(fabs(m->fcy[f2] - yf) - fabs(trans[3 * q1 + 1]));
        const double d3 = fabs(fabs(m->fcz[f2] - zf) - fabs(trans[3 * q1 + 2]));
        double dif = d1;
        if (d2 > dif) dif = d2;
        if (d3 > dif) dif = d3;
        if (dif < best) {
          best = dif;
          fd = f2;
        }
      }
      if (fd < 0) return -1;
      c->pbf[f1 - Fi] = fd - Fi;
      c->pbf[fd - Fi] = f1 - Fi;
      c->pact[f1 - Fi] = 1;
    }
  }
  for (int b = 0; b < Fb; ++b) {
    if (c->btype[b] != OR_BC_PERIODIC) continue;
    if (c->pbf[b] < 0) return -1;
    c->pj1[b] = m->owner[Fi + c->pbf[b]];
    c->nperiodic++;
  }
  /* x/y/zTranslationUse: the first periodic patch's translation */
  for (int q = 0; q < npatches; ++q)
    if (pair[q] >= 0) {
      c->tuse[0] = fabs(trans[3 * q]);
      c->tuse[1] = fabs(trans[3 * q + 1]);
      c->tuse[2] = fabs(trans[3 * q + 2]);
      break;
    }
  for (int q = 0; q < npatches; ++q) {
    if (pair[q] < 0) continue;
    for (int f = m->patch_start[q]; f < m->patch_start[q + 1]; ++f) {
      const int b = f - Fi, i1 = m->owner[f];
      periodic_geometry(c, b, pair[q], trans);
      int found = 0;
      for (int s = 0; s < m->nen[i1] && s < m->w; ++s)
        if (m->gef[(size_t)s * N + i1] == f + 1) {
          m->enbr[(size_t)s * N + i1] = c->pj1[b] + 1;
          found = 1;
          break;
        }
      if (!found) return -1;
    }
  }
  free(c->ppair);
  free(c->ptrans);
  c->ppair = (int *)malloc(sizeof(int) * npatches);
  c->ptrans = dz(3 * (size_t)npatches);
  memcpy(c->ppair, pair, sizeof(int) * npatches);
  memcpy(c->ptrans, trans, sizeof(double) * 3 * npatches);
  c->beta = beta0;
  c->beta_calc = calc_beta;
  c->beta_mdot = mdot_target;
  c->beta_relax = relax;
  c->beta_iters = iters;
  c->beta_every = every > 0 ? every : 1;
  return 0;
}

static int is_periodic(const or_case *c, int b) { return c->pbf && c->pbf[b] >= 0; }

/* g phi_C + (1 - g) phi_F of a periodic face */
static double per_mix(const or_case *c, int b, double pi, double pj) {
  return c->pg[b] * pi + (1. - c->pg[b]) * pj;
}

/* the rotation of a translational pair is the identity: the reference's
 * a1r x + b1r y + c1r z with (1, 0, 0), (0, 1, 0), (0, 0, 1) */
static double rot(int comp, double x, double y, double z) {
  const double a = comp == 0 ? 1. : 0., bb = comp == 1 ? 1. : 0., cc = comp == 2 ? 1. : 0.;
  return a * x + bb * y + cc * z;
}

/* UpdatePeriodic, translational (PeriodicBoundary.for:54-79): the boundary
 * velocities of the lower patch interpolated between the two cells, copied
 * to the partner faces. */
static void update_periodic(or_case *c) {
  if (!c->pbf) return;
  const or_mesh *m = c->m;
  for (int b = 0; b < c->Fb; ++b) {
    if (!c->pact[b]) continue;
    const int i1 = m->owner[c->Fi + b], j1 = c->pj1[b], b2 = c->pbf[b];
    c->bu[b] = per_mix(c, b, c->u[i1], c->u[j1]);
    c->bv[b] = per_mix(c, b, c->v[i1], c->v[j1]);
    c->bw[b] = per_mix(c, b, c->w[i1], c->w[j1]);
    c->bu[b2] = c->bu[b];
    c->bv[b2] = c->bv[b];
    c->bw[b2] = c->bw[b];
  }
}

/* CorrectGradientAtPeriodicBoundary, translational (CorrectGradients.for
 * :141-167): the lower patch's boundary gradients interpolated between the
 * two cells; the partner faces receive the x and y components only (the
 * z component keeps the owner's copy). */
static void periodic_gradient(or_case *c, const double *gr, double *bgr) {
  if (!c->pbf) return;
  const or_mesh *m = c->m;
  const int N = c->N, Fb = c->Fb;
  for (int b = 0; b < Fb; ++b) {
    if (!c->pact[b]) continue;
    const int i1 = m->owner[c->Fi + b], j1 = c->pj1[b], b2 = c->pbf[b];
    bgr[b] = per_mix(c, b, gr[i1], gr[j1]);
    bgr[Fb + b] = per_mix(c, b, gr[N + i1], gr[N + j1]);
    bgr[2 * Fb + b] = per_mix(c, b, gr[2 * N + i1], gr[2 * N + j1]);
    bgr[b2] = bgr[b];
    bgr[Fb + b2] = bgr[Fb + b];
  }
}

/* a periodic face's cell-to-cell value on the lower patch, copied to the
 * partner (UpdateBoundaryPressure :293-317, UpdateBoundaryPressureCorrection
 * :263-290 for p and p') */
static void periodic_values(or_case *c, const double *phi, double *bphi) {
  if (!c->pbf) return;
  const or_mesh *m = c->m;
  for (int b = 0; b < c->Fb; ++b) {
    if (!c->pact[b]) continue;
    const int i1 = m->owner[c->Fi + b], j1 = c->pj1[b];
    bphi[b] = per_mix(c, b, phi[i1], phi[j1]);
    bphi[c->pbf[b]] = bphi[b];
  }
}

void or_case_delete(or_case *c) {
  free(c->du2);
  free(c->dv2);
  free(c->dw2);
  free(c->btype);
  free(c->bzq);
  free(c->pbf);
  free(c->pact);
  free(c->pj1);
  free(c->ppair);
  free(c->ptrans);
  double *parr[] = {c->pg, c->pdx, c->pdy, c->pdz, c->pdcf, c->pux, c->puy, c->puz, c->pgd, c->ptx, c->pty,
                    c->ptz, c->bmdot};
  for (size_t i = 0; i < sizeof(parr) / sizeof(parr[0]); ++i) free(parr[i]);
  double *arrs[] = {c->bu0, c->bv0, c->bw0, c->u, c->v, c->w, c->p, c->pc, c->mdot, c->gu,
                    c->gv, c->gw, c->gp, c->gpc, c->bu, c->bv, c->bw, c->bp, c->bpc, c->bgu,
                    c->bgv, c->bgw, c->bgp, c->bgpc, c->fgu, c->fgv, c->fgw, c->fgp, c->du,
                    c->dv, c->dw, c->us, c->vs, c->ws, c->visc, c->fcf, c->fff, c->fvf, c->ftf,
                    c->fce, c->fte, c->effdiv, c->ac, c->anb, c->bc, c->dphi};
  for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); ++i) free(arrs[i]);
  double *tarr[] = {c->tk, c->tw, c->mut, c->wd, c->btk, c->btw, c->bmut, c->gk, c->gom, c->bgk, c->bgom,
                    c->fgk, c->fgom, c->rhok, c->brhok, c->grk, c->bgrk, c->fgrkb, c->fgrkc, c->F1, c->bF1,
                    c->F2, c->bF2, c->prod, c->sr, c->bsr, c->med, c->bmed};
  for (size_t i = 0; i < sizeof(tarr) / sizeof(tarr[0]); ++i) free(tarr[i]);
  free(c);
}

/* StandardGreenGaussFaceCorrection1 (StandardGreenGaussCorrection1Gradient.for
 * :35-186): Green-Gauss with the face value re-evaluated at the face
 * centroid, nIterGradient correction sweeps, boundary copies of the owner's
 * gradient.  gr = [gx | gy | gz] (3N), bgr = [bgx | bgy | bgz] (3Fb). */
static void gradient_n(or_case *c, const double *phi, const double *bphi, double *gr, double *bgr, int ngrad);

void or_gradient(or_case *c, const double *phi, const double *bphi, double *gr, double *bgr) {
  gradient_n(c, phi, bphi, gr, bgr, c->ngrad);
}

/* MethodCalcGradient 1 is the first pass alone (StandardGreenGaussGradient.for
 * :35-96 forms the same sums): ngrad = 0 */
static void gradient_n(or_case *c, const double *phi, const double *bphi, double *gr, double *bgr, int ngrad) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb;
  double *gx = gr, *gy = gr + N, *gz = gr + 2 * N;
  double *fp = dz(Fi), *fx = dz(N), *fy = dz(N), *fz = dz(N);
  memset(gr, 0, sizeof(double) * 3 * N);
  for (int k = 0; k < Fi; ++k) {
    const int j = m->owner[k], kk = m->neigh[k];
    fp[k] = phi[j] * m->g[k] + phi[kk] * (1. - m->g[k]);
    const double sx = fp[k] * m->Sx[k], sy = fp[k] * m->Sy[k], sz = fp[k] * m->Sz[k];
    gx[j] = gx[j] + sx; gy[j] = gy[j] + sy; gz[j] = gz[j] + sz;
    gx[kk] = gx[kk] - sx; gy[kk] = gy[kk] - sy; gz[kk] = gz[kk] - sz;
  }
  for (int b = 0; b < Fb; ++b) {
    const int f = Fi + b, k = m->owner[f];
    const double sx = bphi[b] * m->Sx[f], sy = bphi[b] * m->Sy[f], sz = bphi[b] * m->Sz[f];
    gx[k] = gx[k] + sx; gy[k] = gy[k] + sy; gz[k] = gz[k] + sz;
  }
  for (int i = 0; i < N; ++i) {
    gx[i] = gx[i] / m->vol[i]; gy[i] = gy[i] / m->vol[i]; gz[i] = gz[i] / m->vol[i];
  }
  for (int it = 0; it < ngrad; ++it) {
    memset(fx, 0, sizeof(double) * N);
    memset(fy, 0, sizeof(double) * N);
    memset(fz, 0, sizeof(double) * N);
    for (int k = 0; k < Fi; ++k) {
      const int j = m->owner[k], kk = m->neigh[k];
      const double g = m->g[k];
      const double edn = m->Ex[k] * m->nx[k] + m->Ey[k] * m->ny[k] + m->Ez[k] * m->nz[k];
      const double rdn = (m->fcx[k] - m->xc[j]) * m->nx[k] + (m->fcy[k] - m->yc[j]) * m->ny[k] +
                         (m->fcz[k] - m->zc[j]) * m->nz[k];
      const double xfp = m->xc[j] + (rdn / edn) * m->Ex[k];
      const double yfp = m->yc[j] + (rdn / edn) * m->Ey[k];
      const double zfp = m->zc[j] + (rdn / edn) * m->Ez[k];
      double cx = gx[j] * g + gx[kk] * (1. - g);
      cx = cx * (m->fcx[k] - xfp);
      double cy = gy[j] * g + gy[kk] * (1. - g);
      cy = cy * (m->fcy[k] - yfp);
      double cz = gz[j] * g + gz[kk] * (1. - g);
      cz = cz * (m->fcz[k] - zfp);
      const double phif = fp[k] + cx + cy + cz;
      const double sx = phif * m->Sx[k], sy = phif * m->Sy[k], sz = phif * m->Sz[k];
      fx[j] = fx[j] + sx; fy[j] = fy[j] + sy; fz[j] = fz[j] + sz;
      fx[kk] = fx[kk] - sx; fy[kk] = fy[kk] - sy; fz[kk] = fz[kk] - sz;
    }
    for (int b = 0; b < Fb; ++b) {
      const int f = Fi + b, k = m->owner[f];
      const double sx = bphi[b] * m->Sx[f], sy = bphi[b] * m->Sy[f], sz = bphi[b] * m->Sz[f];
      fx[k] = fx[k] + sx; fy[k] = fy[k] + sy; fz[k] = fz[k] + sz;
    }
    for (int i = 0; i < N; ++i) {
      gx[i] = fx[i] / m->vol[i]; gy[i] = fy[i] / m->vol[i]; gz[i] = fz[i] / m->vol[i];
    }
  }
  for (int b = 0; b < Fb; ++b) {
    const int k = m->owner[Fi + b];
    bgr[b] = gx[k];
    bgr[Fb + b] = gy[k];
    bgr[2 * Fb + b] = gz[k];
  }
  free(fp);
  free(fx);
  free(fy);
  free(fz);
}

/* InterpolateGradientToFace (InterpolateGradientFromElementToFace.for:16-108):
 * 'average', 'upwind' and 'downwind' (the weight follows mdot > 0),
 * 'averagecorrected' (the average plus the centroid-line correction) and
 * 'harmonic' (density-weighted; the density is constant here). */
static void face_gradient(or_case *c, const double *phi, const double *gr, double *fg, int scheme) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi;
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    double gf = m->g[k];
    if (scheme == OR_GI_UPWIND) gf = c->mdot[k] > 0. ? 1. : 0.;
    if (scheme == OR_GI_DOWNWIND) gf = c->mdot[k] > 0. ? 0. : 1.;
    double fx, fy, fz;
    if (scheme == OR_GI_HARMONIC) {
      const double r = c->rho; /* Density(i) = Density(j) = Densityf(k) */
      fx = r * (gf * gr[i] / r + (1. - gf) * gr[j] / r);
      fy = r * (gf * gr[N + i] / r + (1. - gf) * gr[N + j] / r);
      fz = r * (gf * gr[2 * N + i] / r + (1. - gf) * gr[2 * N + j] / r);
    } else {
      fx = gf * gr[i] + (1. - gf) * gr[j];
      fy = gf * gr[N + i] + (1. - gf) * gr[N + j];
      fz = gf * gr[2 * N + i] + (1. - gf) * gr[2 * N + j];
    }
    if (scheme == OR_GI_AVERAGECORRECTED) {
      double corr = (phi[j] - phi[i]) / m->dcf[k];
      corr = corr - (fx * m->ex[k] + fy * m->ey[k] + fz * m->ez[k]);
      fx = fx + corr * m->ex[k];
      fy = fy + corr * m->ey[k];
      fz = fz + corr * m->ez[k];
    }
    fg[k] = fx;
    fg[Fi + k] = fy;
    fg[2 * Fi + k] = fz;
  }
}

int or_case_set_gradient_interpolation(or_case *c, int momentum, int continuity) {
  if (momentum < 0 || momentum > 4 || continuity < 0 || continuity > 4) return -1;
  c->gi_mom = momentum;
  c->gi_cont = continuity;
  return 0;
}

/* MIN and MAX as the reference build evaluates them: left to right, a NaN
 * operand after the first is passed over. */
static double mn2(double a, double b) { return (a < b || b != b) ? a : b; }
static double mx2(double a, double b) { return (a > b || b != b) ? a : b; }

/* TVDLimiter (LimitersTVDType.for:4-115): psi(r) of the TVD form. */
double or_tvd_psi(int scheme, double r) {
  switch (scheme) {
  case OR_CONV_UPWIND: return 0.;
  case OR_CONV_DOWNWIND: return 2.;
  case OR_CONV_FROMM: return 0.5 * (1. + r);
  case OR_CONV_SOU: return r;
  case OR_CONV_CD: return 1.;
  case OR_CONV_QUICK: return (3. + r) / 4.;
  case OR_CONV_SUPERBEE: return mx2(mx2(0., mn2(1., 2. * r)), mn2(2., r));
  case OR_CONV_MINMOD: return mx2(0., mn2(1., r));
  case OR_CONV_OSHER: return mx2(0., mn2(2., r));
  case OR_CONV_VANLEER: return (r + fabs(r)) / (1. + fabs(r));
  case OR_CONV_MUSCL: return mx2(0., mn2(mn2(2. * r, 0.5 * (1. + r)), 2.));
  case OR_CONV_SMART: return mx2(0., mn2(mn2(2. * r, 0.75 * r + 0.25), 4.));
  case OR_CONV_HQUICK: return 2. * (r + fabs(r)) / (r + 3.);
  case OR_CONV_UMIST: return mx2(0., mn2(mn2(mn2(2. * r, 0.25 + 0.75 * r), 0.75 + 0.25 * r), 2.));
  case OR_CONV_CHARM: return r <= 0. ? 0. : r * (3. * r + 1.) / ((r + 1.) * (r + 1.));
  default: return NAN;
  }
}

/* NVFLimiter (LimitersNVFType.for:4-112): the normalised face value of the
 * NVF form; BettaConvection is the REAL(4) literal 0.1 (User0.F90:408). */
double or_nvf_phif(int scheme, double t) {
  const int in01 = t >= 0. && t <= 1.;
  switch (scheme) {
  case OR_CONV_UPWIND: return t;
  case OR_CONV_SMART: return in01 ? mn2(mn2(3. * t, 0.375 + 0.75 * t), (2. + t) / 3.) : t;
  case OR_CONV_MINMOD: return in01 ? mn2(1.5 * t, 0.5 * t + 0.5) : t;
  case OR_CONV_BOUNDEDCD: return in01 ? 0.5 * t + 0.5 : t;
  case OR_CONV_CD: return 0.5 * t + 0.5;
  case OR_CONV_OSHER: return in01 ? mn2(1.5 * t, 1.) : t;
  case OR_CONV_STOIC: return in01 ? mn2(mn2(0.5 * t + 0.5, 0.375 + 0.75 * t), (2. + t) / 3.) : t;
  case OR_CONV_MUSCL: return in01 ? mn2(mn2(2. * t, 0.25 + t), 1.) : t;
  case OR_CONV_SUPERBEE: return in01 ? mn2(mn2(mn2(2. * t, 0.5 * t + 0.5), 1.5 * t), 1.) : t;
  case OR_CONV_DOWNWIND: return 1.;
  case OR_CONV_GAMMA: {
    const double beta = (double)0.1f;
    if (t > 0. && t < beta) {
      const double factor = 0.5 / beta;
      return (-factor * t + (1. + factor)) * t;
    }
    if (t >= beta && t < 1.) return 0.5 + 0.5 * t;
    return t;
  }
  default: return NAN;
  }
}

static int tvd_defines(int s) { return s >= OR_CONV_UPWIND && s <= OR_CONV_CHARM; }
static int nvf_defines(int s) {
  switch (s) {
  case OR_CONV_UPWIND: case OR_CONV_SMART: case OR_CONV_MINMOD: case OR_CONV_BOUNDEDCD: case OR_CONV_CD:
  case OR_CONV_OSHER: case OR_CONV_STOIC: case OR_CONV_MUSCL: case OR_CONV_SUPERBEE: case OR_CONV_DOWNWIND:
  case OR_CONV_GAMMA: return 1;
  default: return 0;
  }
}

int or_case_set_convection(or_case *c, int framework, int scheme, double bleed, int start_iter) {
  if (framework == OR_HR_TVD ? !tvd_defines(scheme) : framework == OR_HR_NVF ? !nvf_defines(scheme)
                                                                               : framework != OR_HR_NONE)
    return -1;
  c->hr_fw = framework;
  c->hr_scheme = scheme;
  c->bleed = bleed;
  c->hr_start = start_iter;
  return 0;
}

int or_case_set_gradient_limiters(or_case *c, int momentum_method, int continuity_method) {
  if (momentum_method < 0 || momentum_method > 2 || continuity_method < 0 || continuity_method > 2) return -1;
  c->lim_mom = momentum_method;
  c->lim_cont = continuity_method;
  return 0;
}

/* LimitGradientMethod1 (LimitGradients.for:68-262, rk = 1) and
 * LimitGradientMethod2 (:263-437 with PhiFieldMinMax :440-476): one factor
 * per cell from the min/max of the cell, its neighbours and its boundary
 * values and the gradient's extrapolation to each face centroid; the
 * boundary gradients become the owners' limited gradients. */
void or_limit_gradient(or_case *c, int method, const double *phi, const double *bphi, double *gr, double *bgr) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb, W = c->W;
  const double big = 1.e30, tiny = 1.e-15;
  double *lo = dz(N), *hi = dz(N), *lim = dz(N);
  double eps2 = 0.;
  if (method == 2) {
    double pmin = big, pmax = -big;
    for (int i = 0; i < N; ++i) { pmin = mn2(pmin, phi[i]); pmax = mx2(pmax, phi[i]); }
    for (int b = 0; b < Fb; ++b) { pmin = mn2(pmin, bphi[b]); pmax = mx2(pmax, bphi[b]); }
    const double e = 1.e-3 * (pmax - pmin);
    eps2 = e * e;
  }
  for (int i = 0; i < N; ++i) {
    lo[i] = mn2(big, phi[i]);
    hi[i] = mx2(-big, phi[i]);
    lim[i] = method == 2 ? (double)1.e30f : 1.;
  }
  for (int i = 0; i < N; ++i)
    for (int s = 0; s < m->nef[i]; ++s) {
      const int ke = m->enbr[(size_t)s * N + i];
      if (ke != 0) { lo[i] = mn2(lo[i], phi[ke - 1]); hi[i] = mx2(hi[i], phi[ke - 1]); }
    }
  for (int b = 0; b < Fb; ++b) {
    const int k = m->owner[Fi + b];
    lo[k] = mn2(lo[k], bphi[b]);
    hi[k] = mx2(hi[k], bphi[b]);
  }
  /* one face's term: the face centroid f seen from cell i, with base the
   * value the method-2 increment starts from */
  for (int pass = 0; pass < 2; ++pass) {
    const int count = pass == 0 ? N : Fb;
    for (int q = 0; q < count; ++q) {
      const int ns = pass == 0 ? m->nef[q] : 1;
      for (int s = 0; s < ns; ++s) {
        int i, f;
        double base;
        if (pass == 0) {
          if (m->enbr[(size_t)s * N + q] == 0) continue;
          i = q;
          f = m->gef[(size_t)s * N + q] - 1;
          base = phi[i];
        } else {
          f = Fi + q;
          i = m->owner[f];
          base = bphi[q];
        }
        const double gx = gr[i], gy = gr[N + i], gz = gr[2 * N + i];
        const double d = gx * (m->fcx[f] - m->xc[i]) + gy * (m->fcy[f] - m->yc[i]) + gz * (m->fcz[f] - m->zc[i]);
        if (method == 1) {
          const double dmax = hi[i] - phi[i], dmin = lo[i] - phi[i];
          if (d > dmax + tiny) lim[i] = mn2(lim[i], dmax / (d + tiny));
          else if (d < dmin - tiny) lim[i] = mn2(lim[i], dmin / (d + tiny));
        } else {
          const double dp = d >= 0. ? hi[i] - base : lo[i] - base;
          const double dp2 = dp * dp, dm2 = d * d;
          const double part1 = (dp2 + eps2) * d + 2. * dm2 * dp;
          const double part2 = dp2 + 2. * dm2 + dp * d + eps2;
          const double r = part1 / (copysign(1., d) * (fabs(d) + tiny) * part2);
          lim[i] = mn2(mn2(1., r), lim[i]);
        }
      }
    }
  }
  (void)W;
  for (int i = 0; i < N; ++i) {
    gr[i] = lim[i] * gr[i];
    gr[N + i] = lim[i] * gr[N + i];
    gr[2 * N + i] = lim[i] * gr[2 * N + i];
  }
  for (int b = 0; b < Fb; ++b) {
    const int k = m->owner[Fi + b];
    bgr[b] = gr[k];
    bgr[Fb + b] = gr[N + k];
    bgr[2 * Fb + b] = gr[2 * N + k];
  }
  free(lo);
  free(hi);
  free(lim);
}

/* HRSchemesCorrections (AssembleHRSchemeCorrection.for:31-133) for a velocity
 * component: the deferred correction mdot (phi_HR - phi_C) of every interior
 * face added to FluxTf, phi_U extrapolated from the upwind cell's gradient. */
static void hr_correction(or_case *c, int fw, int scheme, double bleed, const double *phi, const double *gr,
                          double *ftf);
void or_hr_correction(or_case *c, const double *phi, const double *gr, double *ftf) {
  hr_correction(c, c->hr_fw, c->hr_scheme, c->bleed, phi, gr, ftf);
}

/* the correction with a variable's own framework, scheme and bleed */
static void hr_correction(or_case *c, int fw, int scheme, double bleed, const double *phi, const double *gr,
                          double *ftf) {
  if (fw == OR_HR_NONE || scheme == OR_CONV_UPWIND) return;
  const or_mesh *m = c->m;
  const int N = c->N;
  const double small = (double)1.e-9f;
  for (int k = 0; k < c->Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    int iu;
    double phic, phid, dcx, dcy, dcz;
    if (c->mdot[k] > 0.) {
      iu = i; phic = phi[i]; phid = phi[j];
      dcx = m->dx[k]; dcy = m->dy[k]; dcz = m->dz[k];
    } else {
      iu = j; phic = phi[j]; phid = phi[i];
      dcx = -m->dx[k]; dcy = -m->dy[k]; dcz = -m->dz[k];
    }
    const double phiu = phid - 2. * (gr[iu] * dcx + gr[N + iu] * dcy + gr[2 * N + iu] * dcz);
    double phihr;
    if (fw == OR_HR_NVF) {
      const double diff = phid - phiu;
      const double t = fabs(diff) < small ? 0. : (phic - phiu) / diff;
      phihr = or_nvf_phif(scheme, t) * diff + phiu;
      if (scheme != OR_CONV_GAMMA) phihr = (1. - bleed) * phihr + bleed * phic;
    } else {
      const double diff = phid - phic;
      const double r = fabs(diff) < small ? 0. : (phic - phiu) / diff;
      phihr = phic + 0.5 * or_tvd_psi(scheme, r) * diff;
      phihr = (1. - bleed) * phihr + bleed * phic;
    }
    ftf[k] = ftf[k] + c->mdot[k] * (phihr - phic) * 1.;
  }
}

/* The boundary gradients of p (and of 2/3 rho k, CalculateGradients.for
 * :264-269) at symmetry planes and walls: CorrectPressureGradientAtSymmetry
 * Boundary (CorrectGradients.for:270-281, its operand order and its z-normal
 * of face (i,i)) and CorrectPresureGradientAtWalls, the tangential part of
 * the owner's gradient. */
static void pressure_boundary_gradient(or_case *c, const double *gr, double *bgr) {
  const or_mesh *m = c->m;
  const int Fi = c->Fi, Fb = c->Fb, N = c->N;
  for (int b = 0; b < Fb; ++b) {
    if (is_periodic(c, b)) continue;
    const int f = Fi + b, i1 = m->owner[f];
    const double nx = m->nx[f], ny = m->ny[f], nz = m->nz[f];
    const double gx = gr[i1], gy = gr[N + i1], gz = gr[2 * N + i1];
    if (c->btype && c->btype[b] == OR_BC_SYMMETRY) {
      const double zq = m->nz[Fi + c->bzq[b]];
      bgr[b] = gx * (1. - nx * nx) - gy * nx * ny - gz * nx * zq;
      bgr[Fb + b] = gy * (1. - ny * ny) - gx * ny * nx - gz * ny * zq;
      bgr[2 * Fb + b] = gz * (1. - zq * zq) - gx * zq * nx - gy * zq * ny;
      continue;
    }
    bgr[b] = (1 - nx * nx) * gx - nx * ny * gy - nx * nz * gz;
    bgr[Fb + b] = (1 - ny * ny) * gy - ny * nx * gx - ny * nz * gz;
    bgr[2 * Fb + b] = (1 - nz * nz) * gz - nz * nx * gx - nz * ny * gy;
  }
}

/* The gradient block at the head of SolveMomentumX (:182-222), with the
 * tangential projection of the boundary pressure gradient at walls
 * (CorrectGradients.for, CorrectPresureGradientAtWalls). */
void or_case_momentum_gradients(or_case *c) {
  or_gradient(c, c->u, c->bu, c->gu, c->bgu);
  or_gradient(c, c->v, c->bv, c->gv, c->bgv);
  or_gradient(c, c->w, c->bw, c->gw, c->bgw);
  gradient_n(c, c->p, c->bp, c->gp, c->bgp, c->ngrad_cont);
  /* LimitGradientVariable after the Green-Gauss sweeps, before the boundary
   * corrections (CalculateGradients.for:252-272) */
  if (c->lim_mom) {
    or_limit_gradient(c, c->lim_mom, c->u, c->bu, c->gu, c->bgu);
    or_limit_gradient(c, c->lim_mom, c->v, c->bv, c->gv, c->bgv);
    or_limit_gradient(c, c->lim_mom, c->w, c->bw, c->gw, c->bgw);
  }
  if (c->lim_cont) or_limit_gradient(c, c->lim_cont, c->p, c->bp, c->gp, c->bgp);
  /* CorrectGradientAtPeriodicBoundary (CalculateGradients.for:263), before
   * the pressure's symmetry and wall projections */
  periodic_gradient(c, c->gu, c->bgu);
  periodic_gradient(c, c->gv, c->bgv);
  periodic_gradient(c, c->gw, c->bgw);
  periodic_gradient(c, c->gp, c->bgp);
  pressure_boundary_gradient(c, c->gp, c->bgp);
  face_gradient(c, c->u, c->gu, c->fgu, c->gi_mom);
  face_gradient(c, c->v, c->gv, c->fgv, c->gi_mom);
  face_gradient(c, c->w, c->gw, c->fgw, c->gi_mom);
  face_gradient(c, c->p, c->gp, c->fgp, c->gi_cont);
  if (c->turb) {
    const or_mesh *m = c->m;
    const int N = c->N, Fb = c->Fb;
    /* CalculateEffectiveDiffusionCoefficient, velocities (Calculate
     * EffectiveDiffusionCoefficient: Viscosity + TurbulentViscosity) */
    for (int i = 0; i < N; ++i) c->med[i] = c->visc[i] + c->mut[i];
    for (int b = 0; b < Fb; ++b) c->bmed[b] = c->mu + c->bmut[b];
    /* MinusTwoThirdsRhoK('velx') (MinusTwoThirdsRhoK: rhok, its gradient
     * with the momentum's gradient settings and the pressure's boundary
     * projections); its face interpolations 'average' and
     * 'averagecorrected' (AssembleMdotUsingRhieChow.for:294-306) */
    const double t23 = (double)0.666666666666666666666667f; /* Constants1 TwoThird, a REAL(4) literal */
    for (int i = 0; i < N; ++i) c->rhok[i] = t23 * c->rho * c->tk[i];
    for (int b = 0; b < Fb; ++b) c->brhok[b] = t23 * c->rho * c->btk[b];
    gradient_n(c, c->rhok, c->brhok, c->grk, c->bgrk, c->ngrad);
    if (c->lim_mom) or_limit_gradient(c, c->lim_mom, c->rhok, c->brhok, c->grk, c->bgrk);
    periodic_gradient(c, c->grk, c->bgrk);
    pressure_boundary_gradient(c, c->grk, c->bgrk);
    face_gradient(c, c->rhok, c->grk, c->fgrkb, OR_GI_AVERAGE);
    face_gradient(c, c->rhok, c->grk, c->fgrkc, OR_GI_AVERAGECORRECTED);
    (void)m;
  }
}

/* AssembleGlobalMatrixFaceFluxes + ElementFluxes
 * (AssembleGolbalElementAndFaceFluxes.for:4-125) on zeroed ac/anb/bc. */
static void scatter(or_case *c) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi;
  memset(c->ac, 0, sizeof(double) * N);
  memset(c->bc, 0, sizeof(double) * N);
  memset(c->anb, 0, sizeof(double) * (size_t)N * c->W);
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const int i1 = m->iown[k] - 1, j1 = m->inbr[k] - 1;
    c->ac[i] = c->ac[i] + c->fcf[k];
    c->anb[(size_t)i1 * N + i] = c->anb[(size_t)i1 * N + i] + c->fff[k];
    c->bc[i] = c->bc[i] - c->ftf[k];
    c->ac[j] = c->ac[j] - c->fff[k];
    c->anb[(size_t)j1 * N + j] = c->anb[(size_t)j1 * N + j] - c->fcf[k];
    c->bc[j] = c->bc[j] + c->ftf[k];
  }
  for (int f = Fi; f < m->nfaces; ++f) {
    const int k = m->owner[f];
    c->ac[k] = c->ac[k] + c->fcf[f];
    c->bc[k] = c->bc[k] - c->ftf[f];
  }
  /* LPeriodicImplicit: FluxFf of a periodic face into its slot (:99-121) */
  if (c->pbf)
    for (int b = 0; b < c->Fb; ++b) {
      if (!is_periodic(c, b)) continue;
      const int f = Fi + b, i1 = m->owner[f];
      for (int s = 0; s < m->nen[i1] && s < m->w; ++s)
        if (m->gef[(size_t)s * N + i1] == f + 1) {
          c->anb[(size_t)s * N + i1] = c->anb[(size_t)s * N + i1] + c->fff[f];
          break;
        }
    }
  for (int i = 0; i < N; ++i) {
    c->ac[i] = c->ac[i] + c->fce[i];
    c->bc[i] = c->bc[i] - c->fte[i];
  }
}

/* Momentum component comp (0,1,2) up to the linear solve: diffusion
 * (AssembleIsotropicDiffusion.for:99-117, walls :148-219), explicit and
 * bulk-viscosity terms (AssembleExplicitDiffusion.for, AssembleBulkViscosity
 * .for), upwind convection and effective divergence (AssembleConvection.for
 * :72-178), pressure gradient (AssemblePressureGradient.for:18-22), scatter,
 * implicit under-relaxation (UnderRelaxation.for) and D = V/a_P
 * (StoreTheDValues.for, 'simple'). */
void or_case_momentum_assemble(or_case *c, int comp) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb, nf = m->nfaces;
  double *phi = comp == 0 ? c->u : comp == 1 ? c->v : c->w;
  double *bphi = comp == 0 ? c->bu : comp == 1 ? c->bv : c->bw;
  double *fg = comp == 0 ? c->fgu : comp == 1 ? c->fgv : c->fgw;
  double *d = comp == 0 ? c->du : comp == 1 ? c->dv : c->dw;
  double *st = comp == 0 ? c->us : comp == 1 ? c->vs : c->ws;
  const double *S = comp == 0 ? m->Sx : comp == 1 ? m->Sy : m->Sz;
  memset(c->fcf, 0, sizeof(double) * nf);
  memset(c->fff, 0, sizeof(double) * nf);
  memset(c->fvf, 0, sizeof(double) * nf);
  memset(c->ftf, 0, sizeof(double) * nf);
  memset(c->fce, 0, sizeof(double) * N);
  memset(c->fte, 0, sizeof(double) * N);
  /* eDiffCoefficient: the viscosity, or with the SST model Viscosity +
   * TurbulentViscosity in cells and on boundary faces */
  const double *gam = c->turb ? c->med : c->visc;
  /* isotropic diffusion, interior */
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gf = m->g[k];
    const double gamf = gf * gam[i] + (1. - gf) * gam[j];
    const double cfl = gamf * m->gdiff[k], ffl = -(gamf * m->gdiff[k]);
    const double vfl = -(gamf * (fg[k] * m->Tx[k] + fg[Fi + k] * m->Ty[k] + fg[2 * Fi + k] * m->Tz[k]));
    c->fcf[k] = c->fcf[k] + cfl;
    c->fff[k] = c->fff[k] + ffl;
    c->fvf[k] = c->fvf[k] + vfl;
    c->ftf[k] = c->ftf[k] + cfl * phi[i] + ffl * phi[j] + vfl;
  }
  /* no-slip walls, symmetry planes, periodic faces */
  for (int b = 0; b < Fb; ++b) {
    const int f = Fi + b, i1 = m->owner[f];
    if (is_periodic(c, b)) {
      /* AssembleIsotropicDiffusion.for:630-700,746-770 (translational,
       * implicit): BGamaFace interpolated between the two cells
       * (InterpolateGamaFromCentroidToFace.for:129-178) */
      const int j1 = c->pj1[b];
      const double gamf = per_mix(c, b, gam[i1], gam[j1]);
      const double *bg = comp == 0 ? c->bgu : comp == 1 ? c->bgv : c->bgw;
      const double cfl = gamf * c->pgd[b], ffl = -(gamf * c->pgd[b]);
      const double vfl = -(gamf * (bg[b] * c->ptx[b] + bg[Fb + b] * c->pty[b] + bg[2 * Fb + b] * c->ptz[b]));
      c->fcf[f] = c->fcf[f] + cfl;
      c->fff[f] = c->fff[f] + ffl;
      c->fvf[f] = c->fvf[f] + vfl;
      c->ftf[f] = c->ftf[f] + cfl * phi[i1] + ffl * phi[j1] + vfl;
      continue;
    }
    const double gamf = c->turb ? c->bmed[b] : c->mu; /* BGamaFace = BeDiffCoefficient */
    const double nx = m->nx[f], ny = m->ny[f], nz = m->nz[f], area = m->A[f];
    const double dn = m->dx[f] * nx + m->dy[f] * ny + m->dz[f] * nz;
    double cfl, vfl;
    if (c->btype && c->btype[b] == OR_BC_SYMMETRY) {
      /* symmetry faces (AssembleIsotropicDiffusion.for:562-620) */
      if (comp == 0) {
        cfl = 2. * gamf * area * (nx * nx) / dn;
        vfl = -((-(2. * gamf * area / dn)) * (c->v[i1] * ny + c->w[i1] * nz) * nx);
      } else if (comp == 1) {
        cfl = 2. * gamf * area * (ny * ny) / dn;
        vfl = -((-(2. * gamf * area / dn)) * (c->u[i1] * nx + c->w[i1] * nz) * ny);
      } else {
        cfl = 2. * gamf * area * (nz * nz) / dn;
        vfl = -((-(2. * gamf * area / dn)) * (c->u[i1] * nx + c->v[i1] * ny) * nz);
      }
    } else if (comp == 0) {
      cfl = gamf * area * (1. - nx * nx) / dn;
      vfl = -((gamf * area / dn) * (bphi[b] * (1. - nx * nx) + (c->v[i1] - c->bv[b]) * nx * ny +
                                     (c->w[i1] - c->bw[b]) * nx * nz));
    } else if (comp == 1) {
      cfl = gamf * area * (1. - ny * ny) / dn;
      vfl = -((gamf * area / dn) * ((c->u[i1] - c->bu[b]) * nx * ny + bphi[b] * (1. - ny * ny) +
                                     (c->w[i1] - c->bw[b]) * nz * ny));
    } else {
      cfl = gamf * area * (1. - nz * nz) / dn;
      vfl = -((gamf * area / dn) * ((c->u[i1] - c->bu[b]) * nx * nz + (c->v[i1] - c->bv[b]) * ny * nz +
                                     bphi[b] * (1. - nz * nz)));
    }
    const double ffl = 0.;
    c->fcf[f] = c->fcf[f] + cfl;
    c->fff[f] = c->fff[f] + ffl;
    c->fvf[f] = c->fvf[f] + vfl;
    c->ftf[f] = c->ftf[f] + cfl * phi[i1] + ffl * bphi[b] + vfl;
  }
  /* explicit diffusion: -Gamma_f (grad u_c)^T . S */
  const double *ga = comp == 0 ? c->fgu : comp == 1 ? c->fgu + Fi : c->fgu + 2 * Fi;
  const double *gb = comp == 0 ? c->fgv : comp == 1 ? c->fgv + Fi : c->fgv + 2 * Fi;
  const double *gc = comp == 0 ? c->fgw : comp == 1 ? c->fgw + Fi : c->fgw + 2 * Fi;
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gamf = m->g[k] * gam[i] + (1. - m->g[k]) * gam[j];
    const double vfl = -(gamf * (ga[k] * m->Sx[k] + gb[k] * m->Sy[k] + gc[k] * m->Sz[k]));
    c->fcf[k] = c->fcf[k] + 0.;
    c->fff[k] = c->fff[k] + 0.;
    c->fvf[k] = c->fvf[k] + vfl;
    c->ftf[k] = c->ftf[k] + vfl;
  }
  const double *bga = c->bgu + comp * Fb, *bgb = c->bgv + comp * Fb, *bgc = c->bgw + comp * Fb;
  for (int b = 0; b < Fb; ++b) {
    const int f = Fi + b;
    const double gamf = c->turb ? c->bmed[b] : c->mu;
    const double vfl = -(gamf * (bga[b] * m->Sx[f] + bgb[b] * m->Sy[f] + bgc[b] * m->Sz[f]));
    c->fcf[f] = c->fcf[f] + 0.;
    c->fff[f] = c->fff[f] + 0.;
    c->fvf[f] = c->fvf[f] + vfl;
    c->ftf[f] = c->ftf[f] + vfl;
  }
  /* bulk viscosity: -(-2/3) Gamma_f (div u)_f S_comp */
  const double c23 = c_m23();
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gamf = m->g[k] * gam[i] + (1. - m->g[k]) * gam[j];
    const double vfl = -(c23 * gamf * (c->fgu[k] + c->fgv[Fi + k] + c->fgw[2 * Fi + k]) * S[k]);
    c->fcf[k] = c->fcf[k] + 0.;
    c->fff[k] = c->fff[k] + 0.;
    c->fvf[k] = c->fvf[k] + vfl;
    c->ftf[k] = c->ftf[k] + vfl;
  }
  for (int b = 0; b < Fb; ++b) {
    const int f = Fi + b;
    const double gamf = c->turb ? c->bmed[b] : c->mu;
    const double vfl = -(c23 * gamf * (c->bgu[b] + c->bgv[Fb + b] + c->bgw[2 * Fb + b]) * S[f]);
    c->fcf[f] = c->fcf[f] + 0.;
    c->fff[f] = c->fff[f] + 0.;
    c->fvf[f] = c->fvf[f] + vfl;
    c->ftf[f] = c->ftf[f] + vfl;
  }
  /* upwind convection */
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double md = c->mdot[k];
    const double cfl = md > 0. ? md : 0.;
    const double ffl = -(-md > 0. ? -md : 0.);
    c->fcf[k] = c->fcf[k] + cfl;
    c->fff[k] = c->fff[k] + ffl;
    c->ftf[k] = c->ftf[k] + cfl * phi[i] + ffl * phi[j];
  }
  /* deferred high-resolution correction (AssembleConvection.for:156-165) */
  if (c->hr_fw != OR_HR_NONE && c->iter >= c->hr_start)
    or_hr_correction(c, phi, comp == 0 ? c->gu : comp == 1 ? c->gv : c->gw, c->ftf);
  /* effective divergence (CalculateEffectiveDivergence.for:16-28); the
   * boundary mass fluxes of walls are zero */
  memset(c->effdiv, 0, sizeof(double) * N);
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    c->effdiv[i] = c->effdiv[i] + c->mdot[k];
    c->effdiv[j] = c->effdiv[j] - c->mdot[k];
  }
  for (int b = 0; b < Fb; ++b) {
    const int k = m->owner[Fi + b];
    c->effdiv[k] = c->effdiv[k] + c->bmdot[b];
  }
  for (int i = 0; i < N; ++i) {
    const double e = c->effdiv[i], ep = e > 0.0 ? e : 0.0;
    const double fcel = ep - e;
    const double fvel = -(ep * phi[i]);
    const double ftel = fcel * phi[i] + fvel;
    c->fce[i] = c->fce[i] + fcel;
    c->fte[i] = c->fte[i] + ftel;
  }
  /* periodic faces (AssembleConvection.for:582-627): upwind with the
   * partner cell's value */
  if (c->pbf)
    for (int b = 0; b < Fb; ++b) {
      if (!is_periodic(c, b)) continue;
      const int f = Fi + b, i1 = m->owner[f], j1 = c->pj1[b];
      const double phib = rot(comp, c->u[j1], c->v[j1], c->w[j1]);
      const double bm = c->bmdot[b];
      const double cfl = bm > 0. ? bm : 0., ffl = -(-bm > 0. ? -bm : 0.);
      c->fcf[f] = c->fcf[f] + cfl;
      c->fff[f] = c->fff[f] + ffl;
      c->ftf[f] = c->ftf[f] + (cfl * phi[i1] + ffl * phib);
    }
  /* AssembleLinearPressureTerm (AssembleLinearPressure.for:18-44) */
  if (c->pbf) {
    const double lt = sqrt(c->tuse[0] * c->tuse[0] + c->tuse[1] * c->tuse[1] + c->tuse[2] * c->tuse[2]);
    const double beta = c->beta * c->tuse[comp] / lt;
    for (int i = 0; i < N; ++i) c->fte[i] = c->fte[i] - beta * m->vol[i];
  }
  /* pressure gradient */
  const double *gpc = c->gp + comp * N;
  for (int i = 0; i < N; ++i) c->fte[i] = c->fte[i] + gpc[i] * m->vol[i];
  /* MinusTwoThirdsRhoK: the gradient of 2/3 rho k */
  if (c->turb) {
    const double *grc = c->grk + comp * N;
    for (int i = 0; i < N; ++i) c->fte[i] = c->fte[i] + grc[i] * m->vol[i];
  }
  scatter(c);
  for (int i = 0; i < N; ++i) c->ac[i] = c->ac[i] / c->urf_u;
  for (int i = 0; i < N; ++i) {
    d[i] = m->vol[i] / c->ac[i];
    st[i] = phi[i];
  }
  if (c->du2) { /* SIMPLEC: V/(a_P + sum over the slots of anb) */
    double *d2 = comp == 0 ? c->du2 : comp == 1 ? c->dv2 : c->dw2;
    for (int i = 0; i < N; ++i) {
      double a = c->ac[i];
      for (int s = 0; s < m->nef[i]; ++s) a = a + c->anb[(size_t)s * N + i];
      d2[i] = m->vol[i] / a;
    }
  }
}

int or_case_set_gradient_method(or_case *c, int mom_method, int mom_iters, int cont_method, int cont_iters) {
  if (mom_method < 1 || mom_method > 2 || cont_method < 1 || cont_method > 2 || mom_iters < 0 || cont_iters < 0)
    return -1;
  c->ngrad = mom_method == 1 ? 0 : mom_iters;
  c->ngrad_cont = cont_method == 1 ? 0 : cont_iters;
  return 0;
}

int or_case_set_decompose(or_case *c, int method) {
  if (method < 1 || method > 3) return -1;
  or_mesh_decompose(c->m, method);
  c->decomp = method;
  if (c->pbf) /* the periodic faces' E/T split follows MethodDecomposeS */
    for (int b = 0; b < c->Fb; ++b)
      if (c->pbf[b] >= 0) {
        int qq = 0;
        while (c->m->patch_start[qq + 1] <= c->Fi + b) ++qq;
        periodic_geometry(c, b, c->ppair[qq], c->ptrans);
      }
  return 0;
}

int or_case_set_algorithm(or_case *c, int simplec) {
  if (simplec != 0 && simplec != 1) return -1;
  free(c->du2);
  free(c->dv2);
  free(c->dw2);
  c->du2 = c->dv2 = c->dw2 = NULL;
  if (simplec) {
    c->du2 = dz(c->N);
    c->dv2 = dz(c->N);
    c->dw2 = dz(c->N);
  }
  return 0;
}

void or_case_momentum_solve(or_case *c, int comp) {
  double *phi = comp == 0 ? c->u : comp == 1 ? c->v : c->w;
  double *hist = dz(c->mitmax + 1);
  c->it_mom[comp] = or_solve_equation(c->N, c->W, c->m->enbr, c->m->nef, c->ac, c->anb, c->bc, phi,
                                      c->mrrf, c->mitmax, c->msolver, comp + 1, c->dphi,
                                      c->res_mom[comp], hist);
  free(hist);
  update_periodic(c); /* SolveXMomentumEquation.for:299 */
}

/* AssembleMdot (AssembleMdotUsingRhieChow.for:262-292 face interpolations,
 * :554-759 face loop, steady with LRelaxMomentum) followed by
 * AssemblePressureCorrection (AssemblePressureCorrectionFluxes.for:81-164)
 * and the scatter. */
/* geoDiff of the D S split by MethodDecomposeS (AssembleMdotUsingRhieChow.for
 * :575-615, AssemblePressureCorrectionFluxes.for:99-141); the E part of D S
 * goes to de[3]. */
static double geo_split(int method, double ex, double ey, double ez, double cf, double dusf, double dvsf,
                        double dwsf, double *de) {
  double due, dve, dwe;
  if (method == 1) {
    const double dot = ex * dusf + ey * dvsf + ez * dwsf;
    due = dot * ex; dve = dot * ey; dwe = dot * ez;
  } else if (method == 2) {
    const double mag = sqrt(dusf * dusf + dvsf * dvsf + dwsf * dwsf);
    due = ex * mag; dve = ey * mag; dwe = ez * mag;
  } else {
    const double mag = sqrt(dusf * dusf + dvsf * dvsf + dwsf * dwsf);
    const double eu = dusf / mag, ev = dvsf / mag, ew = dwsf / mag;
    const double dot = ex * eu + ey * ev + ez * ew;
    due = mag * ex / dot; dve = mag * ey / dot; dwe = mag * ez / dot;
  }
  de[0] = due; de[1] = dve; de[2] = dwe;
  return sqrt(due * due + dve * dve + dwe * dwe) / cf;
}

static double geo_diff(const or_case *c, int k, double du, double dv, double dw) {
  const or_mesh *m = c->m;
  double de[3];
  return geo_split(c->decomp, m->ex[k], m->ey[k], m->ez[k], m->dcf[k], du * m->Sx[k], dv * m->Sy[k],
                   dw * m->Sz[k], de);
}

void or_case_pc_assemble(or_case *c) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, nf = m->nfaces;
  memset(c->fcf, 0, sizeof(double) * nf);
  memset(c->fff, 0, sizeof(double) * nf);
  memset(c->fvf, 0, sizeof(double) * nf);
  memset(c->ftf, 0, sizeof(double) * nf);
  memset(c->fce, 0, sizeof(double) * N);
  memset(c->fte, 0, sizeof(double) * N);
  memset(c->pc, 0, sizeof(double) * N); /* InitializePressureCorrection */
  const double rhof = c->rho;
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gf = m->g[k], og = 1. - gf;
    const double du = gf * c->du[i] + og * c->du[j], dv = gf * c->dv[i] + og * c->dv[j],
                 dw = gf * c->dw[i] + og * c->dw[j];
    const double ub = gf * c->u[i] + og * c->u[j], vb = gf * c->v[i] + og * c->v[j],
                 wb = gf * c->w[i] + og * c->w[j];
    const double usf = gf * c->us[i] + og * c->us[j], vsf = gf * c->vs[i] + og * c->vs[j],
                 wsf = gf * c->ws[i] + og * c->ws[j];
    const double pbx = gf * c->gp[i] + og * c->gp[j], pby = gf * c->gp[N + i] + og * c->gp[N + j],
                 pbz = gf * c->gp[2 * N + i] + og * c->gp[2 * N + j];
    const double sx = m->Sx[k], sy = m->Sy[k], sz = m->Sz[k];
    const double mbar = ub * sx + vb * sy + wb * sz;
    double fv = rhof * mbar;
    const double dusf = du * sx, dvsf = dv * sy, dwsf = dw * sz;
    double de[3];
    const double geo = geo_split(c->decomp, m->ex[k], m->ey[k], m->ez[k], m->dcf[k], dusf, dvsf, dwsf, de);
    const double dut = dusf - de[0], dvt = dvsf - de[1], dwt = dwsf - de[2];
    const double cfl = rhof * geo, ffl = -(rhof * geo);
    fv = fv - rhof * (c->fgp[k] * dut + c->fgp[Fi + k] * dvt + c->fgp[2 * Fi + k] * dwt);
    fv = fv + rhof * (pbx * dusf + pby * dvsf + pbz * dwsf);
    if (c->turb) { /* the -2/3 rho k gradient's difference (:627-635) */
      fv = fv - rhof * (c->fgrkc[k] * dut + c->fgrkc[Fi + k] * dvt + c->fgrkc[2 * Fi + k] * dwt);
      fv = fv + rhof * (c->fgrkb[k] * dusf + c->fgrkb[Fi + k] * dvsf + c->fgrkb[2 * Fi + k] * dwsf);
      fv = fv + cfl * c->rhok[i] + ffl * c->rhok[j];
    }
    const double mstar = rhof * (usf * sx + vsf * sy + wsf * sz);
    fv = fv + (1. - c->urf_u) * (c->mdot[k] - mstar);
    c->mdot[k] = cfl * c->p[i] + ffl * c->p[j] + fv;
  }
  /* periodic faces, translational (AssembleMdotUsingRhieChow.for:1086-1416,
   * steady with LRelaxMomentum): the lower patch's Bmdot, its negative on
   * the partner face */
  if (c->pbf)
    for (int b = 0; b < c->Fb; ++b) {
      if (!c->pact[b]) continue;
      const int f = Fi + b, i1 = m->owner[f], j1 = c->pj1[b], b2 = c->pbf[b];
      const double gf = c->pg[b];
      const double sx = m->Sx[f], sy = m->Sy[f], sz = m->Sz[f];
      const double ex = c->pux[b], ey = c->puy[b], ez = c->puz[b], cf = c->pdcf[b];
      const double uf1 = rot(0, c->u[j1], c->v[j1], c->w[j1]), vf1 = rot(1, c->u[j1], c->v[j1], c->w[j1]),
                   wf1 = rot(2, c->u[j1], c->v[j1], c->w[j1]);
      const double ub = gf * c->u[i1] + (1. - gf) * uf1, vb = gf * c->v[i1] + (1. - gf) * vf1,
                   wb = gf * c->w[i1] + (1. - gf) * wf1;
      const double mbar = ub * sx + vb * sy + wb * sz;
      double fv = rhof * mbar;
      double dusf = gf * c->du[i1] + (1. - gf) * c->du[j1];
      dusf = dusf * sx;
      double dvsf = gf * c->dv[i1] + (1. - gf) * c->dv[j1];
      dvsf = dvsf * sy;
      double dwsf = gf * c->dw[i1] + (1. - gf) * c->dw[j1];
      dwsf = dwsf * sz;
      double de[3];
      const double geo = geo_split(c->decomp, ex, ey, ez, cf, dusf, dvsf, dwsf, de);
      const double dut = dusf - de[0], dvt = dvsf - de[1], dwt = dwsf - de[2];
      const double pbx = c->bgp[b], pby = c->bgp[c->Fb + b], pbz = c->bgp[2 * c->Fb + b];
      double corr = (c->p[j1] - c->p[i1]) / c->pdcf[b];
      corr = corr - (pbx * c->pux[b] + pby * c->puy[b] + pbz * c->puz[b]);
      const double pfx = pbx + corr * c->pux[b], pfy = pby + corr * c->puy[b], pfz = pbz + corr * c->puz[b];
      const double cfl = rhof * geo, ffl = -(rhof * geo);
      fv = fv - rhof * (pfx * dut + pfy * dvt + pfz * dwt);
      fv = fv + rhof * (pbx * dusf + pby * dvsf + pbz * dwsf);
      if (c->turb) { /* :1222-1242 */
        const double kbx = c->bgrk[b], kby = c->bgrk[c->Fb + b], kbz = c->bgrk[2 * c->Fb + b];
        double kc = (c->rhok[j1] - c->rhok[i1]) / c->pdcf[b];
        kc = kc - (kbx * c->pux[b] + kby * c->puy[b] + kbz * c->puz[b]);
        const double kfx = kbx + kc * c->pux[b], kfy = kby + kc * c->puy[b], kfz = kbz + kc * c->puz[b];
        fv = fv - rhof * (kfx * dut + kfy * dvt + kfz * dwt);
        fv = fv + rhof * (kbx * dusf + kby * dvsf + kbz * dwsf);
        fv = fv + cfl * c->rhok[i1] + ffl * c->rhok[j1];
      }
      const double uf1s = rot(0, c->us[j1], c->vs[j1], c->ws[j1]), vf1s = rot(1, c->us[j1], c->vs[j1], c->ws[j1]),
                   wf1s = rot(2, c->us[j1], c->vs[j1], c->ws[j1]);
      const double usf = gf * c->us[i1] + (1. - gf) * uf1s, vsf = gf * c->vs[i1] + (1. - gf) * vf1s,
                   wsf = gf * c->ws[i1] + (1. - gf) * wf1s;
      const double mstar = rhof * (usf * sx + vsf * sy + wsf * sz);
      fv = fv + (1. - c->urf_u) * (c->bmdot[b] - mstar);
      c->bmdot[b] = cfl * c->p[i1] + ffl * c->p[j1] + fv;
      c->bmdot[b2] = -c->bmdot[b];
    }
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gf = m->g[k], og = 1. - gf;
    const double *d2u = c->du2 ? c->du2 : c->du, *d2v = c->du2 ? c->dv2 : c->dv, *d2w = c->du2 ? c->dw2 : c->dw;
    const double geo = geo_diff(c, k, gf * d2u[i] + og * d2u[j], gf * d2v[i] + og * d2v[j],
                                gf * d2w[i] + og * d2w[j]);
    c->fcf[k] = rhof * geo;
    c->fff[k] = -(rhof * geo);
    c->ftf[k] = c->mdot[k];
  }
  for (int f = Fi; f < nf; ++f) c->ftf[f] = c->bmdot[f - Fi]; /* FluxTf = Bmdot (:155-164) */
  /* periodic faces (AssemblePressureCorrectionFluxes.for:748-830),
   * translational: the partner's D2 rotated by the identity */
  if (c->pbf)
    for (int b = 0; b < c->Fb; ++b) {
      if (!is_periodic(c, b)) continue;
      const int f = Fi + b, i1 = m->owner[f], j1 = c->pj1[b];
      const double *d2u = c->du2 ? c->du2 : c->du, *d2v = c->du2 ? c->dv2 : c->dv, *d2w = c->du2 ? c->dw2 : c->dw;
      const double gf = c->pg[b];
      const double sx = m->Sx[f], sy = m->Sy[f], sz = m->Sz[f];
      double dusf, dvsf, dwsf;
      if (c->decomp == 3) {
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
      dusf = dusf * sx;
      dvsf = dvsf * sy;
      dwsf = dwsf * sz;
      double de[3];
      const double geo = geo_split(c->decomp, c->pux[b], c->puy[b], c->puz[b], c->pdcf[b], dusf, dvsf, dwsf, de);
      c->fcf[f] = rhof * geo;
      c->fff[f] = -(rhof * geo);
    }
  scatter(c);
}

/* geoDiff of the D2 S split of a periodic face, as updateperiodicBeta forms
 * it (PeriodicBeta.for:75-134) */
static double periodic_beta_geo(const or_case *c, int b) {
  const or_mesh *m = c->m;
  const int f = c->Fi + b, i1 = m->owner[f], j1 = c->pj1[b];
  const double *d2u = c->du2 ? c->du2 : c->du, *d2v = c->du2 ? c->dv2 : c->dv, *d2w = c->du2 ? c->dw2 : c->dw;
  const double gf = c->pg[b];
  double dusf = gf * d2u[i1] + (1. - gf) * d2u[j1];
  dusf = dusf * m->Sx[f];
  double dvsf = gf * d2v[i1] + (1. - gf) * d2v[j1];
  dvsf = dvsf * m->Sy[f];
  double dwsf = gf * d2w[i1] + (1. - gf) * d2w[j1];
  dwsf = dwsf * m->Sz[f];
  double de[3];
  return geo_split(c->decomp, c->pux[b], c->puy[b], c->puz[b], c->pdcf[b], dusf, dvsf, dwsf, de);
}

/* updateperiodicBeta (PeriodicBeta.for:4-386): the pressure-gradient
 * correction dp'/dx that brings the mass flow through the lower periodic
 * patch to periodicMdot, added to periodicBeta, to the lower patch's
 * boundary velocities and mass fluxes, to every cell velocity along the
 * translation and to every interior mass flux. */
static void periodic_beta(or_case *c) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb;
  const double *d2u = c->du2 ? c->du2 : c->du, *d2v = c->du2 ? c->dv2 : c->dv, *d2w = c->du2 ? c->dw2 : c->dw;
  double dpdx = 0.;
  for (int k = 0; k < c->beta_iters; ++k) {
    double mdin = 0., rgd = 0.;
    for (int b = 0; b < Fb; ++b) {
      if (!c->pact[b]) continue;
      const double rhof = c->rho; /* BDensity */
      const double geo = periodic_beta_geo(c, b);
      mdin = mdin + c->bmdot[b];
      rgd = rgd + rhof * geo;
    }
    dpdx = (c->beta_mdot - mdin) / rgd;
    c->beta = c->beta + c->beta_relax * dpdx;
  }
  for (int b = 0; b < Fb; ++b) {
    if (!c->pact[b]) continue;
    const int i1 = m->owner[Fi + b], b2 = c->pbf[b];
    c->bu[b] = c->bu[b] + d2u[i1] * dpdx;
    c->bv[b] = c->bv[b] + d2v[i1] * dpdx;
    c->bw[b] = c->bw[b] + d2w[i1] * dpdx;
    c->bu[b2] = c->bu[b];
    c->bv[b2] = c->bv[b];
    c->bw[b2] = c->bw[b];
    const double rhof = c->rho;
    const double geo = periodic_beta_geo(c, b);
    c->bmdot[b] = c->bmdot[b] + rhof * geo * dpdx;
    c->bmdot[b2] = -c->bmdot[b];
  }
  const double lt = sqrt(c->tuse[0] * c->tuse[0] + c->tuse[1] * c->tuse[1] + c->tuse[2] * c->tuse[2]);
  const double bx = c->tuse[0] / lt, by = c->tuse[1] / lt, bz = c->tuse[2] / lt;
  for (int i = 0; i < N; ++i) {
    c->u[i] = c->u[i] + d2u[i] * dpdx * bx;
    c->v[i] = c->v[i] + d2v[i] * dpdx * by;
    c->w[i] = c->w[i] + d2w[i] * dpdx * bz;
  }
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gf = m->g[k], rhof = c->rho; /* Densityf */
    const double d1u = gf * d2u[i] + (1. - gf) * d2u[j], d1v = gf * d2v[i] + (1. - gf) * d2v[j],
                 d1w = gf * d2w[i] + (1. - gf) * d2w[j];
    double de[3];
    const double geo = geo_split(c->decomp, m->ex[k], m->ey[k], m->ez[k], m->dcf[k], d1u * m->Sx[k],
                                 d1v * m->Sy[k], d1w * m->Sz[k], de);
    c->mdot[k] = c->mdot[k] + rhof * geo * dpdx;
  }
}

/* The solve and corrections of SolveContinuity (:76-103):
 * UpdateBoundaryPressureCorrection (walls), CorrectPressure, CorrectVelocity,
 * Correctmdot, UpdateBoundaryPressure (walls). */
void or_case_pc_solve_correct(or_case *c) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb;
  double *hist = dz(c->pitmax + 1);
  if (c->plmg) {
    /* not used by the restatement: the reference multigrid path is covered by
     * or_amg.c */
  }
  c->it_pc = or_solve_equation(N, c->W, m->enbr, m->nef, c->ac, c->anb, c->bc, c->pc, c->prrf,
                               c->pitmax, c->psolver, 4, c->dphi, c->res_pc, hist);
  free(hist);
  for (int b = 0; b < Fb; ++b) c->bpc[b] = c->pc[m->owner[Fi + b]];
  periodic_values(c, c->pc, c->bpc);
  const double rpc = c->refcell >= 0 ? c->pc[c->refcell] : 0.0;
  for (int i = 0; i < N; ++i) c->p[i] = c->p[i] + c->urf_p * (c->pc[i] - rpc);
  gradient_n(c, c->pc, c->bpc, c->gpc, c->bgpc, c->ngrad_cont);
  if (c->lim_cont) or_limit_gradient(c, c->lim_cont, c->pc, c->bpc, c->gpc, c->bgpc);
  periodic_gradient(c, c->gpc, c->bgpc);
  for (int i = 0; i < N; ++i) {
    const double d2u = c->du2 ? c->du2[i] : c->du[i], d2v = c->du2 ? c->dv2[i] : c->dv[i],
                 d2w = c->du2 ? c->dw2[i] : c->dw[i];
    c->u[i] = c->u[i] - d2u * c->gpc[i];
    c->v[i] = c->v[i] - d2v * c->gpc[N + i];
    c->w[i] = c->w[i] - d2w * c->gpc[2 * N + i];
  }
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    c->mdot[k] = c->mdot[k] + c->fcf[k] * c->pc[i] + c->fff[k] * c->pc[j];
  }
  for (int b = 0; b < Fb; ++b) {
    if (is_periodic(c, b)) continue;
    const int f = Fi + b, i1 = m->owner[f];
    c->bp[b] = c->p[i1] + c->bgp[b] * m->dx[f] + c->bgp[Fb + b] * m->dy[f] + c->bgp[2 * Fb + b] * m->dz[f];
  }
  periodic_values(c, c->p, c->bp);
  if (c->pbf) {
    /* CorrectBoundarymdot, periodic faces (CorrectMdotandVelocity.for:115-145) */
    const double *d2u = c->du2 ? c->du2 : c->du, *d2v = c->du2 ? c->dv2 : c->dv, *d2w = c->du2 ? c->dw2 : c->dw;
    for (int b = 0; b < Fb; ++b) {
      if (!c->pact[b]) continue;
      const int f = Fi + b, i1 = m->owner[f], j1 = c->pj1[b], b2 = c->pbf[b];
      const double rhof = 1.;
      c->bu[b] = c->bu[b] - d2u[i1] * c->gpc[i1];
      c->bv[b] = c->bv[b] - d2v[i1] * c->gpc[N + i1];
      c->bw[b] = c->bw[b] - d2w[i1] * c->gpc[2 * N + i1];
      c->bmdot[b] = c->bmdot[b] + rhof * c->fcf[f] * c->pc[i1] + rhof * c->fff[f] * c->pc[j1];
      c->bu[b2] = rot(0, c->bu[b], c->bv[b], c->bw[b]);
      c->bv[b2] = rot(1, c->bu[b], c->bv[b], c->bw[b]);
      c->bw[b2] = rot(2, c->bu[b], c->bv[b], c->bw[b]);
      c->bmdot[b2] = -c->bmdot[b];
    }
    if (c->beta_calc && c->iter % c->beta_every == 0) periodic_beta(c);
  }
  or_case_update_symmetry(c); /* SolveContinuityEquation.for:96 */
}

/* L1 of the cell mass imbalance sum(mdot out). */
double or_case_mass_imbalance(or_case *c) {
  const or_mesh *m = c->m;
  double *r = dz(c->N), s = 0;
  for (int k = 0; k < c->Fi; ++k) {
    r[m->owner[k]] += c->mdot[k];
    r[m->neigh[k]] -= c->mdot[k];
  }
  for (int i = 0; i < c->N; ++i) s += fabs(r[i]);
  free(r);
  return s;
}

void or_case_iterate(or_case *c, int niter) {
  /* the UpdateSymmetry of the start-up (MainProgram.for:128) */
  if (c->sym_dirty) or_case_update_symmetry(c);
  for (int it = 0; it < niter; ++it) {
    c->iter = c->iter + 1; /* nIter of MainProgram.for:257 */
    or_case_momentum_gradients(c);
    for (int comp = 0; comp < 3; ++comp) {
      or_case_momentum_assemble(c, comp);
      or_case_momentum_solve(c, comp);
    }
    or_case_pc_assemble(c);
    or_case_pc_solve_correct(c);
    if (c->turb) or_case_sst(c);
  }
  c->mass_l1 = or_case_mass_imbalance(c);
}

/* ===================================================================== */
/* k-omega SST, ModelNumber 14, low-Reynolds-number walls ('menter').     */

/* SetTurbulenceModelCoefficients, 'komegasst' (SetModelCoefficients.for
 * :266-300): REAL(4) literals widened to double on assignment. */
#define SST_SIGK1 ((double)0.85f)
#define SST_SIGK2 1.0
#define SST_SIGW1 ((double)0.5f)
#define SST_SIGW2 ((double)0.856f)
#define SST_ALPHA1 ((double)(5.0f / 9.0f))
#define SST_ALPHA2 ((double)0.44f)
#define SST_BETA1 ((double)0.075f)
#define SST_BETA2 ((double)0.0828f)
#define SST_BETTA ((double)0.075f)
#define SST_A1 ((double)0.31f)
#define SST_CMU ((double)0.09f)
#define SST_C1LIM 10.0 /* c1limiter (SetModelCoefficients.for:18) */
#define SST_TINY 1.e-15 /* Constants1 tiny */

int or_case_set_sst(or_case *c, const int *io, const double *d, const double *k, const double *omega,
                    const double *mut, const double *wd, const double *bk, const double *bomega,
                    const double *bmut) {
  for (int q = 0; q < 2; ++q) {
    const int fw = io[4 + 2 * q], sc = io[5 + 2 * q];
    if (io[2 * q] < 0 || io[2 * q] > 5 || io[1 + 2 * q] < 0 || fw < OR_HR_NONE || fw > OR_HR_NVF ||
        sc < OR_CONV_UPWIND || sc > OR_CONV_GAMMA || io[8 + 2 * q] < 1 || io[8 + 2 * q] > 2 ||
        io[9 + 2 * q] < 0 || io[12 + q] < 0 || io[12 + q] > 2)
      return -1;
  }
  if (!(d[0] > 0.) || !(d[1] > 0.) || !(d[2] > 0.) || !(d[5] > 0.)) return -1;
  const int N = c->N, Fb = c->Fb, Fi = c->Fi;
  c->turb = 1;
  c->k_solver = io[0]; c->k_itmax = io[1]; c->w_solver = io[2]; c->w_itmax = io[3];
  c->k_fw = io[4]; c->k_scheme = io[5]; c->w_fw = io[6]; c->w_scheme = io[7];
  c->k_ngrad = io[8] == 1 ? 0 : io[9];
  c->w_ngrad = io[10] == 1 ? 0 : io[11];
  c->k_lim = io[12]; c->w_lim = io[13];
  c->urf_k = d[0]; c->urf_w = d[1]; c->urf_mut = d[2]; c->k_rrf = d[3]; c->w_rrf = d[4];
  c->max_visc_ratio = d[5];
  double **cells[] = {&c->tk, &c->tw, &c->mut, &c->wd, &c->rhok, &c->F1, &c->F2, &c->prod, &c->sr, &c->med};
  for (size_t q = 0; q < sizeof(cells) / sizeof(cells[0]); ++q) { free(*cells[q]); *cells[q] = dz(N); }
  double **bfs[] = {&c->btk, &c->btw, &c->bmut, &c->brhok, &c->bF1, &c->bF2, &c->bsr, &c->bmed};
  for (size_t q = 0; q < sizeof(bfs) / sizeof(bfs[0]); ++q) { free(*bfs[q]); *bfs[q] = dz(Fb); }
  double **g3[] = {&c->gk, &c->gom, &c->grk};
  for (size_t q = 0; q < 3; ++q) { free(*g3[q]); *g3[q] = dz(3 * (size_t)N); }
  double **b3[] = {&c->bgk, &c->bgom, &c->bgrk};
  for (size_t q = 0; q < 3; ++q) { free(*b3[q]); *b3[q] = dz(3 * (size_t)Fb); }
  double **f3[] = {&c->fgk, &c->fgom, &c->fgrkb, &c->fgrkc};
  for (size_t q = 0; q < 4; ++q) { free(*f3[q]); *f3[q] = dz(3 * (size_t)Fi); }
  memcpy(c->tk, k, sizeof(double) * N);
  memcpy(c->tw, omega, sizeof(double) * N);
  memcpy(c->mut, mut, sizeof(double) * N);
  memcpy(c->wd, wd, sizeof(double) * N);
  memcpy(c->btk, bk, sizeof(double) * Fb);
  memcpy(c->btw, bomega, sizeof(double) * Fb);
  memcpy(c->bmut, bmut, sizeof(double) * Fb);
  return 0; /* F1, F2 start at 0 and fr1Coefficient at 1 (AllocateStorageRoutine.for:2618-2631) */
}

static int is_wall(const or_case *c, int b) { return !c->btype || c->btype[b] == OR_BC_WALL; }

/* CalculateStrainRateTensor (CalculateStrainRate.for): S_ij of the velocity
 * gradients and |S| = sqrt(2 S:S), the terms summed in the reference's order */
static double strain_rate(double ux, double uy, double uz, double vx, double vy, double vz, double wx, double wy,
                          double wz) {
  const double s11 = ux, s12 = 0.5 * (uy + vx), s13 = 0.5 * (uz + wx), s22 = vy, s23 = 0.5 * (vz + wy), s33 = wz;
  return sqrt(2. * (s11 * s11 + s12 * s12 + s13 * s13 + s12 * s12 + s22 * s22 + s23 * s23 + s13 * s13 + s23 * s23 +
                    s33 * s33));
}

static void sst_strain(or_case *c) {
  const int N = c->N, Fb = c->Fb;
  for (int i = 0; i < N; ++i)
    c->sr[i] = strain_rate(c->gu[i], c->gu[N + i], c->gu[2 * N + i], c->gv[i], c->gv[N + i], c->gv[2 * N + i],
                           c->gw[i], c->gw[N + i], c->gw[2 * N + i]);
  for (int b = 0; b < Fb; ++b)
    c->bsr[b] = strain_rate(c->bgu[b], c->bgu[Fb + b], c->bgu[2 * Fb + b], c->bgv[b], c->bgv[Fb + b],
                            c->bgv[2 * Fb + b], c->bgw[b], c->bgw[Fb + b], c->bgw[2 * Fb + b]);
}

/* CalculateTurbulenceProduction, case 14 (CalculateProductionOfTurbulence.for
 * :297-316): P_k = max(tau:grad u, 0) with tau = 2 mu_t S - 2/3 rho k I
 * (CalculateShearStressValues.for, the non-k-epsilon branch), limited to
 * c1limiter cmu rho max(k omega, 0). */
static void sst_production(or_case *c) {
  const int N = c->N;
  const double t23 = (double)0.666666666666666666666667f;
  for (int i = 0; i < N; ++i) {
    const double ux = c->gu[i], uy = c->gu[N + i], uz = c->gu[2 * N + i];
    const double vx = c->gv[i], vy = c->gv[N + i], vz = c->gv[2 * N + i];
    const double wx = c->gw[i], wy = c->gw[N + i], wz = c->gw[2 * N + i];
    const double s11 = ux, s12 = 0.5 * (uy + vx), s13 = 0.5 * (uz + wx), s22 = vy, s23 = 0.5 * (vz + wy), s33 = wz;
    const double mt = c->mut[i];
    double t11 = mt * 2. * s11, t12 = mt * 2. * s12, t13 = mt * 2. * s13, t22 = mt * 2. * s22,
           t23v = mt * 2. * s23, t33 = mt * 2. * s33;
    const double k1 = mx2(c->tk[i], 0.);
    t11 = t11 - t23 * c->rho * k1;
    t22 = t22 - t23 * c->rho * k1;
    t33 = t33 - t23 * c->rho * k1;
    const double p = t11 * ux + t12 * uy + t13 * uz + t12 * vx + t22 * vy + t23v * vz + t13 * wx + t23v * wy + t33 * wz;
    const double pk = mx2(p, 0.);
    const double lim = SST_C1LIM * SST_CMU * c->rho * mx2(c->tk[i] * c->tw[i], 0.);
    c->prod[i] = mn2(pk, lim);
  }
}

/* CalculateTurbulenceInterpolationFactors (CalculateTurbulenceModelCoefficients
 * .for:3063-3290) without roughness or rotation corrections: F1 and F2 of
 * cells and boundary faces (the boundary faces with the owner's wall
 * distance); F3 = F4 = 1. */
static double sst_f1(double k, double w, double d, double mu, double rho, double kx, double ky, double kz,
                     double wx, double wy, double wz) {
  const double tke1 = mx2(k, 0.), tom1 = mx2(w, SST_TINY);
  const double f1 = sqrt(tke1) / (SST_CMU * tom1 * d);
  const double f2 = 500. * mu / (rho * (d * d) * tom1);
  const double f4 = 2. * rho * SST_SIGW2 / tom1;
  const double f5 = kx * wx + ky * wy + kz * wz;
  const double f6 = f4 * f5;
  const double cd = mx2(f6, 1.e-10);
  const double f3 = 4. * rho * SST_SIGW2 * tke1 / (cd * (d * d));
  const double g1 = mn2(mn2(mx2(f1, f2), f3), 10.);
  return tanh(g1 * g1 * g1 * g1);
}
static double sst_f2(double k, double w, double d, double mu, double rho) {
  const double tke1 = mx2(k, 0.), tom1 = mx2(w, SST_TINY);
  const double f1 = 2. * sqrt(tke1) / (SST_CMU * tom1 * d);
  const double f2 = 500. * mu / (rho * (d * d) * tom1);
  const double g2 = mn2(mx2(f1, f2), 100.);
  return tanh(g2 * g2);
}
static void sst_blending(or_case *c) {
  const or_mesh *m = c->m;
  const int N = c->N, Fb = c->Fb;
  for (int i = 0; i < N; ++i)
    c->F1[i] = sst_f1(c->tk[i], c->tw[i], c->wd[i], c->visc[i], c->rho, c->gk[i], c->gk[N + i], c->gk[2 * N + i],
                      c->gom[i], c->gom[N + i], c->gom[2 * N + i]);
  for (int b = 0; b < Fb; ++b) {
    const int o = m->owner[c->Fi + b];
    c->bF1[b] = sst_f1(c->btk[b], c->btw[b], c->wd[o], c->mu, c->rho, c->bgk[b], c->bgk[Fb + b], c->bgk[2 * Fb + b],
                       c->bgom[b], c->bgom[Fb + b], c->bgom[2 * Fb + b]);
  }
  for (int i = 0; i < N; ++i) c->F2[i] = sst_f2(c->tk[i], c->tw[i], c->wd[i], c->visc[i], c->rho);
  for (int b = 0; b < Fb; ++b)
    c->bF2[b] = sst_f2(c->btk[b], c->btw[b], c->wd[m->owner[c->Fi + b]], c->mu, c->rho);
}

/* Gradient of k or omega (CalculateGradients.for: Green-Gauss, the limiter,
 * CorrectGradientAtPeriodicBoundary) and its face interpolation 'average' */
static void sst_gradient(or_case *c, const double *phi, const double *bphi, double *gr, double *bgr, double *fg,
                         int ngrad, int lim) {
  gradient_n(c, phi, bphi, gr, bgr, ngrad);
  if (lim) or_limit_gradient(c, lim, phi, bphi, gr, bgr);
  periodic_gradient(c, gr, bgr);
  face_gradient(c, phi, gr, fg, OR_GI_AVERAGE);
}

/* One transport equation of the SST model, var 0 = k ('tke',
 * SolveTKE.for), 1 = omega ('tomega', SolveTOmega.for): gradient, effective
 * diffusion coefficient, sources (AssembleTKESources.for:985-1096,
 * AssembleTOmegaSources.for:558-713, low-Reynolds-number walls), diffusion
 * (AssembleIsotropicDiffusion.for:2172-2505), convection
 * (AssembleConvection.for:90-110,156-178,808-966), the empty user source,
 * scatter, under-relaxation, SolveEquation, bounding and the boundary values
 * (UpdateVariablesFromBoundaryConditions.for:76-146,431-501). */
static void sst_equation(or_case *c, int var) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb, nf = m->nfaces;
  double *phi = var ? c->tw : c->tk, *bphi = var ? c->btw : c->btk;
  double *gr = var ? c->gom : c->gk, *bgr = var ? c->bgom : c->bgk, *fg = var ? c->fgom : c->fgk;
  memset(c->fcf, 0, sizeof(double) * nf);
  memset(c->fff, 0, sizeof(double) * nf);
  memset(c->fvf, 0, sizeof(double) * nf);
  memset(c->ftf, 0, sizeof(double) * nf);
  memset(c->fce, 0, sizeof(double) * N);
  memset(c->fte, 0, sizeof(double) * N);
  sst_gradient(c, phi, bphi, gr, bgr, fg, var ? c->w_ngrad : c->k_ngrad, var ? c->w_lim : c->k_lim);
  /* CalculateEffectiveDiffusionCoefficient: mu + mu_t sigma(F1), F1 of the
   * last omega assembly */
  const double s1 = var ? SST_SIGW1 : SST_SIGK1, s2 = var ? SST_SIGW2 : SST_SIGK2;
  for (int i = 0; i < N; ++i) {
    const double sl = c->F1[i] * s1 + (1. - c->F1[i]) * s2;
    c->med[i] = c->visc[i] + c->mut[i] * sl;
  }
  for (int b = 0; b < Fb; ++b) {
    const double sl = c->bF1[b] * s1 + (1. - c->bF1[b]) * s2;
    c->bmed[b] = c->mu + c->bmut[b] * sl;
  }
  if (var == 0) {
    sst_production(c);
    for (int i = 0; i < N; ++i) {
      const double rted = SST_CMU * c->rho * mx2(c->tw[i], 0.);
      const double pt = 1. * c->prod[i]; /* fr1Coefficient = 1 */
      c->fce[i] = c->fce[i] + rted * m->vol[i];
      c->fte[i] = c->fte[i] - pt * m->vol[i] + rted * m->vol[i] * mx2(c->tk[i], 0.);
    }
    for (int b = 0; b < Fb; ++b) /* SetTKEWallBC */
      if (is_wall(c, b)) c->btk[b] = 0.;
  } else {
    sst_blending(c);
    for (int i = 0; i < N; ++i) {
      const double v = m->vol[i], w = c->tw[i], f1 = c->F1[i];
      double t1 = c->rho * c->prod[i] * 1. / mx2(c->mut[i], 1.e-10);
      const double al = f1 * SST_ALPHA1 + (1. - f1) * SST_ALPHA2;
      c->fte[i] = c->fte[i] - v * mx2(al * t1, 0.);
      t1 = c->rho * mx2(w, 0.) * v;
      double be = f1 * SST_BETA1 + (1. - f1) * SST_BETA2;
      be = 1. * be; /* F4factor = 1 */
      c->fce[i] = c->fce[i] + 2. * mx2(be * t1, 0.);
      c->fte[i] = c->fte[i] + mx2(be * t1 * w, 0.);
      t1 = 2. * c->rho * SST_SIGW2 / mx2(w, SST_TINY);
      const double t2 = c->gk[i] * c->gom[i] + c->gk[N + i] * c->gom[N + i] + c->gk[2 * N + i] * c->gom[2 * N + i];
      t1 = (1. - f1) * t1 * t2 * v;
      if (t1 > 0.) {
        c->fte[i] = c->fte[i] - t1;
      } else {
        const double ted1 = mx2(w, SST_TINY);
        c->fce[i] = c->fce[i] - t1 / ted1;
        c->fte[i] = c->fte[i] - t1;
      }
    }
    for (int b = 0; b < Fb; ++b) /* SetTOmegaWallBC, 'menter' */
      if (is_wall(c, b)) {
        const double d = c->wd[m->owner[Fi + b]];
        c->btw[b] = 10. * 6. * c->mu / (c->rho * SST_BETTA * (d * d));
      }
  }
  /* the wall faces' diffusion, assembled with the sources */
  for (int b = 0; b < Fb; ++b) {
    if (!is_wall(c, b)) continue;
    const int f = Fi + b, i1 = m->owner[f];
    const double gamf = c->bmed[b];
    const double cfl = gamf * m->gdiff[f], ffl = -(gamf * m->gdiff[f]);
    const double vfl = -(gamf * (bgr[b] * m->Tx[f] + bgr[Fb + b] * m->Ty[f] + bgr[2 * Fb + b] * m->Tz[f]));
    c->fcf[f] = c->fcf[f] + cfl;
    c->fff[f] = c->fff[f] + ffl;
    c->fvf[f] = c->fvf[f] + vfl;
    c->ftf[f] = c->ftf[f] + cfl * phi[i1] + ffl * bphi[b] + vfl;
  }
  /* diffusion: interior faces */
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double gf = m->g[k];
    const double gamf = gf * c->med[i] + (1. - gf) * c->med[j];
    const double cfl = gamf * m->gdiff[k], ffl = -(gamf * m->gdiff[k]);
    const double vfl = -(gamf * (fg[k] * m->Tx[k] + fg[Fi + k] * m->Ty[k] + fg[2 * Fi + k] * m->Tz[k]));
    c->fcf[k] = c->fcf[k] + cfl;
    c->fff[k] = c->fff[k] + ffl;
    c->fvf[k] = c->fvf[k] + vfl;
    c->ftf[k] = c->ftf[k] + cfl * phi[i] + ffl * phi[j] + vfl;
  }
  /* symmetry planes: no flux, the boundary value the owner's */
  for (int b = 0; b < Fb; ++b)
    if (c->btype && c->btype[b] == OR_BC_SYMMETRY) bphi[b] = phi[m->owner[Fi + b]];
  /* periodic faces: BGamaFace interpolated between the two cells */
  if (c->pbf)
    for (int b = 0; b < Fb; ++b) {
      if (!is_periodic(c, b)) continue;
      const int f = Fi + b, i1 = m->owner[f], j1 = c->pj1[b];
      const double gamf = per_mix(c, b, c->med[i1], c->med[j1]);
      const double cfl = gamf * c->pgd[b], ffl = -(gamf * c->pgd[b]);
      const double vfl = -(gamf * (bgr[b] * c->ptx[b] + bgr[Fb + b] * c->pty[b] + bgr[2 * Fb + b] * c->ptz[b]));
      c->fcf[f] = c->fcf[f] + cfl;
      c->fff[f] = c->fff[f] + ffl;
      c->fvf[f] = c->fvf[f] + vfl;
      const double ftl = cfl * phi[i1] + ffl * phi[j1] + vfl;
      c->ftf[f] = c->ftf[f] + ftl;
    }
  /* convection: upwind, the deferred correction, effective divergence */
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    const double md = c->mdot[k];
    const double cfl = md > 0. ? md : 0.;
    const double ffl = -(-md > 0. ? -md : 0.);
    c->fcf[k] = c->fcf[k] + cfl;
    c->fff[k] = c->fff[k] + ffl;
    c->ftf[k] = c->ftf[k] + cfl * phi[i] + ffl * phi[j];
  }
  if (c->iter >= c->hr_start)
    hr_correction(c, var ? c->w_fw : c->k_fw, var ? c->w_scheme : c->k_scheme, 0., phi, gr, c->ftf);
  memset(c->effdiv, 0, sizeof(double) * N);
  for (int k = 0; k < Fi; ++k) {
    const int i = m->owner[k], j = m->neigh[k];
    c->effdiv[i] = c->effdiv[i] + c->mdot[k];
    c->effdiv[j] = c->effdiv[j] - c->mdot[k];
  }
  for (int b = 0; b < Fb; ++b) {
    const int k = m->owner[Fi + b];
    c->effdiv[k] = c->effdiv[k] + c->bmdot[b];
  }
  for (int i = 0; i < N; ++i) {
    const double e = c->effdiv[i], ep = e > 0.0 ? e : 0.0;
    const double fcel = ep - e;
    const double fvel = -(ep * phi[i]);
    const double ftel = fcel * phi[i] + fvel;
    c->fce[i] = c->fce[i] + fcel;
    c->fte[i] = c->fte[i] + ftel;
  }
  if (c->pbf)
    for (int b = 0; b < Fb; ++b) {
      if (!is_periodic(c, b)) continue;
      const int f = Fi + b, i1 = m->owner[f], j1 = c->pj1[b];
      const double bm = c->bmdot[b];
      const double cfl = bm > 0. ? bm : 0., ffl = -(-bm > 0. ? -bm : 0.);
      c->fcf[f] = c->fcf[f] + cfl;
      c->fff[f] = c->fff[f] + ffl;
      c->ftf[f] = c->ftf[f] + cfl * phi[i1] + ffl * phi[j1];
    }
  /* AssembleSourceTerm with the zero ScTKE/SbTKE (AssembleSources.for) */
  for (int i = 0; i < N; ++i) {
    c->fce[i] = c->fce[i] + 0.;
    c->fte[i] = c->fte[i] - (0. + 0. * phi[i]) * m->vol[i];
  }
  scatter(c);
  const double urf = var ? c->urf_w : c->urf_k;
  for (int i = 0; i < N; ++i) c->ac[i] = c->ac[i] / urf;
  /* ModifyCoefficientsAlongWalls: nothing for the low-Reynolds-number SST */
  const int itmax = var ? c->w_itmax : c->k_itmax;
  double *hist = dz(itmax + 1);
  const int it = or_solve_equation(N, c->W, m->enbr, m->nef, c->ac, c->anb, c->bc, phi, var ? c->w_rrf : c->k_rrf,
                                   itmax, var ? c->w_solver : c->k_solver, var ? 10 : 8, c->dphi,
                                   var ? c->res_w : c->res_k, hist);
  free(hist);
  if (var) c->it_w = it; else c->it_k = it;
  /* BoundTurbulentKE / BoundTurbulentOmega (BoundTurbulentVariables.for) */
  for (int i = 0; i < N; ++i) phi[i] = mx2(phi[i], SST_TINY);
  for (int b = 0; b < Fb; ++b) bphi[b] = mx2(bphi[b], 0.);
  for (int b = 0; b < Fb; ++b) {
    if (c->btype && c->btype[b] == OR_BC_SYMMETRY) bphi[b] = phi[m->owner[Fi + b]];
    else if (is_periodic(c, b)) bphi[b] = per_mix(c, b, phi[m->owner[Fi + b]], phi[c->pj1[b]]);
  }
}

/* CalculateTurbulentViscosity, case 14 (CalculateTurbulentViscosity.for
 * :483-513,925-938): relaxed rho a1 k / max(a1 omega, |S| F2 F3, tiny),
 * capped at MaximumViscosityRatio times the viscosity. */
static void sst_viscosity(or_case *c) {
  const int N = c->N, Fb = c->Fb;
  const double u = c->urf_mut;
  for (int i = 0; i < N; ++i) {
    const double den = mx2(mx2(SST_A1 * c->tw[i], c->sr[i] * c->F2[i] * 1.), SST_TINY);
    c->mut[i] = (1. - u) * c->mut[i] + u * c->rho * SST_A1 * mx2(c->tk[i], 0.) / den;
  }
  for (int b = 0; b < Fb; ++b) {
    const double den = mx2(mx2(SST_A1 * c->btw[b], c->bsr[b] * c->bF2[b] * 1.), SST_TINY);
    c->bmut[b] = (1. - u) * c->bmut[b] + u * c->rho * SST_A1 * mx2(c->btk[b], 0.) / den;
  }
  for (int i = 0; i < N; ++i) c->mut[i] = mn2(c->mut[i], c->max_visc_ratio * c->visc[i]);
  for (int b = 0; b < Fb; ++b) c->bmut[b] = mn2(c->bmut[b], c->max_visc_ratio * c->mu);
}

void or_case_sst(or_case *c) {
  if (!c->turb) return;
  sst_strain(c);
  sst_equation(c, 0);
  sst_equation(c, 1);
  sst_viscosity(c);
}

/* CalculateNormalDistanceToWall, MethodCalculateNormalDistance 2
 * (CalculateDistanceToWall.for:154-470): the Poisson problem lap(phi) = -1
 * with phi = 0 on the no-slip walls and no flux elsewhere (periodic faces
 * included), Green-Gauss gradients (method 2, two sweeps) and 'average' face
 * gradients, outer passes of assembly and solve until the solve's final L1
 * residual is at most tol; then d = -|grad phi| + sqrt(|grad phi|^2 + 2 phi)
 * with the gradient of the last pass, and the cells next to a wall the
 * normal distance of their (last) wall face.  The inner solves use
 * or_solve_equation (solver, rrf, itmax) where the reference uses its
 * algebraic multigrid: the converged result is the same to within the
 * tolerance.  Returns the outer passes run, or -1 when maxouter passes do not
 * converge.  phi_out may be NULL. */
int or_case_wall_distance(or_case *c, double tol, int maxouter, int solver, double rrf, int itmax, double *wd,
                          double *phi_out) {
  const or_mesh *m = c->m;
  const int N = c->N, Fi = c->Fi, Fb = c->Fb, nf = m->nfaces;
  double *phi = dz(N), *bphi = dz(Fb), *gr = dz(3 * (size_t)N), *bgr = dz(3 * (size_t)Fb);
  double *fg = dz(3 * (size_t)Fi), *hist = dz(itmax + 1);
  for (int i = 0; i < N; ++i) phi[i] = 1.;
  for (int b = 0; b < Fb; ++b) bphi[b] = is_wall(c, b) ? 0. : 1.;
  double res = 1.e10;
  int outer = 1, passes = 0;
  const double stop = tol < 0. ? -tol * 0.75 : tol;
  while (outer < maxouter && res > stop) {
    memset(c->fcf, 0, sizeof(double) * nf);
    memset(c->fff, 0, sizeof(double) * nf);
    memset(c->fvf, 0, sizeof(double) * nf);
    memset(c->ftf, 0, sizeof(double) * nf);
    memset(c->fce, 0, sizeof(double) * N);
    memset(c->fte, 0, sizeof(double) * N);
    gradient_n(c, phi, bphi, gr, bgr, 2);
    periodic_gradient(c, gr, bgr);
    face_gradient(c, phi, gr, fg, OR_GI_AVERAGE);
    for (int k = 0; k < Fi; ++k) {
      const int i = m->owner[k], j = m->neigh[k];
      const double cfl = m->gdiff[k], ffl = -m->gdiff[k];
      const double vfl = -fg[k] * m->Tx[k] - fg[Fi + k] * m->Ty[k] - fg[2 * Fi + k] * m->Tz[k];
      c->fcf[k] = c->fcf[k] + cfl;
      c->fff[k] = c->fff[k] + ffl;
      c->fvf[k] = c->fvf[k] + vfl;
      c->ftf[k] = c->ftf[k] + cfl * phi[i] + ffl * phi[j] + vfl;
    }
    for (int b = 0; b < Fb; ++b) {
      const int f = Fi + b, i1 = m->owner[f];
      if (is_wall(c, b)) {
        bphi[b] = 0.;
        const double cfl = m->gdiff[f], ffl = -m->gdiff[f];
        const double vfl = -(bgr[b] * m->Tx[f] + bgr[Fb + b] * m->Ty[f] + bgr[2 * Fb + b] * m->Tz[f]);
        c->fcf[f] = c->fcf[f] + cfl;
        c->fff[f] = c->fff[f] + ffl;
        c->fvf[f] = c->fvf[f] + vfl;
        c->ftf[f] = c->ftf[f] + cfl * phi[i1] + ffl * bphi[b] + vfl;
      } else {
        bphi[b] = phi[i1];
        c->fcf[f] = c->fcf[f] + 0.;
        c->fff[f] = c->fff[f] + 0.;
        c->fvf[f] = c->fvf[f] + 0.;
        c->ftf[f] = c->ftf[f] + 0.;
      }
    }
    /* AssembleSourceTerm with ScWall = 0, SbWall = 1 */
    for (int i = 0; i < N; ++i) {
      c->fce[i] = c->fce[i] + 0.;
      c->fte[i] = c->fte[i] - (1. + 0. * phi[i]) * m->vol[i];
    }
    /* no implicit periodic coupling: the periodic faces carry no flux, so
     * the scatter's partner slots receive FluxFf = 0 */
    scatter(c);
    double r4[4];
    double ires = 0.;
    for (int i = 0; i < N; ++i) ires += fabs(c->bc[i]);
    or_solve_equation(N, c->W, m->enbr, m->nef, c->ac, c->anb, c->bc, phi, rrf, itmax, solver, 5, c->dphi, r4,
                      hist);
    /* the solve's final L1 residual (CheckConvergence) of the updated phi */
    double s = 0.;
    for (int i = 0; i < N; ++i) {
      double r = c->bc[i] - c->ac[i] * c->dphi[i];
      for (int q = 0; q < m->nef[i]; ++q) {
        const int kq = m->enbr[(size_t)q * N + i];
        if (kq) r = r - c->anb[(size_t)q * N + i] * c->dphi[kq - 1];
      }
      s += fabs(r);
    }
    res = tol < 0. ? ires : s; /* tol < 0: stop on the pass's initial residual |bc|_1 <= -tol */
    if (tol < 0. && res <= -tol) res = -tol * 0.5;
    outer = outer + 1;
    ++passes;
  }
  for (int i = 0; i < N; ++i) {
    const double t1 = gr[i] * gr[i] + gr[N + i] * gr[N + i] + gr[2 * N + i] * gr[2 * N + i];
    const double t2 = t1 + 2. * phi[i];
    wd[i] = -sqrt(t1) + sqrt(t2);
  }
  for (int b = 0; b < Fb; ++b)
    if (is_wall(c, b)) {
      const int f = Fi + b;
      wd[m->owner[f]] = m->dx[f] * m->nx[f] + m->dy[f] * m->ny[f] + m->dz[f] * m->nz[f];
    }
  if (phi_out) memcpy(phi_out, phi, sizeof(double) * N);
  free(phi); free(bphi); free(gr); free(bgr); free(fg); free(hist);
  return res > stop ? -1 : passes;
}
