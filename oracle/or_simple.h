/*
 * oracle/or_simple.h -- CPU restatement of the reference's mesh topology,
 * geometry and steady SIMPLE iteration for a laminar, incompressible,
 * constant-property flow in a box bounded by no-slip walls (moving or at
 * rest) and symmetry planes.  TEST INFRASTRUCTURE ONLY.
 */
#ifndef OR_SIMPLE_H
#define OR_SIMPLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* polyMesh input (0-based) */
  int npoints, nfaces, nif, ncells, npatches;
  double *x, *y, *z;
  int *fptr, *fnodes, *owner, *neigh, *patch_start;
  /* topology in the reference's numbering (cell numbers 1-based) */
  int w;        /* NElementFaces */
  int *nef;     /* NumbOfElementFaces(N) */
  int *gef;     /* NGlobalEFaces(N,W) column-major, 1-based face numbers */
  int *enbr;    /* ElementNeighbor(N,W), 0 = boundary slot */
  int *iown, *inbr; /* iOwnerNeighbor / iNeighborOwner (Fi), 1-based slots */
  int *nen, *len, len_w; /* NumbOfElementNodes, ListOfElementNodes (N,len_w), 0-based nodes */
  /* geometry over all faces (interior first) and cells */
  double *Sx, *Sy, *Sz, *A, *fcx, *fcy, *fcz, *nx, *ny, *nz;
  double *vol, *xc, *yc, *zc;
  double *dcf, *dx, *dy, *dz, *ex, *ey, *ez, *g;
  double *Ex, *Ey, *Ez, *E, *gdiff, *Tx, *Ty, *Tz, *T;
} or_mesh;

or_mesh *or_mesh_new(int npoints, const double *x, const double *y, const double *z, int nfaces,
                     const int *fptr, const int *fnodes, const int *owner, int nif,
                     const int *neigh, int npatches, const int *patch_size);
void or_mesh_delete(or_mesh *m);
void or_mesh_free(or_mesh *m);
int or_mesh_topology(or_mesh *m);
int or_mesh_geometry(or_mesh *m);
void or_mesh_decompose(or_mesh *m, int method);

/* one SIMPLE case on a mesh */
typedef struct {
  or_mesh *m;
  int N, Fi, Fb, W;
  /* boundary faces (index b = face - Fi): wall velocity */
  double *bu0, *bv0, *bw0;
  /* settings */
  double rho, mu, urf_u, urf_p;
  int msolver, mitmax, psolver, pitmax, plmg, ngrad, refcell;
  double mrrf, prrf;
  /* fields */
  double *u, *v, *w, *p, *pc, *mdot;
  double *gu, *gv, *gw, *gp, *gpc;      /* 3N: x block, y block, z block */
  double *bu, *bv, *bw, *bp, *bpc;      /* Fb */
  double *bgu, *bgv, *bgw, *bgp, *bgpc; /* 3Fb */
  double *fgu, *fgv, *fgw, *fgp;        /* 3Fi face gradients */
  double *du, *dv, *dw, *us, *vs, *ws;  /* D1 (= D2 for SIMPLE), star */
  double *visc;                          /* cell viscosity */
  /* assembly work */
  double *fcf, *fff, *fvf, *ftf;        /* all faces */
  double *fce, *fte, *effdiv;           /* cells */
  double *ac, *anb, *bc, *dphi;
  double res_mom[3][4], res_pc[4], mass_l1;
  int it_mom[3], it_pc;
  /* boundary condition per boundary face (NULL = every face a no-slip
   * wall): OR_BC_WALL or OR_BC_SYMMETRY; bzq[b] = the boundary face whose
   * z-normal the symmetry pressure-gradient projection reads */
  int *btype, *bzq;
  int sym_dirty; /* symmetry boundary velocities not yet derived */
  /* high-resolution convection of the momentum equations (HRFramework,
   * ConvectionScheme, Bleed, nIterStartApplyingHR) and the gradient limiters
   * (LimitGradientMomentum/ContinuityMethod, 0 = off); iter = nIter, the
   * number of the outer iteration in progress */
  int hr_fw, hr_scheme, hr_start, lim_mom, lim_cont, iter;
  double bleed;
  /* Algorithm 'simplec' (StoreTheDValues.for:30-39): D2 = V/(a_P + sum anb)
   * in the p' coefficients and the velocity correction; NULL for 'simple'
   * (D2 = D1) */
  double *du2, *dv2, *dw2;
  /* Green-Gauss correction sweeps of p and p' (nIterGradientContinuity;
   * ngrad above is nIterGradientMomentum); 0 = MethodCalcGradient 1 */
  int ngrad_cont;
  int decomp; /* MethodDecomposeS of the mesh's E/T split and of Rhie-Chow */
  /* GradientInterpolationSchemeMomentum / Continuity (OR_GI_*) */
  int gi_mom, gi_cont;
  /* Translational periodicity with the implicit coupling (the case file's
   * LTranslationalPeriodicity, LPeriodicImplicit; BoundaryConditions2.F90
   * :130-146).  Per boundary face b: pbf[b] the corresponding face on the
   * partner patch (Icorrespondingface, a boundary-face index; -1 off the
   * periodic patches), pact[b] 1 on the lower patch of a pair (the faces
   * the reference's j2 > i2 branches visit), pj1[b] the partner face's owner;
   * the geometry every periodic branch computes from the partner's centroid
   * carried by the partner patch's translation: pg (GFactCF), pdx/pdy/pdz/
   * pdcf (DistCF), pux/puy/puz (its unit vector), and the E/T split of the
   * face area by MethodDecomposeS (AssembleIsotropicDiffusion.for:663-700):
   * pgd (fgDiff), ptx/pty/ptz (FT). */
  int *pbf, *pact, *pj1;
  double *pg, *pdx, *pdy, *pdz, *pdcf, *pux, *puy, *puz, *pgd, *ptx, *pty, *ptz;
  double *bmdot;          /* Bmdot over the boundary faces (0 on walls) */
  double tuse[3];         /* x/y/zTranslationUse */
  double beta, beta_mdot, beta_relax; /* periodicBeta, periodicMdot, relaxBeta */
  int beta_calc, beta_iters, beta_every; /* LcalculateBeta, BetaIterations, nIterCorrectBeta */
  int nperiodic;
  int *ppair;             /* partner patch per patch (-1 none) */
  double *ptrans;         /* x/y/zTranslation per patch */
  /* k-omega SST (TurbulenceModel 'komegasst', ModelNumber 14) with
   * low-Reynolds-number walls (WallBCModel 'menter'); turb = 0 laminar.
   * Cells: k, omega, mu_t, the wall distance, F1/F2, TurbulenceProduction,
   * StrainRate, 2/3 rho k and the effective diffusion coefficient of the
   * equation in assembly (eDiffCoefficient); boundary faces: their B*
   * counterparts.  Gradients as gu (3N / 3Fb / 3Fi); the 2/3 rho k face
   * gradients interpolated 'average' (fgrkb) and 'averagecorrected' (fgrkc). */
  int turb;
  double *tk, *tw, *mut, *wd, *btk, *btw, *bmut;
  double *gk, *gom, *bgk, *bgom, *fgk, *fgom;
  double *rhok, *brhok, *grk, *bgrk, *fgrkb, *fgrkc;
  double *F1, *bF1, *F2, *bF2, *prod, *sr, *bsr, *med, *bmed;
  /* ASSolver, ASIter, HRFramework, ConvectionScheme, nIterGradient (0 =
   * MethodCalcGradient 1) and LimitGradientMethod (0 off) of k and omega */
  int k_solver, k_itmax, w_solver, w_itmax, k_fw, k_scheme, w_fw, w_scheme, k_ngrad, w_ngrad, k_lim, w_lim;
  /* urfTKE, urfTOmega, urfTViscosity, rrFTKE, rrFTOmega, MaximumViscosityRatio */
  double urf_k, urf_w, urf_mut, k_rrf, w_rrf, max_visc_ratio;
  int it_k, it_w;
  double res_k[4], res_w[4];
} or_case;

enum { OR_GI_AVERAGE = 0, OR_GI_UPWIND = 1, OR_GI_DOWNWIND = 2, OR_GI_AVERAGECORRECTED = 3, OR_GI_HARMONIC = 4 };

enum { OR_BC_WALL = 0, OR_BC_SYMMETRY = 1, OR_BC_PERIODIC = 2 };
enum { OR_HR_NONE = 0, OR_HR_TVD = 1, OR_HR_NVF = 2 };
enum {
  OR_CONV_UPWIND = 0, OR_CONV_DOWNWIND, OR_CONV_FROMM, OR_CONV_SOU, OR_CONV_CD, OR_CONV_QUICK,
  OR_CONV_SUPERBEE, OR_CONV_MINMOD, OR_CONV_OSHER, OR_CONV_VANLEER, OR_CONV_MUSCL, OR_CONV_SMART,
  OR_CONV_HQUICK, OR_CONV_UMIST, OR_CONV_CHARM, OR_CONV_BOUNDEDCD, OR_CONV_STOIC, OR_CONV_GAMMA
};

or_case *or_case_new(or_mesh *m, const double *wall_u, const double *wall_v,
                     const double *wall_w, double rho, double mu, double urf_u,
                     double urf_p, int msolver, double mrrf, int mitmax,
                     int psolver, double prrf, int pitmax);
void or_case_delete(or_case *c);
/* Boundary condition of each patch (patches cover the boundary faces in
 * order); returns 0, or -1 for sizes that do not cover the boundary or a
 * symmetry patch p (1-based) with fewer than p faces. */
int or_case_set_patch_types(or_case *c, int npatches, const int *patch_nfaces, const int *patch_type);
void or_case_update_symmetry(or_case *c);
/* Translational periodic pairs after or_case_set_patch_types marked their
 * patches OR_BC_PERIODIC: pair[q] the partner patch of patch q (0-based, -1
 * none), trans[3q..3q+2] its x/y/zTranslation.  Pairs the faces and couples
 * the cells implicitly as RotationMatrixCoefficients does
 * (CalculateRotationMatrixCoefficients.for:171-291; this modifies the mesh's
 * ElementNeighbor), then sets the fixed-bulk-flux controls of
 * updateperiodicBeta (PeriodicBeta.for:4).  Returns -1 when a face finds no
 * partner or a pair is not mutual. */
int or_case_set_periodic(or_case *c, int npatches, const int *pair, const double *trans, double beta0,
                         int calc_beta, double mdot_target, double relax, int iters, int every);
/* Returns -1 for a scheme the framework does not define (or a steady
 * single-phase case cannot use) and for an unknown framework. */
int or_case_set_convection(or_case *c, int framework, int scheme, double bleed, int start_iter);
int or_case_set_gradient_limiters(or_case *c, int momentum_method, int continuity_method);
/* MethodCalcGradient 1 (plain Green-Gauss) or 2 (with nIterGradient
 * face-centroid corrections) for the velocities and for p and p' */
int or_case_set_gradient_method(or_case *c, int mom_method, int mom_iters, int cont_method, int cont_iters);
int or_case_set_gradient_interpolation(or_case *c, int momentum, int continuity);
/* MethodDecomposeS 1, 2 or 3: recomputes the mesh's E/T split */
int or_case_set_decompose(or_case *c, int method);
/* 0 'simple', 1 'simplec' */
int or_case_set_algorithm(or_case *c, int simplec);
double or_tvd_psi(int scheme, double r);
double or_nvf_phif(int scheme, double phic_tilde);
void or_limit_gradient(or_case *c, int method, const double *phi, const double *bphi, double *gr, double *bgr);
void or_hr_correction(or_case *c, const double *phi, const double *gr, double *ftf);
void or_gradient(or_case *c, const double *phi, const double *bphi, double *gr, double *bgr);
void or_case_momentum_gradients(or_case *c);
void or_case_momentum_assemble(or_case *c, int comp);
void or_case_momentum_solve(or_case *c, int comp);
void or_case_pc_assemble(or_case *c);
void or_case_pc_solve_correct(or_case *c);
void or_case_iterate(or_case *c, int niter);
/* The k-omega SST case: initial cell fields k, omega, mu_t and the wall
 * distance, initial boundary-face fields k, omega, mu_t; every outer
 * iteration then ends with CalculateEffectsOfTurbulence.  iopt: k solver,
 * k itmax, omega solver, omega itmax, k framework, k scheme, omega
 * framework, omega scheme, k MethodCalcGradient, k nIterGradient, omega
 * MethodCalcGradient, omega nIterGradient, k limiter, omega limiter (0 off);
 * dopt: urfTKE, urfTOmega, urfTViscosity, rrFTKE, rrFTOmega,
 * MaximumViscosityRatio.  Returns -1 on an invalid option. */
int or_case_set_sst(or_case *c, const int *iopt, const double *dopt, const double *k, const double *omega,
                    const double *mut, const double *wd, const double *bk, const double *bomega,
                    const double *bmut);
/* CalculateEffectsOfTurbulence for the SST case
 * (SolveSelectedTurbulenceModelEquations.for:4-13,150-160). */
void or_case_sst(or_case *c);
/* CalculateNormalDistanceToWall, method 2 (the Poisson wall distance) of the
 * case's no-slip walls into wd (N); phi_out (N, or NULL) the Poisson
 * solution.  Returns the outer passes, -1 without convergence. */
int or_case_wall_distance(or_case *c, double tol, int maxouter, int solver, double rrf, int itmax, double *wd,
                          double *phi_out);
double or_case_mass_imbalance(or_case *c);

#ifdef __cplusplus
}
#endif
#endif
