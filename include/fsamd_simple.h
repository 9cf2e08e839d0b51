/*
 * fsamd_simple.h -- C ABI of the device-resident SIMPLE outer iteration.
 *
 * The reference runs one steady SIMPLE outer iteration per pass of its main
 * loop (MainProgram/MainProgram.for:257-348): SolveMomentumX/Y/Z
 * (SolveEquations/Momentum/SolveXMomentumEquation.for:4 and twins), then
 * SolveContinuity (SolveEquations/ContinuityEquation/
 * SolveContinuityEquation.for:4).  fs_case_iterate replaces that pass for a
 * laminar, incompressible, constant-property flow bounded by no-slip walls
 * (moving or at rest), with every array resident in HBM:
 *   gradients  Green-Gauss + 2 face-centroid corrections
 *              (GradientFiles/StandardGreenGaussCorrection1Gradient.for:3)
 *   face grads 'average' (velocity) / 'averagecorrected' (pressure)
 *              (Interpolation/InterpolateGradientFromElementToFace.for:3)
 *   momentum   diffusion with E/T split, explicit and bulk-viscosity terms,
 *              upwind convection, effective divergence, pressure gradient,
 *              under-relaxation, D = V/aP
 *              (Diffusion/AssembleIsotropicDiffusion.for:3,
 *               Diffusion/AssembleExplicitDiffusion.for:3,
 *               Diffusion/AssembleBulkViscosity.for:3,
 *               Convection/AssembleConvection.for:4,
 *               Momentum/AssemblePressureGradient.for:4,
 *               UnderRelaxVariables/UnderRelaxation.for:4,
 *               Momentum/StoreTheDValues.for:4)
 *   continuity Rhie-Chow mass flux, p' coefficients, p' solve, pressure,
 *              velocity and mass-flux corrections, wall pressure
 *              (ContinuityEquation/AssembleMdotUsingRhieChow.for:4,
 *               AssemblePressureCorrectionFluxes.for:4, CorrectPressure.for:4,
 *               CorrectMdotandVelocity.for:4,722)
 * with the reference's settings MethodCalcGradient = 2, nIterGradient = 2,
 * MethodDecomposeS = 3, LRelaxMomentum (User0.F90), and by default
 * Algorithm 'simple' and upwind convection; fs_case_set_algorithm,
 * fs_case_set_convection and fs_case_set_gradient_limiters select SIMPLEC,
 * the TVD/NVF high-resolution schemes and the gradient limiters.
 *
 * The mesh is a polyMesh in memory (0-based): face_ptr/face_nodes list each
 * face's nodes, faces 0..nif-1 are interior (owner < neighbour), the rest are
 * boundary faces patch by patch.  Topology and geometry follow the
 * reference's Readit (GridAndGeometry/ReadFoamMesh.f90:254-619) and
 * ProcessGeometry (GridAndGeometry/ProcessGeometryRoutine.for:3).
 */
#ifndef FSAMD_SIMPLE_H
#define FSAMD_SIMPLE_H

#include "fsamd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs_case fs_case;

typedef struct fs_simple_stats {
  int iterations;            /* SIMPLE outer iterations run                  */
  int momentum_iterations;   /* summed over u, v, w solves                   */
  int pc_iterations;         /* summed p' solver iterations                  */
  int pc_unconverged;        /* p' solves that stopped at itmax              */
  int pc_levels;             /* AMG levels of the last p' solve              */
  double seconds;            /* wall time of the iterations                  */
  double momentum_seconds;
  double pc_seconds;
  double pc_initial_l1;      /* last p' solve: sum|bc|                       */
  double pc_final_l1;        /* last p' solve: sum|bc - A p'|                */
  double res_momentum[3];    /* last ResorAbs of u, v, w                     */
  double mass_l1;            /* sum over cells of |net mass outflow| at end  */
  int turbulence_iterations; /* summed k and omega solver iterations         */
} fs_simple_stats;

enum fs_phase {
  FS_PHASE_MOMENTUM_GRADIENTS = 0,
  FS_PHASE_ASSEMBLE_U = 1,
  FS_PHASE_ASSEMBLE_V = 2,
  FS_PHASE_ASSEMBLE_W = 3,
  FS_PHASE_SOLVE_U = 4,
  FS_PHASE_SOLVE_V = 5,
  FS_PHASE_SOLVE_W = 6,
  FS_PHASE_PC_ASSEMBLE = 7,
  FS_PHASE_PC_SOLVE_CORRECT = 8,
  FS_PHASE_UPDATE_SYMMETRY = 9,
  FS_PHASE_TURBULENCE = 10    /* CalculateEffectsOfTurbulence (SST)        */
};

/* wall_u/v/w: velocity of every boundary face (nfaces - nif entries). */
int fs_case_create(int npoints, const double *x, const double *y, const double *z,
                   int nfaces, const int *face_ptr, const int *face_nodes,
                   const int *owner, int nif, const int *neighbour,
                   const double *wall_u, const double *wall_v,
                   const double *wall_w, double rho, double mu, fs_case **out);
int fs_case_destroy(fs_case *c);

/* Boundary condition types (BCType of the reference's case definition). */
enum fs_bc_type { FS_BC_WALL = 0, FS_BC_SYMMETRY = 1, FS_BC_PERIODIC = 2 };
/* Boundary condition of each patch, the patches covering the boundary faces
 * in order (npatches sizes summing to NFacesTotal - NIFaces).  A case starts
 * with every boundary face a no-slip wall of the velocity given at creation.
 * A symmetry face takes the owner's velocity less its normal component
 * (UpdateSymmetry, SymmetryAndAxisBoundary.for:4-44; applied before the next
 * iteration and after every pressure correction), slip diffusion
 * (AssembleIsotropicDiffusion.for:562-620) and the symmetry projection of the
 * boundary pressure gradient (CorrectGradients.for:244-289, which reads the
 * z-normal of face i of patch i: a symmetry patch p, counted from 1, needs
 * at least p faces).  Fails for sizes that do not cover the boundary. */
int fs_case_set_patch_types(fs_case *c, int npatches, const int *patch_nfaces,
                            const int *patch_type);
/* Translational periodicity with the implicit coupling (the case file's
 * LTranslationalPeriodicity and LPeriodicImplicit, BoundaryConditions2.F90
 * :130-146, set up by RotationMatrixCoefficients,
 * CalculateRotationMatrixCoefficients.for:171-291) on patches that
 * fs_case_set_patch_types marked FS_BC_PERIODIC: pair[q] is the partner
 * patch of patch q (0-based, -1 for none; mutual, of equal size),
 * translation[3q..3q+2] its x/y/zTranslation (the vector that carries the
 * partner patch's cells across patch q).  Each face of the lower patch of a
 * pair takes the partner face whose centroid offset is closest to the
 * translation; the partner cell joins the face's slot of ElementNeighbor
 * (every solver then couples the pair) and the periodic branches of the
 * gradients, diffusion, convection, Rhie-Chow, p' coefficients and
 * corrections run (PeriodicBoundary.for, CorrectGradients.for:66-167,
 * AssembleIsotropicDiffusion.for:630-857, AssembleConvection.for:582-627,
 * AssembleMdotUsingRhieChow.for:1086-1416,
 * AssemblePressureCorrectionFluxes.for:748-830, CorrectMdotandVelocity.for
 * :115-145, CorrectPressure.for:293-317).  The momentum equations take the
 * pressure-gradient source -beta V along the translation
 * (AssembleLinearPressure.for:4) with beta = periodicBeta, initially `beta`;
 * calc_beta (LcalculateBeta) updates it every `every` iterations
 * (nIterCorrectBeta) towards the mass flow `mdot_target` (periodicMdot,
 * counted outward through the lower periodic patch) with relaxation `relax`
 * and `iters` passes (relaxBeta, BetaIterations; updateperiodicBeta,
 * PeriodicBeta.for:4).  Once per case, after fs_case_set_patch_types. */
int fs_case_set_periodic(fs_case *c, int npatches, const int *patch_nfaces, const int *pair,
                         const double *translation, double beta, int calc_beta,
                         double mdot_target, double relax, int iters, int every);
/* The k-omega SST model (TurbulenceModel 'komegasst', ModelNumber 14;
 * SolveSelectedTurbulenceModelEquations.for:150-160) with low-Reynolds-number
 * walls (WallTreatment 'lowreynoldsnumber', WallBCModel 'menter'): every
 * outer iteration then ends with CalculateEffectsOfTurbulence (strain rate,
 * the k and omega equations, mu_t), and the momentum equations take
 * Viscosity + TurbulentViscosity and the 2/3 rho k gradient
 * (MinusTwoThirdsRhoK, with its difference in the Rhie-Chow flux).
 * iopt[14]: k solver, k itmax, omega solver, omega itmax (FS_* solver
 * codes), k HR framework, k scheme, omega framework, omega scheme (as
 * fs_case_set_convection), k MethodCalcGradient (1 or 2), k nIterGradient,
 * omega MethodCalcGradient, omega nIterGradient, k limiter method, omega
 * limiter method (0 off, 1, 2).  dopt[6]: urfTKE, urfTOmega, urfTViscosity,
 * rrFTKE, rrFTOmega, MaximumViscosityRatio.  Initial fields (host arrays):
 * k, omega, mu_t and the wall distance of every cell; k, omega, mu_t of every
 * boundary face.  Fields by name afterwards: k omega mut wd bk bomega bmut F1
 * F2.  wall_distance may be NULL after fs_case_wall_distance: the computed
 * distance is used.  Returns -1 on an invalid option, -3 for a solver code
 * the k and omega solves cannot take. */
/* The distance of every cell from the no-slip walls of the case's patch
 * types by the reference's Poisson method (CalculateNormalDistanceToWall,
 * MethodCalculateNormalDistance 2, CalculateDistanceToWall.for:154-470):
 * lap(phi) = -1, phi = 0 at the walls, no flux elsewhere, passes of
 * assembly and CG + AMG solve until a pass starts from at most tol times the
 * first pass's residual |bc|_1; d = -|grad phi| + sqrt(|grad phi|^2 + 2 phi),
 * the cells at a wall their face's normal distance.  Kept on the device for
 * fs_case_set_sst; copied to wd (N, host) unless NULL; passes: the passes
 * run.  Returns 0, 1 when maxouter passes did not reach tol, < 0 on error. */
int fs_case_wall_distance(fs_case *c, double tol, int maxouter, double *wd, int *passes);
int fs_case_set_sst(fs_case *c, const int *iopt, const double *dopt, const double *k,
                    const double *omega, const double *mut, const double *wall_distance,
                    const double *bk, const double *bomega, const double *bmut);
/* High-resolution convection of the momentum equations: HRFrameworkMomentum
 * ('none', 'tvd', 'nvf'), ConvectionSchemeMomentum, BleedMomentum and
 * nIterStartApplyingHR (User0.F90:367-410).  Upwind fluxes stay implicit and
 * mdot (phi_HR - phi_C) of every interior face joins FluxTf as a deferred
 * correction (AssembleConvection.for:156-165,
 * AssembleHRSchemeCorrection.for:4-133) from the outer iteration numbered
 * start_iter on (iterations count from 1 over the case's life).  A case
 * starts with upwind convection.  Fails for a scheme the framework does not
 * define: TVD (LimitersTVDType.for:4) defines upwind .. charm, NVF
 * (LimitersNVFType.for:4) upwind, smart, minmod, boundedcd, cd, osher, stoic,
 * muscl, superbee, downwind and gamma; the volume-fraction schemes stacs,
 * hric and cicsam are not provided. */
enum fs_hr_framework { FS_HR_NONE = 0, FS_HR_TVD = 1, FS_HR_NVF = 2 };
enum fs_conv_scheme {
  FS_CONV_UPWIND = 0, FS_CONV_DOWNWIND = 1, FS_CONV_FROMM = 2, FS_CONV_SOU = 3, FS_CONV_CD = 4,
  FS_CONV_QUICK = 5, FS_CONV_SUPERBEE = 6, FS_CONV_MINMOD = 7, FS_CONV_OSHER = 8, FS_CONV_VANLEER = 9,
  FS_CONV_MUSCL = 10, FS_CONV_SMART = 11, FS_CONV_HQUICK = 12, FS_CONV_UMIST = 13, FS_CONV_CHARM = 14,
  FS_CONV_BOUNDEDCD = 15, FS_CONV_STOIC = 16, FS_CONV_GAMMA = 17
};
int fs_case_set_convection(fs_case *c, int framework, int scheme, double bleed, int start_iter);
/* Gradient limiters (LimitGradientVariable, LimitGradients.for:3):
 * momentum_method limits the u, v, w gradients, continuity_method those of
 * p and p' (LimitGradientMomentum/ContinuityMethod, User0.F90:132-149);
 * 0 = off (the reference's default), 1 = min/max limiter with rk = 1
 * (:68-262), 2 = the smooth limiter with eps = 1e-3 of the field's range
 * (:263-437).  The limited gradient replaces the Green-Gauss one before the
 * boundary corrections, and the boundary gradients become the owners'. */
int fs_case_set_gradient_limiters(fs_case *c, int momentum_method, int continuity_method);
/* Algorithm (User0.F90:33): SIMPLE (the default) or SIMPLEC, whose p'
 * coefficients and velocity correction take D2 = V/(a_P + sum anb) while the
 * Rhie-Chow flux keeps D1 = V/a_P (StoreTheDValues.for:20-39,
 * AssemblePressureCorrectionFluxes.for:81,
 * CorrectMdotandVelocity.for:774-778).  Fields "du2" "dv2" "dw2" hold D2. */
enum fs_algorithm { FS_ALGORITHM_SIMPLE = 0, FS_ALGORITHM_SIMPLEC = 1 };
/* MethodCalcGradientMomentum/Continuity and nIterGradientMomentum/Continuity
 * (User0.F90:98-115; CalculateGradients.for:212-235): method 1 is plain
 * Green-Gauss (StandardGreenGaussGradient.for:3), method 2 adds iters
 * face-centroid correction sweeps (StandardGreenGaussCorrection1Gradient.for
 * :3); the momentum pair applies to u, v, w, the continuity pair to p and
 * p'.  A case starts with 2, 2 for both.  Methods 3-6 are not provided. */
int fs_case_set_gradient_method(fs_case *c, int momentum_method, int momentum_iters, int continuity_method,
                                int continuity_iters);
/* MethodDecomposeS (User0.F90:70): the split of each face's area vector into
 * E (along the centroid line) and T, 1 minimum correction, 2 orthogonal
 * correction, 3 over-relaxed (the default); the geometry's E, T and gDiff
 * are recomputed (ProcessGeometryRoutine.for:428-575) and the Rhie-Chow and
 * p' coefficients split D S the same way (AssembleMdotUsingRhieChow.for
 * :575-615, AssemblePressureCorrectionFluxes.for:99-141). */
int fs_case_set_decompose(fs_case *c, int method);
/* GradientInterpolationSchemeMomentum / Continuity (User0.F90:419-420): the
 * face gradients of u, v, w (diffusion, explicit and bulk terms) and of p
 * (Rhie-Chow), InterpolateGradientFromElementToFace.for:16-108.  A case
 * starts with average / averagecorrected. */
enum fs_gradient_interpolation {
  FS_GRADINT_AVERAGE = 0, FS_GRADINT_UPWIND = 1, FS_GRADINT_DOWNWIND = 2, FS_GRADINT_AVERAGECORRECTED = 3,
  FS_GRADINT_HARMONIC = 4
};
int fs_case_set_gradient_interpolation(fs_case *c, int momentum, int continuity);
/* The reference's restart file (writeSolution, WriteSolutionToContinue.for:3,
 * the records a steady, laminar, incompressible flow writes): a Fortran
 * unformatted sequential file of the state after the last iteration, its
 * boundary arrays laid out (NumberOfBCSets, NBFacesMax) for the patch sizes
 * given.  It holds face gradients and the effective divergence that an
 * iteration forms but does not keep: fs_case_set_restart_capture(c, 1)
 * keeps them (9 + 3 face arrays and one cell array), and the last iteration
 * before fs_case_write_solution must run with it on. */
int fs_case_set_restart_capture(fs_case *c, int on);
int fs_case_write_solution(fs_case *c, const char *path, int npatches, const int *patch_nfaces);
/* Resume from such a file (readSolution, WriteSolutionToContinue.for): the
 * velocities, pressure, mass fluxes and boundary values of u, v, w and p;
 * everything else is formed again by the next iteration.  Fails unless the
 * file has the steady laminar layout for this mesh and these patches. */
int fs_case_read_solution(fs_case *c, const char *path, int npatches, const int *patch_nfaces);
int fs_case_set_algorithm(fs_case *c, int algorithm);
/* urfMomentum, urfPressure, ASSolverMomentum/rrFMomentum/ASIterMomentum/
 * LMultigridMomentum and the Continuity equivalents (User0.F90:271-318). */
int fs_case_configure(fs_case *c, double urf_u, double urf_p, int msolver,
                      double mrrf, int mitmax, int mlmg, int psolver,
                      double prrf, int pitmax, int plmg);
/* The linear-solve context of the case (for fs_set_option). */
fs_mesh *fs_case_mesh(fs_case *c);
int fs_case_iterate(fs_case *c, int niter, fs_simple_stats *stats);
/* One phase of the iteration, for tests (enum fs_phase). */
int fs_case_phase(fs_case *c, int phase);
/* Host copies of device fields by name (u v w p pc mdot gu gv gw gp gpc bu bv
 * bw bp bgp du dv dw us vs ws ac anb bc cf ff tf and geometry vol Sx Sy
 * Sz g gdiff Tx Ty Tz dcf A kx ky kz); count must equal the field length. */
int fs_case_get(fs_case *c, const char *name, double *out, long long count);
/* Setting u, v or w re-derives the symmetry boundary velocities before the
 * next iteration, as the reference's start-up does. */
int fs_case_set(fs_case *c, const char *name, const double *in, long long count);
/* Average duration (ms) of the ELL product y = A*x with the last assembled
 * system (the p' operator after an iteration), reps launches timed with HIP
 * events on the case's stream. */
int fs_case_time_spmv(fs_case *c, int reps, double *avg_ms);
/* dims = {N, W, NFacesTotal, NIFaces} */
int fs_case_dims(fs_case *c, int *dims);
/* Topology arrays in the reference's numbering: which = 0 NumbOfElementFaces,
 * 1 NGlobalEFaces, 2 ElementNeighbor, 3 iOwnerNeighbor, 4 iNeighborOwner. */
int fs_case_topology(fs_case *c, int which, int *out, long long count);

/* Host-only topology and geometry (no device needed): the same arrays the
 * case uploads.  Face fields over NFacesTotal (interior first): Sx Sy Sz A
 * fcx fcy fcz nx ny nz dcf dx dy dz ex ey ez g Ex Ey Ez E gdiff Tx Ty Tz T;
 * cell fields: vol xc yc zc (ProcessGeometryRoutine.for:76-499). */
typedef struct fs_geom fs_geom;
int fs_geom_create(int npoints, const double *x, const double *y, const double *z,
                   int nfaces, const int *face_ptr, const int *face_nodes,
                   const int *owner, int nif, const int *neighbour, fs_geom **out);
int fs_geom_destroy(fs_geom *g);
int fs_geom_dims(fs_geom *g, int *dims);
int fs_geom_get(fs_geom *g, const char *name, double *out, long long count);
int fs_geom_topology(fs_geom *g, int which, int *out, long long count);

/* An OpenFOAM polyMesh directory (ASCII points, faces, owner, neighbour,
 * boundary), the on-disk input of the reference's Readit
 * (GridAndGeometry/ReadFoamMesh.f90:31-230).  counts = {points, faces,
 * face-node entries, interior faces, patches}.  fs_polymesh_get copies the
 * arrays in the layout fs_case_create takes (0-based; face_ptr has
 * faces + 1 entries) and the patch names, name_len bytes each. */
typedef struct fs_polymesh fs_polymesh;
int fs_polymesh_read(const char *dir, int *counts, fs_polymesh **out);
int fs_polymesh_get(fs_polymesh *m, double *x, double *y, double *z, int *face_ptr, int *face_nodes,
                    int *owner, int *neighbour, int *patch_start, int *patch_nfaces, char *patch_names,
                    int name_len);
int fs_polymesh_destroy(fs_polymesh *m);

#ifdef __cplusplus
}
#endif
#endif
