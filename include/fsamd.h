/*
 * fsamd.h -- C ABI of the MI355X (gfx950) library for the per-iteration hot
 * path of the FlowSolver finite-volume code.
 *
 * Array conventions are the reference's module globals, unchanged, so a
 * Fortran host can pass its own arrays through ISO_C_BINDING:
 *   - cell arrays ac(N), bc(N), FiT(N), dphi(N)            (double)
 *   - ELL arrays anb(N,W), ElementNeighbor(N,W)             (column-major:
 *     element (i,j) at [j*N + i]); neighbour numbers are 1-based, 0 = the slot
 *     of a boundary face (DeclareVariables/Variables2.F90:10-20,
 *     DeclareVariables/Geometry3.F90:34-35).
 *   - default INTEGER is int32; LOGICAL is passed as int32 (0 / nonzero).
 * Strings of the reference (solver names, variable names) become enum codes.
 *
 * Return codes: 0 = converged; 1 = stopped at itmax (the residual is still
 * reported, as the reference does); negative = invalid input or device error
 * (fs_last_error() gives the message): -1 invalid argument, also a diagonal
 * ac <= 0 for a solver that divides by it (Gauss-Seidel, D-ILU, PBCG and
 * every multigrid smoother), -2 device error, -3 unsupported solver, -4 a
 * non-finite ac or bc, or a NaN residual.  A refused input leaves the
 * caller's arrays untouched so the host can fall back to its own solver.
 *
 * Threading: one call in flight per fs_mesh; each mesh owns one HIP stream
 * and every entry point synchronises before it returns.
 */
#ifndef FSAMD_H
#define FSAMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs_mesh fs_mesh;

/* solver codes: the reference's ASSolver* strings (User0.F90:288-301) */
enum fs_solver {
  FS_SOR = 0,    /* 'sor'    symmetric Gauss-Seidel (SORMethod.for:4)        */
  FS_ILU = 1,    /* 'ilu'    D-ILU (FactorizeUsingILU.for:4, ILUMethod.for:4) */
  FS_PBCG = 2,   /* 'pbcg'   ILU-diagonal biconjugate gradient (PBCGMethod.for:96) */
  FS_DIRECT = 3, /* 'direct' dense LU -- not offered on the device             */
  FS_GMRES = 4,  /* 'gmres'  restarted GMRES (matrixSolverOMP.f90:48)          */
  FS_BICGS = 5,  /* 'bicgs'  BiCGSTAB (matrixSolverOMP.f90:306)                */
  FS_CG = 6      /* preconditioned conjugate gradient for symmetric systems
                    (the p' equation) -- no reference counterpart            */
};

/* multigrid modes for the LMultigrid argument */
enum fs_multigrid {
  FS_MG_OFF = 0,     /* no multigrid                                           */
  FS_MG_SOLVER = 1,  /* the reference's AlgebraicMultigrid: V-cycles as the
                        solver, smoother chosen by the solver code
                        (AlgebraicMultiGridMainRoutine.for:4)                 */
  FS_MG_PRECOND = 2  /* one AMG V-cycle as the preconditioner of the Krylov
                        solver (PBCG, GMRES, BiCGSTAB) -- no reference
                        counterpart                                           */
};

/* option keys for fs_set_option (defaults in brackets) */
enum fs_option {
  FS_OPT_MG_PRESWEEP = 1,     /* MultiGridpreSweep [2 on device]                */
  FS_OPT_MG_POSTSWEEP = 2,    /* MultiGridpostSweep [2 on device]               */
  FS_OPT_MG_MAXCYCLES = 3,    /* MaxMultiGridCycles [30]                        */
  FS_OPT_MG_MAXLEVELS = 4,    /* maxNumberofCoarseLevels [10]                   */
  FS_OPT_MG_MINPARENTS = 5,   /* minNumberofParents [10]                        */
  FS_OPT_MG_REAGGLOMERATE = 6,/* reAgglomorate: rebuild the aggregates every n
                                 calls; otherwise only the Galerkin operators
                                 are recomputed [300]                          */
  FS_OPT_MG_FP32 = 7,         /* 1 = hierarchy and smoother in FP32 inside the
                                 FP64 Krylov loop [0]                          */
  FS_OPT_MG_OMEGA = 8,        /* damping of the device Jacobi smoother [0.8]    */
  FS_OPT_MG_SMOOTHER = 9,     /* 0 = damped Jacobi, 1 = multicolour symmetric
                                 Gauss-Seidel on levels with <= 4 colours [0]  */
  FS_OPT_GMRES_RESTART = 10,  /* GMRES restart length [30]                      */
  FS_OPT_CHECK_EVERY = 11,    /* iterations launched between host checks [8];
                                 a solve's first batch is the iteration count
                                 of the previous solve of the same variable
                                 and solver (at most 4x this), later batches
                                 grow 2, 4, ... up to this                     */
  FS_OPT_MG_TAIL_CELLS = 12,  /* levels with at most this many cells run in one
                                 single-workgroup kernel [1024]                 */
  FS_OPT_MG_COARSE_SWEEPS = 13,/* smoothing sweeps on a coarsest level too large
                                 for the direct FP64 solve (> 64 cells) [16]   */
  FS_OPT_MG_ALPHA = 14,       /* over-correction of the prolongated coarse
                                 correction; 0 = 1.0 with multigrid as the
                                 solver (the reference), 1.7 as a
                                 preconditioner [0]                           */
  FS_OPT_MG_COARSE_PRESWEEP = 15, /* pre-sweeps on levels >= 1 (-1: as level 0) [-1] */
  FS_OPT_MG_COARSE_POSTSWEEP = 16, /* post-sweeps on levels >= 1 (-1: as level 0) [-1] */
  FS_OPT_MG_SELL_ROWS = 17,   /* coarse levels with at least this many rows keep
                                 their operator in sliced ELL (64-row slices,
                                 rows sorted by length within 256-row windows,
                                 one row per lane); 0 = CSR everywhere [100000] */
  FS_OPT_MG_AGGREGATION = 18, /* 0 = the reference's greedy agglomeration
                                 (AgglomorateALevelAlgebraicly, bitwise its
                                 maps); 1 = pairwise matching of each cell with
                                 its strongest negative coupling, 2 = two
                                 pairwise passes per level (aggregates of up to
                                 4 cells) -- no reference counterpart [0]      */
  FS_OPT_MG_FUSE_PROLONG = 19, /* 1 = the prolongation of the coarse correction
                                 fused into the first post-sweep of grid-wide
                                 levels, 0 = a pass of its own before it [1]   */
  FS_OPT_WARM_PREVIOUS = 20   /* device SIMPLE case (fsamd_simple.h) without a
                                 warm-start network: 1 = start each p' solve
                                 from the previous iteration's p', scaled to
                                 its best multiple and dropped when worse than
                                 the zero start; 2..4 = from the combination of
                                 the last 2..4 p' fields with the smallest L2
                                 residual (projection), scaled and dropped the
                                 same way [0]                                  */,
  FS_OPT_MG_CYCLE = 21        /* device AMG cycle: 0 = V, 1 = F, 2 = W in the
                                 pattern of MultiGridCycle.for:84-127, 162-220
                                 (back up to level l and down again, for l from
                                 the deepest grid-wide level up to 1; W once
                                 more from 2 down), every coarser level of a
                                 descent starting from zero; the one-workgroup
                                 tail counts as one level [0]                  */
};

/* Statistics of the last solve. */
typedef struct fs_solve_stats {
  int iterations;       /* Krylov / stationary iterations or MG cycles       */
  int levels;           /* AMG levels in use (1 = no hierarchy)              */
  double initial_l1;    /* sum |bc - A*0|                                    */
  double final_l1;      /* sum |bc - A*dphi| (true residual)                 */
  double seconds;       /* wall time of the solve on the device              */
} fs_solve_stats;

/* Topology of the mesh (ReadFoamMesh.f90:288-619; Geometry3.F90:21-37).
 * elem_nbr: ElementNeighbor(N,W) column-major, 1-based, 0 = boundary slot.
 * n_nbr   : NumberofElementNeighbors(N).
 * The face arrays may be NULL when only the linear-solve entry points are
 * used: owner/neighbour = NIFaceOwner/NIFaceNeighbor(Fi), i_own_nbr /
 * i_nbr_own = iOwnerNeighbor/iNeighborOwner(Fi) (1-based slots). */
int fs_mesh_create(int n_cells, int width, const int *elem_nbr,
                   const int *n_nbr, int n_ifaces, const int *owner,
                   const int *neighbour, const int *i_own_nbr,
                   const int *i_nbr_own, fs_mesh **out);
int fs_mesh_destroy(fs_mesh *mesh);
int fs_set_option(fs_mesh *mesh, int key, double value);

/* Drop-in for SUBROUTINE SolveEquation(Variable,FiT,rrF,itmax,solver,
 * LMultigrid) (SolveEquations/AlgebraicSolvers/SolveAlegbraicEquations.for:4)
 * with its module inputs ac, anb, bc (Variables2) passed explicitly.
 * nf is the GetNFfromVariable code (AssignNFvaluesToVariables.for:4; 4 = p').
 * On return fit += dphi, dphi holds the correction, res4 = ResorAbs, ResorMax,
 * ResorRMS, ResorScaled of the initial residual (CalculateResidualsRoutine.for
 * :19-67).  Host arrays; copies in and out over PCIe. */
int fs_solve_equation(fs_mesh *mesh, int nf, double *fit, double rrf,
                      int itmax, int solver, int lmultigrid, const double *ac,
                      const double *anb, const double *bc, double *dphi,
                      double *res4, fs_solve_stats *stats);

/* The same solve on device-resident arrays (hipMalloc'ed, same layouts). */
int fs_solve_equation_device(fs_mesh *mesh, int nf, double *d_fit, double rrf,
                             int itmax, int solver, int lmultigrid,
                             const double *d_ac, const double *d_anb,
                             const double *d_bc, double *d_dphi, double *res4,
                             fs_solve_stats *stats);

/* The same device solve started from the initial guess the caller leaves in
 * d_dphi instead of from zero.  No counterpart in the reference (its
 * SolveEquation always starts at dphi = 0, SolveAlegbraicEquations.for:47);
 * it is the entry of a predicted p' (the learned warm start).  Every solver
 * first scales the guess to its multiple with the smallest L2 residual and
 * drops it for the zero start when its L1 residual still exceeds |bc|_1 (a
 * non-finite guess is always dropped); the residual norms and the stop rule
 * are those of the zero start, so only the iteration count can change. */
int fs_solve_equation_device_from_guess(fs_mesh *mesh, int nf, double *d_fit, double rrf,
                                        int itmax, int solver, int lmultigrid,
                                        const double *d_ac, const double *d_anb,
                                        const double *d_bc, double *d_dphi, double *res4,
                                        fs_solve_stats *stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the solve of fs_solve_equation_device, same arguments
 * and results, with lmultigrid 0 and solver SOR, PBCG, BICGS or CG, whose work
 * vectors, D-ILU diagonal, reduction partials, scalar slots and flags live in
 * one static device array of the library instead of the mesh's buffers.  The
 * call makes no hipMalloc/hipFree.  Limits: at most FS_NOALLOC_MAX_CELLS
 * cells (larger meshes are refused with -1); the array takes
 * FS_NOALLOC_SCRATCH_BYTES on every GPU the library runs on, until the
 * process ends; the mesh's device must be current.  All calls on one GPU
 * share that array (it keeps the last call's contents; everything a solve
 * reads is initialised by the solve), so calls that overlap on one GPU race. */
#define FS_NOALLOC_MAX_CELLS 524288
#define FS_NOALLOC_SCRATCH_BYTES 38077440
int fs_solve_equation_device_noalloc(fs_mesh *mesh, int nf, double *d_fit, double rrf,
                                     int itmax, int solver, int lmultigrid,
                                     const double *d_ac, const double *d_anb,
                                     const double *d_bc, double *d_dphi, double *res4,
                                     fs_solve_stats *stats);

/* Cell-to-face map NGlobalEFaces(N,W) (column-major, 1-based global face
 * numbers: interior faces 1..Fi first, then boundary faces patch by patch)
 * with NumbOfElementFaces(N) (ReadFoamMesh.f90:329-342; Geometry1.F90:14).
 * Needed by the face-flux assembly; slot j of a cell must be the face across
 * which ElementNeighbor(cell,j) lies.  Requires the interior owner arrays of
 * fs_mesh_create. */
int fs_mesh_set_cell_faces(fs_mesh *mesh, int n_faces_total,
                           const int *global_efaces, const int *n_efaces);

/* Scatter of face fluxes into the ELL system, AssembleGlobalMatrixFaceFluxes
 * and AssembleGlobalMatrixElementFluxes (Assemble/
 * AssembleGolbalElementAndFaceFluxes.for:43-125, :4-39), as a deterministic
 * per-cell gather.  flux_* are FluxCf/FluxFf/FluxTf(NFacesTotal) with the
 * interior faces first (Variables2.F90:4-7); only interior faces and the
 * boundary faces listed in the mesh are used; flux_ce/flux_te are
 * FluxCE/FluxTE(N) (Variables3.F90:4-6).  ac/anb/bc are accumulated into
 * (the reference zeroes them in InitializeCoefficients before assembly).
 * Device pointers. */
int fs_assemble_face_fluxes_device(fs_mesh *mesh, const double *d_flux_cf,
                                   const double *d_flux_ff,
                                   const double *d_flux_tf,
                                   const double *d_flux_ce,
                                   const double *d_flux_te, double *d_ac,
                                   double *d_anb, double *d_bc);

/* Inspection of the AMG hierarchy held by the mesh (after a multigrid
 * solve): dims = {n, w, total children, levels, colours} of level `level`;
 * maps in the reference's numbering (1-based): parent (this level -> next),
 * seed/nchild/children of this level's aggregates (ElementParent,
 * NumberOfChildren, Children), nbr/nnbr (ElementNeighborMG,
 * NumberofElementNeighborsMG), ac/anb (acMG/anbMG of the last solve).
 * Any pointer may be NULL.  (AgglomorateALevelAlgebraicly.for:48-331) */
int fs_amg_export(fs_mesh *mesh, int level, int *dims, int *parent, int *seed,
                  int *nchild, int *child_ptr, int *child, int *nbr,
                  int *nnbr, double *ac, double *anb);

/* Average ms of one plain device-to-device copy of `bytes` (a multiple of
 * 16) over `reps` launches of a streaming copy kernel on a private stream.
 * No counterpart in the reference: it calibrates HBM byte counters
 * (FETCH_SIZE / WRITE_SIZE) and gives the achievable copy bandwidth. */
int fs_time_device_copy(long long bytes, int reps, double *avg_ms);

/* Per-launch timing of kernels of the solve loops, for the roofline figure:
 * each launch of an armed site records two HIP events on the mesh's stream
 * from its own dispatch (hipExtLaunchKernelGGL).  No counterpart in the
 * reference.  fs_kernel_clock_start creates `capacity` event pairs (outside
 * any timed region) and arms the sites of site_mask (bit 1 << FS_CLOCK_*; 0
 * disarms).  fs_kernel_clock_read synchronises the stream and returns for one
 * site out[0] launches that did work (launches of Krylov iterations after
 * convergence, which return at once, are left out), out[1] their mean
 * duration in ms, out[2] their mean algorithmic bytes, out[3] launches of the
 * site not timed (idle or beyond the capacity).  fs_kernel_clock_clear
 * empties the records; the sites stay armed.  fs_kernel_clock_stride times
 * only every stride-th launch of each site (default 1: every launch), which
 * keeps the events' own cost out of the throughput of the timed region. */
enum fs_clock_site {
  FS_CLOCK_OFF = 0,
  FS_CLOCK_SPMV = 1,      /* k_spmv_dot, the FP64 operator product of the
                             BiCGSTAB / CG loops                              */
  FS_CLOCK_FINE_SWEEP = 2 /* k_sweep on the finest AMG level                 */
};
int fs_kernel_clock_start(fs_mesh *mesh, int site_mask, int capacity);
int fs_kernel_clock_read(fs_mesh *mesh, int site, double *out);
int fs_kernel_clock_clear(fs_mesh *mesh);
int fs_kernel_clock_stride(fs_mesh *mesh, int stride);

/* Message of the last failing call on this thread. */
const char *fs_last_error(void);
/* Library version string. */
const char *fs_version(void);

#ifdef __cplusplus
}
#endif
#endif
