// Entry points shared between the solver translation units.
#pragma once

#include <vector>

#include "fs_common.h"

namespace fs {

int mesh_prepare_solver(fs_mesh *m);
void mesh_release_solver(fs_mesh *m);
void greedy_colors(int n, int w, const int *nbr0, std::vector<int> &ptr, std::vector<int> &cells);
void add_inplace(fs_mesh *m, double *a, const double *b);
// one-block convergence update from np L1 partials (iteration count, done)
void launch_fin_conv(fs_mesh *m, int np);

int solve_device(fs_mesh *m, int nf, double *fit, double rrf, int itmax, int solver, int lmg,
                 const double *ac, const double *anb, const double *bc, double *dphi, double *res4,
                 fs_solve_stats *st);

// AMG (fs_amg.hip).  amg_prepare builds or refreshes the hierarchy for the
// fine operator (ac, anb): aggregation on the first call and every
// mg_reagg calls, device Galerkin operators otherwise.
void amg_prepare(fs_mesh *m, const double *ac, const double *anb);
// out = one V-cycle applied to rhs (both FP64 device vectors).
void amg_apply(fs_mesh *m, const double *rhs, double *out, const int *flags);
// The reference's multigrid-as-solver loop; returns cycles run.
int amg_solve(fs_mesh *m, const double *ac, const double *anb, const double *bc, double *dphi,
              int solver);
int amg_levels(fs_mesh *m);
// Average ms of one k_spmv_dot launch (y = A x with (dot_with, y) partials).
int time_spmv(fs_mesh *m, const double *ac, const double *anb, const double *x, const double *dot_with,
              double *y, int reps, double *avg_ms);
// Host copies of the aggregation maps of level l -> l+1 for tests.
int amg_export(fs_mesh *m, int level, int *dims, int *parent, int *seed, int *nchild,
               int *child_ptr, int *child, int *nbr, int *nnbr, double *ac, double *anb);

} // namespace fs
