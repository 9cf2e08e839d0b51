// Entry points shared between the solver translation units.
#pragma once

#include <hip/hip_ext.h>

#include <vector>

#include "fs_common.h"

namespace fs {

// Launch through the mesh's kernel clock when `site` is the armed one.
template <class K, class... A>
void timed_launch(fs_mesh *m, int site, double bytes, K kernel, dim3 grid, dim3 block, unsigned lds,
                  hipStream_t s, A... args) {
  KernelClock &c = m->kclock;
  if (site != 0 && (c.sites >> site & 1) && c.seen[site & 3]++ % c.stride == 0) {
    if (c.used < c.cap) {
      const int u = c.used++;
      c.iter[u] = c.cur_iter;
      c.site[u] = site;
      c.bytes[u] = bytes;
      c.live[u] = 1;
      hipExtLaunchKernelGGL(kernel, grid, block, lds, s, c.ev[2 * u], c.ev[2 * u + 1], 0u, args...);
      return;
    }
    c.dropped[site & 3]++;
  }
  hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}
int kernel_clock_start(fs_mesh *m, int site_mask, int capacity);
int kernel_clock_read(fs_mesh *m, int site, double *out);
int kernel_clock_clear(fs_mesh *m);
int kernel_clock_stride(fs_mesh *m, int stride);
void kernel_clock_release(fs_mesh *m);

int mesh_prepare_solver(fs_mesh *m);
void mesh_release_solver(fs_mesh *m);
void greedy_colors(int n, int w, const int *nbr0, std::vector<int> &ptr, std::vector<int> &cells);
void add_inplace(fs_mesh *m, double *a, const double *b);
// one-block convergence update from np L1 partials (iteration count, done)
void launch_fin_conv(fs_mesh *m, int np);

// warm: dphi holds an initial guess, which every solver starts from after it
// is scaled to its smallest-L2-residual multiple (and dropped for the zero
// start when its L1 residual still exceeds |bc|_1); otherwise dphi starts
// at zero (:47)
int solve_device(fs_mesh *m, int nf, double *fit, double rrf, int itmax, int solver, int lmg,
                 const double *ac, const double *anb, const double *bc, double *dphi, double *res4,
                 fs_solve_stats *st, bool warm = false);

// x0 = the combination of the k (at most kProjMax) device vectors X[0..k-1]
// with the smallest L2 residual |bc - A x0|; vectors that add nothing to the
// span of those before them get weight 0.  Uses wk.v[0..k-1] and the
// partials; synchronises the stream once.  Returns the vectors used.
constexpr int kProjMax = 4;
int project_guess(fs_mesh *m, const double *ac, const double *anb, const double *bc, const double *const *X, int k,
                  double *x0);

// AMG (fs_amg.hip).  amg_prepare builds or refreshes the hierarchy for the
// fine operator (ac, anb): aggregation on the first call and every
// mg_reagg calls, device Galerkin operators otherwise.
void amg_prepare(fs_mesh *m, const double *ac, const double *anb);
// out = one V-cycle applied to rhs (both FP64 device vectors, not aliased),
// coarse corrections scaled by alpha.
void amg_apply(fs_mesh *m, const double *rhs, double *out, const int *flags, double alpha);
// The reference's multigrid-as-solver loop; returns cycles run.  warm: dphi
// holds an initial guess, so the first cycle smooths bc - A dphi.
int amg_solve(fs_mesh *m, const double *ac, const double *anb, const double *bc, double *dphi,
              int solver, bool warm = false);
int amg_levels(fs_mesh *m);
// Average ms of one k_spmv_dot launch (y = A x with (dot_with, y) partials).
int time_spmv(fs_mesh *m, const double *ac, const double *anb, const double *x, const double *dot_with,
              double *y, int reps, double *avg_ms);
// Host copies of the aggregation maps of level l -> l+1 for tests.
int amg_export(fs_mesh *m, int level, int *dims, int *parent, int *seed, int *nchild,
               int *child_ptr, int *child, int *nbr, int *nnbr, double *ac, double *anb);

} // namespace fs
