/*
 * oracle/or_solve.c -- CPU restatement of the reference's linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library; it is imported only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  No product code links it.
 *
 * Array conventions follow the reference's module globals exactly:
 *   ac(N), bc(N), dphi(N)                     -- cell arrays
 *   anb(N,W), ElementNeighbor(N,W)            -- ELL, column-major (cell fastest):
 *                                                element (i,j) at [j*n + i]
 *   ElementNeighbor values are 1-based cell numbers, 0 = boundary slot.
 * Loops run over cells in the reference's order so that every floating-point
 * sum is formed in the same order as the Fortran.
 *
 * Every function cites the reference file:line it restates.
 */
#include "or_solve.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1.0e-15 /* Constants1.F90:5 tiny */

/* CalculateResiduals (CalculateResidualsRoutine.for:3-70).  res[0..3] =
 * ResorAbs, ResorMax, ResorRMS, ResorScaled of bc - A*dphi. */
void or_calculate_residuals(int n, int w, const int *nbr, const int *nnbr,
                            const double *ac, const double *anb,
                            const double *bc, const double *dphi,
                            const double *fit, int nf, double res[4]) {
  double rabs = 0.0, rmax = -1.0, rscaled = (nf == 4) ? 0.0 : -1.0;
  double rms = 0.0, acphic = -1.0;
  (void)w;
  for (int i = 0; i < n; ++i) acphic = fmax(acphic, fabs(ac[i] * fit[i]));
  for (int i = 0; i < n; ++i) {
    double r = bc[i] - ac[i] * dphi[i];
    for (int j = 0; j < nnbr[i]; ++j) {
      int k = nbr[(size_t)j * n + i];
      if (k != 0) r = r - anb[(size_t)j * n + i] * dphi[k - 1];
    }
    r = fabs(r);
    rabs = rabs + r;
    rmax = fmax(rmax, r);
    rms = rms + r * r;
    if (nf != 4) rscaled = fmax(rscaled, r / (acphic + TINY));
  }
  res[0] = rabs;
  res[1] = rmax;
  res[2] = sqrt(rms / n);
  res[3] = rscaled;
}

/* CheckConvergence (CheckConvergenceRoutine.for:3-41): L1 norm of bc - A*dphi. */
double or_check_convergence(int n, int w, const int *nbr, const int *nnbr,
                            const double *ac, const double *anb,
                            const double *bc, const double *dphi) {
  double s = 0.0;
  (void)w;
  for (int i = 0; i < n; ++i) {
    double r = bc[i] - ac[i] * dphi[i];
    for (int j = 0; j < nnbr[i]; ++j) {
      int k = nbr[(size_t)j * n + i];
      if (k != 0) r = r - anb[(size_t)j * n + i] * dphi[k - 1];
    }
    s = s + fabs(r);
  }
  return s;
}

/* One symmetric Gauss-Seidel sweep, forward then backward
 * (SORMethod.for:20-60; SORMultiGrid.for:21-69 for a coarse level). */
void or_sor_sweep(int n, const int *nbr, const int *nnbr, const double *ac,
                  const double *anb, const double *bc, double *dphi) {
  for (int i = 0; i < n; ++i) {
    double s = bc[i];
    for (int j = 0; j < nnbr[i]; ++j) {
      int k = nbr[(size_t)j * n + i];
      if (k != 0) s = s - anb[(size_t)j * n + i] * dphi[k - 1];
    }
    dphi[i] = s / ac[i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = bc[i];
    for (int j = 0; j < nnbr[i]; ++j) {
      int k = nbr[(size_t)j * n + i];
      if (k != 0) s = s - anb[(size_t)j * n + i] * dphi[k - 1];
    }
    dphi[i] = s / ac[i];
  }
}

/* D-ILU diagonal (FactorizeUsingILU.for:20-78).  dc = the inverse ILU
 * diagonal.  Reproduces two behaviours of the reference:
 *  - the update loop is skipped for row N-1 (1-based) (line 36);
 *  - the search for A(j,i) may read one slot past row j's neighbour count
 *    (line 53: `i2.le.j3` lets i2 reach j3+1); the element read there is the
 *    next one in column-major storage, and only a match on i is used. */
void or_ilu_factor(int n, int w, const int *nbr, const int *nnbr,
                   const double *ac, const double *anb, double *dc) {
  size_t total = (size_t)n * w;
  for (int i = 0; i < n; ++i) dc[i] = ac[i];
  for (int i1 = 1; i1 <= n; ++i1) {
    dc[i1 - 1] = 1.0 / dc[i1 - 1];
    int i3 = nnbr[i1 - 1];
    if (i1 != n - 1) {
      for (int j2 = 1; j2 <= i3; ++j2) {
        int j4 = nbr[(size_t)(j2 - 1) * n + (i1 - 1)];
        if (j4 > i1 && j4 <= n) {
          int j3 = nnbr[j4 - 1];
          int i2 = 0, i4 = -1;
          while (i2 <= j3 && i4 != i1) {
            i2 = i2 + 1;
            size_t lin = (size_t)(i2 - 1) * n + (j4 - 1);
            i4 = lin < total ? nbr[lin] : 0;
          }
          if (i4 == i1) {
            size_t lin = (size_t)(i2 - 1) * n + (j4 - 1);
            dc[j4 - 1] = dc[j4 - 1] -
                         anb[lin] * dc[i1 - 1] * anb[(size_t)(j2 - 1) * n + (i1 - 1)];
          }
        }
      }
    }
  }
}

/* One D-ILU iteration (ILUMethod.for:22-139): residual, forward and backward
 * substitution with the factored diagonal dc, dphi += correction.  rc is
 * workspace of length n. */
void or_ilu_sweep(int n, const int *nbr, const int *nnbr, const double *ac,
                  const double *anb, const double *bc, const double *dc,
                  double *rc, double *dphi) {
  for (int i = 0; i < n; ++i) {
    double r = -ac[i] * dphi[i] + bc[i];
    for (int j = 0; j < nnbr[i]; ++j) {
      int k = nbr[(size_t)j * n + i];
      if (k != 0) r = r - anb[(size_t)j * n + i] * dphi[k - 1];
    }
    rc[i] = r;
  }
  for (int i1 = 1; i1 <= n; ++i1) {
    double p1 = dc[i1 - 1] * rc[i1 - 1];
    int i3 = nnbr[i1 - 1];
    for (int j2 = 1; j2 <= i3; ++j2) {
      int j1 = nbr[(size_t)(j2 - 1) * n + (i1 - 1)];
      if (j1 != 0 && j1 > i1 && j1 <= n) {
        int j3 = nnbr[j1 - 1];
        int i2 = 0, k = 0;
        while (i2 + 1 <= j3 && k != i1) {
          i2 = i2 + 1;
          k = nbr[(size_t)(i2 - 1) * n + (j1 - 1)];
        }
        if (k != 0 && k == i1) {
          double p2 = anb[(size_t)(i2 - 1) * n + (j1 - 1)] * p1;
          rc[j1 - 1] = rc[j1 - 1] - p2;
        }
      }
    }
  }
  for (int i1 = n; i1 >= 1; --i1) {
    if (i1 < n) {
      int i3 = nnbr[i1 - 1];
      for (int j2 = 1; j2 <= i3; ++j2) {
        int j = nbr[(size_t)(j2 - 1) * n + (i1 - 1)];
        if (j > i1)
          rc[i1 - 1] = rc[i1 - 1] - anb[(size_t)(j2 - 1) * n + (i1 - 1)] * rc[j - 1];
      }
    }
    double p1 = dc[i1 - 1] * rc[i1 - 1];
    rc[i1 - 1] = p1;
    dphi[i1 - 1] = dphi[i1 - 1] + p1;
  }
}

/* Numerical-Recipes row-indexed sparse storage (PBCGMethod.for:377-438,
 * sprsin; size from GetSizeOfPBCGarrays :51-92).  Returns the array length;
 * sa/ija are allocated here (1-based contents, index 0 unused). */
int or_sprsin(int n, const int *nbr, const int *nnbr, const double *ac,
              const double *anb, double **sa_out, int **ija_out) {
  int k = n + 1;
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= nnbr[i - 1]; ++j) {
      int k1 = nbr[(size_t)(j - 1) * n + (i - 1)];
      if (k1 != 0 && fabs(anb[(size_t)(j - 1) * n + (i - 1)]) >= 0.0 && i != k1) ++k;
    }
  int size = k;
  double *sa = (double *)calloc((size_t)size + 2, sizeof(double));
  int *ija = (int *)calloc((size_t)size + 2, sizeof(int));
  for (int j = 1; j <= n; ++j) sa[j] = ac[j - 1];
  ija[1] = n + 2;
  k = n + 1;
  for (int i = 1; i <= n; ++i) {
    for (int j = 1; j <= nnbr[i - 1]; ++j) {
      int k1 = nbr[(size_t)(j - 1) * n + (i - 1)];
      if (k1 != 0 && fabs(anb[(size_t)(j - 1) * n + (i - 1)]) >= 0.0 && i != k1) {
        ++k;
        sa[k] = anb[(size_t)(j - 1) * n + (i - 1)];
        ija[k] = k1;
      }
    }
    ija[i + 1] = k + 1;
  }
  *sa_out = sa;
  *ija_out = ija;
  return size;
}

/* dsprsax / dsprstx (PBCGMethod.for:303-371): x = A*b and x = A^T*b on the
 * row-indexed storage (1-based sa/ija, 0-based vectors). */
static void sprsax(int n, const double *sa, const int *ija, const double *b,
                   double *x) {
  for (int i = 1; i <= n; ++i) {
    double s = sa[i] * b[i - 1];
    for (int k = ija[i]; k <= ija[i + 1] - 1; ++k) s = s + sa[k] * b[ija[k] - 1];
    x[i - 1] = s;
  }
}
static void sprstx(int n, const double *sa, const int *ija, const double *b,
                   double *x) {
  for (int i = 1; i <= n; ++i) x[i - 1] = sa[i] * b[i - 1];
  for (int i = 1; i <= n; ++i)
    for (int k = ija[i]; k <= ija[i + 1] - 1; ++k) {
      int j = ija[k];
      x[j - 1] = x[j - 1] + sa[k] * b[i - 1];
    }
}

/* One call of SolveEquationsUsingPBCG (PBCGMethod.for:96-216): one iteration
 * of the minimum-residual biconjugate gradient with the D-ILU diagonal dc as
 * preconditioner (asolve, :271-297).  bkden persists across calls: in the
 * reference it is a local read before assignment when iter>0 (:108,183,194),
 * which keeps its value only where locals are static; the restatement keeps
 * it in the state, matching a build with static locals. */
void or_pbcg_iteration(or_pbcg_state *s, int iter, int n, const double *bc,
                       const double *dc, double *dphi) {
  double *p = s->p, *pp = s->pp, *r = s->r, *rr = s->rr, *z = s->z, *zz = s->zz;
  if (iter == 0) {
    sprsax(n, s->sa, s->ija, dphi, r);
    for (int j = 0; j < n; ++j) {
      r[j] = bc[j] - r[j];
      rr[j] = r[j];
    }
    sprsax(n, s->sa, s->ija, r, rr);
  }
  for (int i = 0; i < n; ++i) z[i] = r[i] * dc[i];
  for (int i = 0; i < n; ++i) zz[i] = rr[i] * dc[i];
  double bknum = 0.0;
  for (int j = 0; j < n; ++j) bknum = bknum + z[j] * rr[j];
  if (iter == 0) {
    for (int j = 0; j < n; ++j) {
      p[j] = z[j];
      pp[j] = zz[j];
    }
  } else {
    double bk = bknum / (s->bkden + 1.0e-10);
    for (int j = 0; j < n; ++j) {
      p[j] = bk * p[j] + z[j];
      pp[j] = bk * pp[j] + zz[j];
    }
  }
  s->bkden = bknum;
  sprsax(n, s->sa, s->ija, p, z);
  double akden = 0.0;
  for (int j = 0; j < n; ++j) akden = akden + z[j] * pp[j];
  double ak = bknum / (akden + 1.0e-10);
  sprstx(n, s->sa, s->ija, pp, zz);
  for (int j = 0; j < n; ++j) {
    dphi[j] = dphi[j] + ak * p[j];
    r[j] = r[j] - ak * z[j];
    rr[j] = rr[j] - ak * zz[j];
  }
}

void or_pbcg_init(or_pbcg_state *s, int n, const int *nbr, const int *nnbr,
                  const double *ac, const double *anb) {
  memset(s, 0, sizeof(*s));
  s->n = n;
  s->size = or_sprsin(n, nbr, nnbr, ac, anb, &s->sa, &s->ija);
  s->p = (double *)calloc(n, sizeof(double));
  s->pp = (double *)calloc(n, sizeof(double));
  s->r = (double *)calloc(n, sizeof(double));
  s->rr = (double *)calloc(n, sizeof(double));
  s->z = (double *)calloc(n, sizeof(double));
  s->zz = (double *)calloc(n, sizeof(double));
}

void or_pbcg_free(or_pbcg_state *s) {
  free(s->sa);
  free(s->ija);
  free(s->p);
  free(s->pp);
  free(s->r);
  free(s->rr);
  free(s->z);
  free(s->zz);
  memset(s, 0, sizeof(*s));
}

/* Non-multigrid branch of SolveEquation (SolveAlegbraicEquations.for:46-48,
 * 106-180) for solver = SOR/ILU/PBCG: dphi=0, initial residuals, iterate until
 * the L1 residual falls to rrf times its initial value or itmax, then
 * fit += dphi.  The p' source-resistance hook (:116-117) applies to artery
 * cases only and is not part of this path.  hist (may be NULL) receives the
 * L1 residual after each iteration.  Returns the iteration count. */
int or_solve_equation(int n, int w, const int *nbr, const int *nnbr,
                      const double *ac, const double *anb, const double *bc,
                      double *fit, double rrf, int itmax, int solver, int nf,
                      double *dphi, double res[4], double *hist) {
  memset(dphi, 0, sizeof(double) * n);
  or_calculate_residuals(n, w, nbr, nnbr, ac, anb, bc, dphi, fit, nf, res);
  double ires = res[0], fres = ires;
  int iter = 0;
  double *dc = (double *)calloc(n, sizeof(double));
  double *rc = (double *)calloc(n, sizeof(double));
  or_pbcg_state st;
  memset(&st, 0, sizeof(st));
  if (solver == OR_SOLVER_ILU) or_ilu_factor(n, w, nbr, nnbr, ac, anb, dc);
  if (solver == OR_SOLVER_PBCG) {
    or_pbcg_init(&st, n, nbr, nnbr, ac, anb);
    or_ilu_factor(n, w, nbr, nnbr, ac, anb, dc);
  }
  while (fres > rrf * ires && iter < itmax) {
    if (solver == OR_SOLVER_SOR) or_sor_sweep(n, nbr, nnbr, ac, anb, bc, dphi);
    if (solver == OR_SOLVER_ILU) or_ilu_sweep(n, nbr, nnbr, ac, anb, bc, dc, rc, dphi);
    if (solver == OR_SOLVER_PBCG) or_pbcg_iteration(&st, iter, n, bc, dc, dphi);
    fres = or_check_convergence(n, w, nbr, nnbr, ac, anb, bc, dphi);
    if (hist) hist[iter] = fres;
    ++iter;
  }
  for (int i = 0; i < n; ++i) fit[i] = fit[i] + dphi[i];
  if (solver == OR_SOLVER_PBCG) or_pbcg_free(&st);
  free(dc);
  free(rc);
  return iter;
}
