/*
 * oracle/or_amg.c -- CPU restatement of the reference's additive-correction
 * algebraic multigrid (SolveEquations/AlgebraicMultiGrid/).
 *
 * TEST INFRASTRUCTURE ONLY (see or_solve.c).
 *
 * The reference keeps every level in flat arrays offset by ijbeginE/ijbeginN
 * (MultiGrid1.F90:6-26).  This restatement keeps one struct per level; the
 * contents (parents, seed, children in insertion order, coarse neighbour
 * lists in insertion order, coefficients) are the same, and every sum is
 * formed in the reference's order.  Cell numbers stored in the maps are
 * 1-based and local to their level, as in the reference once ijbeginE is
 * subtracted.
 */
#include "or_solve.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1.0e-15 /* Constants1.F90:5 */

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  return p;
}

/* Children-per-aggregate caps by coarse level (AgglomorateALevelAlgebraicly.for:79-105). */
static int max_children_for_level(int ilevel) {
  if (ilevel == 2) return 8;
  if (ilevel == 3) return 4;
  if (ilevel == 4) return 3;
  return 2;
}

/* One level of greedy aggregation (AgglomorateALevelAlgebraicly.for:48-331)
 * plus the coarse connectivity (:733-787).  f is level ilevel-1, c receives
 * level ilevel's size and neighbour lists.  Returns the number of isolated
 * orphans (aggregates with no seed, :326-330). */
static int aggregate_level(or_amg_level *f, or_amg_level *c, int ilevel) {
  const int nf = f->n;
  const double ratiomin = 0.5; /* :45 */
  const int maxch = max_children_for_level(ilevel);
  double *maxanb = (double *)xcalloc(nf, sizeof(double));
  int *parents = (int *)xcalloc(nf, sizeof(int));
  /* children stored per parent number; a parent never has more than
   * maxch + (orphans) children, bounded by nf */
  int *nchild = (int *)xcalloc(nf + 1, sizeof(int));
  int *seed = (int *)xcalloc(nf + 1, sizeof(int));
  int **clist = (int **)xcalloc(nf + 1, sizeof(int *));
  int *ccap = (int *)xcalloc(nf + 1, sizeof(int));
  int isolated = 0;

#define NBR(L, i, j) ((L)->nbr[(size_t)(j) * (L)->n + (i)])
#define ANB(L, i, j) ((L)->anb[(size_t)(j) * (L)->n + (i)])
#define PUSH_CHILD(P, V)                                                   \
  do {                                                                     \
    if (nchild[P] >= ccap[P]) {                                            \
      ccap[P] = ccap[P] ? 2 * ccap[P] : 8;                                 \
      clist[P] = (int *)realloc(clist[P], sizeof(int) * (size_t)ccap[P]);  \
    }                                                                      \
    clist[P][nchild[P]++] = (V);                                           \
  } while (0)

  for (int i = 0; i < nf; ++i) { /* :52-61 */
    maxanb[i] = -1.0;
    for (int j = 0; j < f->nnbr[i]; ++j) maxanb[i] = fmax(maxanb[i], -ANB(f, i, j));
  }
  int iparent = 1;
  for (int s1 = 1; s1 <= nf; ++s1) { /* step 1, :112-151 */
    int nb_children = 0;
    if (parents[s1 - 1] != 0) continue;
    parents[s1 - 1] = iparent;
    nchild[iparent] = 0;
    for (int l = 0; l < f->nnbr[s1 - 1]; ++l) {
      int inb = NBR(f, s1 - 1, l);
      if (inb == 0 || parents[inb - 1] != 0) continue;
      double ratio = -ANB(f, s1 - 1, l) / fmax(maxanb[s1 - 1], TINY);
      if (ratio > ratiomin && nb_children < maxch) {
        parents[inb - 1] = iparent;
        PUSH_CHILD(iparent, inb);
        nb_children++;
      }
    }
    int nb2 = nb_children; /* step 2, :155-200 */
    for (int cl = 0; cl < nb2; ++cl) {
      int ich = clist[iparent][cl];
      if (ich == 0) continue;
      for (int l = 0; l < f->nnbr[ich - 1]; ++l) {
        int inb = NBR(f, ich - 1, l);
        if (inb == 0 || parents[inb - 1] != 0) continue;
        double ratio = -ANB(f, ich - 1, l) / fmax(maxanb[ich - 1], TINY);
        if (ratio > ratiomin && nb_children < maxch) {
          parents[inb - 1] = iparent;
          nb_children++;
          PUSH_CHILD(iparent, inb);
        }
      }
    }
    if (nb_children == 0) { /* :202-213 */
      parents[s1 - 1] = 0;
      nchild[iparent] = 0;
    } else {
      seed[iparent] = s1;
      nchild[iparent] = nb_children;
      iparent++;
    }
  }
  for (int io = 1; io <= nf; ++io) { /* orphans, :291-331 */
    if (parents[io - 1] == 0) {
      double strength = 0.0;
      int index = 0;
      for (int l = 0; l < f->nnbr[io - 1]; ++l) {
        int inb = NBR(f, io - 1, l);
        if (inb == 0) continue;
        if (parents[inb - 1] != 0 && index != 1) {
          double ratio = -ANB(f, io - 1, l) / maxanb[inb - 1];
          if (strength < ratio) {
            index = 1;
            strength = ratio;
            int p = parents[inb - 1];
            parents[io - 1] = p;
            PUSH_CHILD(p, io);
          }
        }
      }
    }
    if (parents[io - 1] == 0) {
      parents[io - 1] = iparent;
      nchild[iparent] = 0;
      seed[iparent] = 0;
      iparent++;
      isolated++;
    }
  }
  const int nc = iparent - 1;

  /* coarse connectivity in fine visit order (:733-787) */
  int **cn = (int **)xcalloc(nc + 1, sizeof(int *));
  int *cnn = (int *)xcalloc(nc + 1, sizeof(int));
  int *ccn = (int *)xcalloc(nc + 1, sizeof(int));
  int wmax = 0;
  for (int e = 1; e <= nf; ++e) {
    for (int l = 0; l < f->nnbr[e - 1]; ++l) {
      int inb = NBR(f, e - 1, l);
      if (inb == 0) continue;
      int p = parents[e - 1], q = parents[inb - 1];
      if (p == q) continue;
      int found = 0;
      for (int k = 0; k < cnn[p]; ++k)
        if (cn[p][k] == q) found = 1;
      if (!found) {
        if (cnn[p] >= ccn[p]) {
          ccn[p] = ccn[p] ? 2 * ccn[p] : 8;
          cn[p] = (int *)realloc(cn[p], sizeof(int) * (size_t)ccn[p]);
        }
        cn[p][cnn[p]++] = q;
        if (cnn[p] > wmax) wmax = cnn[p];
      }
    }
  }
  if (wmax == 0) wmax = 1;

  /* export */
  f->parent = parents;
  c->n = nc;
  c->w = wmax;
  c->nbr = (int *)xcalloc((size_t)nc * wmax, sizeof(int));
  c->nnbr = (int *)xcalloc(nc, sizeof(int));
  c->seed = (int *)xcalloc(nc, sizeof(int));
  c->nchild = (int *)xcalloc(nc, sizeof(int));
  c->child_ptr = (int *)xcalloc(nc + 1, sizeof(int));
  for (int p = 1; p <= nc; ++p) c->child_ptr[p] = c->child_ptr[p - 1] + nchild[p];
  c->child = (int *)xcalloc(c->child_ptr[nc] + 1, sizeof(int));
  for (int p = 1; p <= nc; ++p) {
    c->seed[p - 1] = seed[p];
    c->nchild[p - 1] = nchild[p];
    for (int k = 0; k < nchild[p]; ++k) c->child[c->child_ptr[p - 1] + k] = clist[p][k];
    c->nnbr[p - 1] = cnn[p];
    for (int k = 0; k < cnn[p]; ++k) c->nbr[(size_t)k * nc + (p - 1)] = cn[p][k];
  }
  c->ac = (double *)xcalloc(nc, sizeof(double));
  c->anb = (double *)xcalloc((size_t)nc * wmax, sizeof(double));
  c->bc = (double *)xcalloc(nc, sizeof(double));
  c->dphi = (double *)xcalloc(nc, sizeof(double));
  c->res = (double *)xcalloc(nc, sizeof(double));
  for (int p = 0; p <= nc; ++p) free(cn[p]);
  for (int p = 0; p <= nf; ++p) free(clist[p]);
  free(cn);
  free(cnn);
  free(ccn);
  free(clist);
  free(ccap);
  free(nchild);
  free(seed);
  free(maxanb);
  return isolated;
#undef PUSH_CHILD
}

/* Galerkin coarse operator A_c = P^T A P for piecewise-constant P
 * (AssembleElementBasedAgglomoratedLHS.for:29-150): the seed's row, then each
 * child's row in list order; intra-aggregate couplings go to the diagonal,
 * the others to the coarse slot holding the neighbour's parent.  An
 * aggregate with no seed (isolated orphan) would index acMG(0) in the
 * reference; here its seed term is skipped. */
void or_amg_galerkin(const or_amg_level *f, or_amg_level *c) {
  const int nc = c->n;
  for (int p = 0; p < nc; ++p) {
    c->ac[p] = 0.0;
    for (int j = 0; j < c->nnbr[p]; ++j) c->anb[(size_t)j * nc + p] = 0.0;
  }
  for (int p = 1; p <= nc; ++p) {
    double acc = 0.0;
    int nmem = c->nchild[p - 1] + 1;
    for (int m = 0; m < nmem; ++m) {
      int e = (m == 0) ? c->seed[p - 1] : c->child[c->child_ptr[p - 1] + m - 1];
      if (e <= 0) continue;
      acc = acc + f->ac[e - 1];
      for (int l = 0; l < f->nnbr[e - 1]; ++l) {
        int inb = NBR(f, e - 1, l);
        if (inb <= 0) continue;
        int q = f->parent[inb - 1];
        if (q <= 0) continue;
        double a = ANB(f, e - 1, l);
        if (q == p) {
          acc = acc + a;
        } else {
          for (int k = 0; k < c->nnbr[p - 1]; ++k)
            if (c->nbr[(size_t)k * nc + (p - 1)] == q)
              c->anb[(size_t)k * nc + (p - 1)] = c->anb[(size_t)k * nc + (p - 1)] + a;
        }
      }
    }
    c->ac[p - 1] = acc;
  }
}

/* Level construction (AgglomorateGrid.for:28-72): add levels while
 * level < maxlevels and N_level > 2*minparents. */
int or_amg_setup(or_amg *h, int n, int w, const int *nbr, const int *nnbr,
                 const double *ac, const double *anb, int maxlevels,
                 int minparents) {
  memset(h, 0, sizeof(*h));
  h->levels = (or_amg_level *)xcalloc(maxlevels + 1, sizeof(or_amg_level));
  or_amg_level *l0 = &h->levels[0];
  l0->n = n;
  l0->w = w;
  l0->nbr = (int *)xcalloc((size_t)n * w, sizeof(int));
  l0->nnbr = (int *)xcalloc(n, sizeof(int));
  memcpy(l0->nbr, nbr, sizeof(int) * (size_t)n * w);
  memcpy(l0->nnbr, nnbr, sizeof(int) * (size_t)n);
  l0->ac = (double *)xcalloc(n, sizeof(double));
  l0->anb = (double *)xcalloc((size_t)n * w, sizeof(double));
  l0->bc = (double *)xcalloc(n, sizeof(double));
  l0->dphi = (double *)xcalloc(n, sizeof(double));
  l0->res = (double *)xcalloc(n, sizeof(double));
  memcpy(l0->ac, ac, sizeof(double) * n);
  memcpy(l0->anb, anb, sizeof(double) * (size_t)n * w);
  int nl = 1;
  h->isolated = 0;
  while (nl < maxlevels && h->levels[nl - 1].n > 2 * minparents) {
    h->isolated += aggregate_level(&h->levels[nl - 1], &h->levels[nl], nl + 1);
    or_amg_galerkin(&h->levels[nl - 1], &h->levels[nl]);
    nl++;
  }
  h->nlevels = nl;
  return nl;
}

/* Reload the fine coefficients and rebuild every coarse operator with the
 * existing maps (AlgebraicMultiGridMainRoutine.for:303-326). */
void or_amg_update(or_amg *h, const double *ac, const double *anb) {
  or_amg_level *l0 = &h->levels[0];
  memcpy(l0->ac, ac, sizeof(double) * l0->n);
  memcpy(l0->anb, anb, sizeof(double) * (size_t)l0->n * l0->w);
  for (int l = 1; l < h->nlevels; ++l) or_amg_galerkin(&h->levels[l - 1], &h->levels[l]);
}

void or_amg_free(or_amg *h) {
  for (int l = 0; l < h->nlevels; ++l) {
    or_amg_level *L = &h->levels[l];
    free(L->nbr);
    free(L->nnbr);
    free(L->ac);
    free(L->anb);
    free(L->bc);
    free(L->dphi);
    free(L->res);
    free(L->parent);
    free(L->seed);
    free(L->nchild);
    free(L->child_ptr);
    free(L->child);
  }
  free(h->levels);
  memset(h, 0, sizeof(*h));
}

static double level_l1(const or_amg_level *L) { /* CheckMultiGridConvergence.for:21-41 */
  return or_check_convergence(L->n, L->w, L->nbr, L->nnbr, L->ac, L->anb, L->bc, L->dphi);
}

/* CalculateResidualsMG (CalculateMultiGridResiduals.for:14-33) */
static void level_residual(or_amg_level *L) {
  const int n = L->n;
  for (int i = 0; i < n; ++i) {
    double r = L->bc[i] - L->ac[i] * L->dphi[i];
    for (int j = 0; j < L->nnbr[i]; ++j) {
      int k = L->nbr[(size_t)j * n + i];
      if (k > 0) r = r - L->anb[(size_t)j * n + i] * L->dphi[k - 1];
    }
    L->res[i] = r;
  }
}

/* SolveEquationMG with the 'sor' smoother (SolveEquationsWithMultiGrid.for:
 * 18-35): symmetric GS sweeps until the level's L1 residual is at most
 * rrf times the fine initial residual, or itmax sweeps. */
static void level_smooth(or_amg_level *L, double ires, double rrf, int itmax) {
  double fres = ires;
  int iter = 0;
  while (fres > rrf * ires && iter < itmax) {
    or_sor_sweep(L->n, L->nbr, L->nnbr, L->ac, L->anb, L->bc, L->dphi);
    fres = level_l1(L);
    iter++;
  }
}

/* AssembleAgglomeratedRHS (AssembleNodeAndElementBasedAgglomoratedRHS.for:17-33) */
static void restrict_rhs(const or_amg_level *f, or_amg_level *c) {
  for (int i = 0; i < c->n; ++i) c->bc[i] = 0.0;
  for (int i = 0; i < f->n; ++i) {
    int p = f->parent[i];
    c->bc[p - 1] = c->bc[p - 1] + f->res[i];
  }
}

/* CorrectFinerLevelSolution (CorrectSolutionOnFinerLevel.for:17-24) */
static void prolong_add(or_amg_level *f, const or_amg_level *c) {
  for (int i = 0; i < f->n; ++i) f->dphi[i] = f->dphi[i] + c->dphi[f->parent[i] - 1];
}

/* Restrict(iLevel) (RestrictSolution.for:17-20), levels numbered from 1 as
 * in the reference: smooth level iLevel, its residual, the next level's
 * right-hand side */
static void restrict_at(or_amg *h, int ilevel, double ires, double rrf, int npre) {
  or_amg_level *f = &h->levels[ilevel - 1];
  level_smooth(f, ires, rrf, npre);
  level_residual(f);
  restrict_rhs(f, &h->levels[ilevel]);
}

/* Prolongate(iLevel) (ProlongateSolution.for:15-16): level iLevel's
 * correction into level iLevel-1, then smooth level iLevel-1 */
static void prolongate_at(or_amg *h, int ilevel, double ires, double rrf, int npost) {
  prolong_add(&h->levels[ilevel - 2], &h->levels[ilevel - 1]);
  level_smooth(&h->levels[ilevel - 2], ires, rrf, npost);
}

/* SolveEquationMG + CalculateResidualsMG on the coarsest level */
static void coarsest(or_amg *h, double ires, double rrf, int npost) {
  level_smooth(&h->levels[h->nlevels - 1], ires, rrf, npost);
  level_residual(&h->levels[h->nlevels - 1]);
}

/* vMultiGridCycle, fMultiGridCycle, wMultiGridCycle (MultiGridCycle.for:32-47,
 * 84-127, 162-220): cycle 0 = V, 1 = F, 2 = W */
static void cycle(or_amg *h, int kind, double ires, double rrf, int npre, int npost) {
  const int lmax = h->nlevels;
  for (int il = 1; il <= lmax - 1; ++il) restrict_at(h, il, ires, rrf, npre);
  coarsest(h, ires, rrf, npost);
  if (kind >= 1) {
    for (int l1 = lmax - 1; l1 >= 2; --l1) {
      for (int il = lmax; il >= l1 + 1; --il) prolongate_at(h, il, ires, rrf, npost);
      for (int il = l1; il <= lmax - 1; ++il) restrict_at(h, il, ires, rrf, npre);
      coarsest(h, ires, rrf, npost);
    }
  }
  if (kind == 2) {
    for (int l1 = 3; l1 <= lmax - 1; ++l1) {
      for (int il = lmax; il >= l1 + 1; --il) prolongate_at(h, il, ires, rrf, npost);
      for (int il = l1; il <= lmax - 1; ++il) restrict_at(h, il, ires, rrf, npre);
      coarsest(h, ires, rrf, npost);
    }
  }
  for (int il = lmax; il >= 2; --il) prolongate_at(h, il, ires, rrf, npost);
}

/* Cycle loop of AlgebraicMultigrid (AlgebraicMultiGridMainRoutine.for:371-416)
 * with V, F or W cycles (kind 0, 1, 2): every level's dphi and residual start at zero once, the
 * fine right-hand side is bc, and cycles repeat while the fine L1 residual
 * exceeds rrf*ires.  ires is ResorAbs(NF) of the caller.  Returns cycles run. */
int or_amg_solve_cycle(or_amg *h, int kind, const double *bc, double ires, double rrf,
                       int maxcycles, int npre, int npost, double *dphi, double *hist) {
  for (int l = 0; l < h->nlevels; ++l) {
    or_amg_level *L = &h->levels[l];
    memset(L->dphi, 0, sizeof(double) * L->n);
    memset(L->res, 0, sizeof(double) * L->n);
  }
  memcpy(h->levels[0].bc, bc, sizeof(double) * h->levels[0].n);
  double fres = ires;
  int ncycle = 1;
  while (ncycle <= maxcycles && fres > rrf * ires) {
    cycle(h, kind, ires, rrf, npre, npost);
    fres = level_l1(&h->levels[0]);
    if (hist) hist[ncycle - 1] = fres;
    ncycle++;
  }
  memcpy(dphi, h->levels[0].dphi, sizeof(double) * h->levels[0].n);
  return ncycle - 1;
}

int or_amg_solve(or_amg *h, const double *bc, double ires, double rrf,
                 int maxcycles, int npre, int npost, double *dphi,
                 double *hist) {
  return or_amg_solve_cycle(h, 0, bc, ires, rrf, maxcycles, npre, npost, dphi, hist);
}

/* ---- flat accessors for the ctypes wrapper ---- */
or_amg *or_amg_new(int n, int w, const int *nbr, const int *nnbr,
                   const double *ac, const double *anb, int maxlevels,
                   int minparents) {
  or_amg *h = (or_amg *)xcalloc(1, sizeof(or_amg));
  or_amg_setup(h, n, w, nbr, nnbr, ac, anb, maxlevels, minparents);
  return h;
}
void or_amg_delete(or_amg *h) {
  or_amg_free(h);
  free(h);
}
int or_amg_nlevels(const or_amg *h) { return h->nlevels; }
int or_amg_isolated(const or_amg *h) { return h->isolated; }
void or_amg_level_dims(const or_amg *h, int l, int dims[3]) {
  const or_amg_level *L = &h->levels[l];
  dims[0] = L->n;
  dims[1] = L->w;
  dims[2] = L->child_ptr ? L->child_ptr[L->n] : 0;
}
void or_amg_level_copy(const or_amg *h, int l, int *nbr, int *nnbr, double *ac,
                       double *anb, int *parent, int *seed, int *nchild,
                       int *child_ptr, int *child) {
  const or_amg_level *L = &h->levels[l];
  memcpy(nbr, L->nbr, sizeof(int) * (size_t)L->n * L->w);
  memcpy(nnbr, L->nnbr, sizeof(int) * L->n);
  memcpy(ac, L->ac, sizeof(double) * L->n);
  memcpy(anb, L->anb, sizeof(double) * (size_t)L->n * L->w);
  if (L->parent) memcpy(parent, L->parent, sizeof(int) * L->n);
  if (L->seed) {
    memcpy(seed, L->seed, sizeof(int) * L->n);
    memcpy(nchild, L->nchild, sizeof(int) * L->n);
    memcpy(child_ptr, L->child_ptr, sizeof(int) * (L->n + 1));
    memcpy(child, L->child, sizeof(int) * L->child_ptr[L->n]);
  }
}
