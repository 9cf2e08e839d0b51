/*
 * oracle/or_solve.h -- CPU restatement of the reference's linear-solve path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked by the product.
 */
#ifndef OR_SOLVE_H
#define OR_SOLVE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_SOLVER_SOR = 0, OR_SOLVER_ILU = 1, OR_SOLVER_PBCG = 2 };

void or_calculate_residuals(int n, int w, const int *nbr, const int *nnbr,
                            const double *ac, const double *anb,
                            const double *bc, const double *dphi,
                            const double *fit, int nf, double res[4]);
double or_check_convergence(int n, int w, const int *nbr, const int *nnbr,
                            const double *ac, const double *anb,
                            const double *bc, const double *dphi);
void or_sor_sweep(int n, const int *nbr, const int *nnbr, const double *ac,
                  const double *anb, const double *bc, double *dphi);
void or_ilu_factor(int n, int w, const int *nbr, const int *nnbr,
                   const double *ac, const double *anb, double *dc);
void or_ilu_sweep(int n, const int *nbr, const int *nnbr, const double *ac,
                  const double *anb, const double *bc, const double *dc,
                  double *rc, double *dphi);

typedef struct {
  int n, size;
  double *sa;
  int *ija;
  double *p, *pp, *r, *rr, *z, *zz;
  double bkden;
} or_pbcg_state;

int or_sprsin(int n, const int *nbr, const int *nnbr, const double *ac,
              const double *anb, double **sa_out, int **ija_out);
void or_pbcg_init(or_pbcg_state *s, int n, const int *nbr, const int *nnbr,
                  const double *ac, const double *anb);
void or_pbcg_iteration(or_pbcg_state *s, int iter, int n, const double *bc,
                       const double *dc, double *dphi);
void or_pbcg_free(or_pbcg_state *s);
int or_solve_equation(int n, int w, const int *nbr, const int *nnbr,
                      const double *ac, const double *anb, const double *bc,
                      double *fit, double rrf, int itmax, int solver, int nf,
                      double *dphi, double res[4], double *hist);

/* one AMG level; maps to the next coarser level live on the finer level
 * (parent) and on the coarser level (seed, children) */
typedef struct {
  int n, w;
  int *nbr, *nnbr;
  double *ac, *anb, *bc, *dphi, *res;
  int *parent;              /* [n] 1-based parent at the next level */
  int *seed;                /* [n] seed cell (1-based, finer level), 0 = none */
  int *nchild;              /* [n] */
  int *child_ptr, *child;   /* CSR of children, insertion order */
} or_amg_level;

typedef struct {
  int nlevels, isolated;
  or_amg_level *levels;
} or_amg;

int or_amg_setup(or_amg *h, int n, int w, const int *nbr, const int *nnbr,
                 const double *ac, const double *anb, int maxlevels,
                 int minparents);
void or_amg_galerkin(const or_amg_level *f, or_amg_level *c);
void or_amg_update(or_amg *h, const double *ac, const double *anb);
int or_amg_solve(or_amg *h, const double *bc, double ires, double rrf,
                 int maxcycles, int npre, int npost, double *dphi,
                 double *hist);
int or_amg_solve_cycle(or_amg *h, int kind, const double *bc, double ires, double rrf,
                       int maxcycles, int npre, int npost, double *dphi, double *hist);
void or_amg_free(or_amg *h);
or_amg *or_amg_new(int n, int w, const int *nbr, const int *nnbr,
                   const double *ac, const double *anb, int maxlevels,
                   int minparents);
void or_amg_delete(or_amg *h);
int or_amg_nlevels(const or_amg *h);
int or_amg_isolated(const or_amg *h);
void or_amg_level_dims(const or_amg *h, int l, int dims[3]);
void or_amg_level_copy(const or_amg *h, int l, int *nbr, int *nnbr, double *ac,
                       double *anb, int *parent, int *seed, int *nchild,
                       int *child_ptr, int *child);

#ifdef __cplusplus
}
#endif
#endif
