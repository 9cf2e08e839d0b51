/*
 * fsamd_learn.h -- C ABI of the learned warm start for the p' solve.
 *
 * No counterpart in the reference.  A message-passing network over the cell
 * graph of a mesh (edges = the ELL slots of ElementNeighbor, weights
 * -anb/ac of the p' operator) predicts the pressure correction from the
 * source bc, the operator and the cell volumes.  Its prediction is the
 * initial guess of the p' Krylov solve at the reference's call site
 * (SolveEquations/ContinuityEquation/SolveContinuityEquation.for:76); the
 * solve still runs to the reference's stop rule (L1 residual <= rrF times
 * the L1 norm of bc), so the network can only change the iteration count.
 * The solver first scales the guess by the factor that minimises the L2
 * residual (so it is never worse than a zero start in that norm) and drops
 * it if its L1 residual still exceeds |bc|_1.
 *
 * Definition (3 layers, hidden 64, BF16 weights and activations, FP32
 * accumulation on the matrix cores), per system of N cells:
 *   s_b = rms(b), s_a = mean(ac), features [b/s_b, ac/s_a, nnb/W, vol/mean(vol)]
 *   and, for a network built with use_prev, xp s_a/s_b with xp the previous
 *   SIMPLE iteration's p' solution (mean-free; zero when there is none)
 *   H0 = relu(X We + be); per layer M = sum_j e_ij H_j, H += relu([H M] W_l + b_l)
 *   y = H wd + bd;  x_hat = (y - mean(y)) s_b/s_a
 * trained on (bc, converged p') pairs to the mean-free MSE of x s_a/s_b.
 * Parameters are one flat vector in the order We(8x64) be(64) {W_l(128x64)
 * b_l(64)} wd(64) bd(1), matrices row-major (input index slow).
 *
 * Arrays use the reference's layouts: elem_nbr / anb are N x W column-major,
 * neighbours 1-based with 0 = none.  Sample arrays hold n_samples systems
 * back to back (b, x, ac, vol: n_samples*N; anb: n_samples*N*W).
 * Return codes: 0 ok, < 0 error (message in fs_last_error()).
 */
#ifndef FSAMD_LEARN_H
#define FSAMD_LEARN_H

#include "fsamd_simple.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs_mpnn fs_mpnn;

/* Network on the topology of a mesh of n_cells cells (every system it sees
 * has this topology); layers in [1, 8]; use_prev 1 adds the previous p'
 * solution to the inputs; seed draws the He initialisation; capacity =
 * number of training samples the device dataset can hold. */
int fs_mpnn_create(int n_cells, int width, const int *elem_nbr, int layers, int use_prev,
                   unsigned long long seed, int capacity, fs_mpnn **out);
int fs_mpnn_destroy(fs_mpnn *net);
int fs_mpnn_param_count(fs_mpnn *net);
int fs_mpnn_get_params(fs_mpnn *net, double *theta, int count);
/* Sets the master parameters and restarts the Adam moments. */
int fs_mpnn_set_params(fs_mpnn *net, const double *theta, int count);

/* Appends samples (host arrays; x_prev may be NULL: zeros) to the device
 * dataset; returns the new sample count or < 0 when the capacity would be
 * exceeded. */
int fs_mpnn_add_samples(fs_mpnn *net, int n_samples, const double *b, const double *x,
                        const double *ac, const double *anb, const double *vol, const double *x_prev);
int fs_mpnn_sample_count(fs_mpnn *net);
/* Copies sample k of the dataset (FP32 on the device) back as FP64; any
 * pointer may be NULL. */
int fs_mpnn_get_sample(fs_mpnn *net, int k, double *b, double *x, double *ac, double *anb, double *vol,
                       double *x_prev);

/* Loss and gradient of the batch `ids` (no update).  grad may be NULL. */
int fs_mpnn_loss_grad(fs_mpnn *net, const int *ids, int batch, double *loss, double *grad);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the batch `ids`. */
int fs_mpnn_train_step(fs_mpnn *net, const int *ids, int batch, double lr, double *loss);

/* Prediction for one system on the network's topology (host arrays;
 * x_prev may be NULL). */
int fs_mpnn_predict(fs_mpnn *net, const double *b, const double *ac, const double *anb,
                    const double *vol, const double *x_prev, double *x);

/* The case's p' solves start from the network's prediction (NULL: from
 * zero, the reference).  The case's mesh must have the network's topology. */
int fs_case_set_warm_start(fs_case *c, fs_mpnn *net);
/* Each p' solve of the case appends (bc, converged p', ac, anb, vol) to the
 * network's dataset until it is full (NULL: no capture). */
int fs_case_set_snapshot_sink(fs_case *c, fs_mpnn *net);

#ifdef __cplusplus
}
#endif
#endif
