/*
 * qdeq_guess.h -- learned initial guess for the RK43 stage solves
 * (libqdeq.so).  No reference counterpart: the reference starts every
 * elliptic solve from the previous solve's displacement
 * (odeSolver.cpp:652-720, u_dev keeps the last solution).  Here a small
 * network predicts the change of the displacement caused by the change of
 * the fault boundary data since the previous solve; multigrid then iterates
 * from the prediction to the same tolerance, so the network never changes
 * what a solve converges to, only where it starts.
 *
 * Network (dimensions for nk = 1025, factor 4):
 *   input  x = ds / max|ds|: ds[j] = fault-column displacement (slip/2) now
 *          minus at the previous solve, j = 0..nk-1 (1025 values);
 *   encoder Conv1d(1 -> 8 channels, width 8, stride 4) + ReLU -> 256 x 8;
 *   MLP 2048 -> 512 -> 512 -> 512 (ReLU) -> nc*nc (nc = (nk-1)/4+1 = 257),
 *          no biases (the map is positively homogeneous, zero in -> zero out);
 *   output c = the displacement change / max|ds| on the coarse grid of
 *          every 4th node (257 x 257), bilinearly interpolated to the solve
 *          grid (the fault column uses the exact boundary change).
 * The uniform change dR of the remote-edge data is added exactly as dR * W_R,
 * W_R = the solution for unit remote data and zero fault data (one solve when
 * a context is attached).  The guess is U_prev + max|ds| * P(c) + dR * W_R.
 * It is used only when its residual norm is below the previous solution's
 * (and finite); otherwise the solve starts from the previous solution
 * exactly as without the network.
 *
 * Arithmetic: BF16 weights and activations, FP32 accumulation (gfx950
 * 16x16x32 BF16 MFMA), FP32 master weights and Adam moments, FP32 encoder.
 * Training pairs (x, coarse displacement change - dR*W_R, both / max|ds|)
 * are harvested by attached contexts from their own solves and kept in HBM.
 */
#ifndef QDEQ_GUESS_H
#define QDEQ_GUESS_H

#include "qdeq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qdeq_guess qdeq_guess;

typedef struct qdeq_guess_config {
  int nk;              /* fault nodes = level-0 rows; (nk - 1) % 32 == 0 */
  int factor;          /* output grid stride on the solve grid (4) */
  int channels;        /* encoder channels (8) */
  int width;           /* encoder kernel width (8); its stride is factor */
  int hidden;          /* MLP width (512, multiple of 128) */
  int capacity;        /* training pairs kept on the device */
  int stride;          /* HARVEST keeps one pair every `stride` solves (1) */
  unsigned long long seed;  /* weight initialisation */
  int device;          /* HIP device ordinal */
} qdeq_guess_config;

typedef struct qdeq_train_config {
  int epochs;
  int batch;           /* pairs per step (rows padded to 128 on the device) */
  double lr, lr_final; /* Adam step size, decayed geometrically to lr_final */
  double beta1, beta2, eps;
  unsigned long long seed;  /* minibatch shuffling */
} qdeq_train_config;

enum { QDEQ_GUESS_OFF = 0, QDEQ_GUESS_HARVEST = 1, QDEQ_GUESS_PREDICT = 2 };

void qdeq_guess_default_config(qdeq_guess_config *cfg, int nk);
void qdeq_train_default_config(qdeq_train_config *cfg);
int qdeq_guess_create(const qdeq_guess_config *cfg, qdeq_guess **out);
int qdeq_guess_destroy(qdeq_guess *g);
const char *qdeq_guess_last_error(const qdeq_guess *g);

/* Parameters as one FP32 vector: conv [channels][width], W1 [hidden][F0],
 * W2, W3 [hidden][hidden], W4 [nc*nc][hidden] (F0 = (nk-1)/factor*channels;
 * row-major [out][in]). */
long long qdeq_guess_num_params(const qdeq_guess *g);
int qdeq_guess_dims(const qdeq_guess *g, int *nout, int *f0);
int qdeq_guess_get_params(qdeq_guess *g, float *out);
int qdeq_guess_set_params(qdeq_guess *g, const float *in);

/* Training pairs harvested so far (at most capacity), and reset. */
int qdeq_guess_pairs(const qdeq_guess *g, int *n);
int qdeq_guess_clear_pairs(qdeq_guess *g);
/* Host access to the stored pairs (x: n x nk, t: n x nc*nc, scale: n;
 * pairs first .. first+n-1); any pointer may be NULL. */
int qdeq_guess_read_pairs(qdeq_guess *g, int first, int n, float *x, float *t,
                          double *scale);

/* Adam on the harvested pairs with a nonzero scale; *final_loss = mean
 * squared error of the last epoch (may be NULL). */
int qdeq_guess_train(qdeq_guess *g, const qdeq_train_config *cfg,
                     double *final_loss);

/* One forward + backward pass on host data (no update): x [batch x nk],
 * t [batch x nc*nc]; *loss = mean squared error, grad = dloss/dparams in the
 * qdeq_guess_get_params layout. */
int qdeq_guess_forward_backward(qdeq_guess *g, int batch, const float *x,
                                const float *t, double *loss, float *grad);
/* Forward pass on host data: y [batch x nc*nc]. */
int qdeq_guess_forward(qdeq_guess *g, int batch, const float *x, float *y);

/* Attaches g to a context (same device, nk == ctx nk) for the RK43 driver's
 * solves: HARVEST stores a pair per solve, PREDICT starts each solve from
 * the network's guess; OFF detaches.  The first attach of a context computes
 * W_R (one solve to a relative residual of 1e-12). */
int qdeq_attach_guess(qdeq_ctx *ctx, qdeq_guess *g, int mode);
/* Solves with a prediction, predictions accepted, pairs harvested. */
int qdeq_guess_stats(qdeq_ctx *ctx, long long *predicted, long long *accepted,
                     long long *harvested);
/* Test hook: multiplies every prediction by factor (NaN makes every guess
 * non-finite, which the solve must reject). */
int qdeq_guess_inject(qdeq_ctx *ctx, double factor);
/* The context's inference path on one host input x [nk] -> y [nc*nc]. */
int qdeq_guess_infer(qdeq_ctx *ctx, const float *x, float *y);

#ifdef __cplusplus
}
#endif
#endif /* QDEQ_GUESS_H */
