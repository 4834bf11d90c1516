/*
 * vvm_net.h -- C ABI of the learned initial guess for the cal_w solve.
 *
 * No counterpart in the reference, whose cal_w always starts Eigen's
 * BiCGSTAB from x = 0 (src/PoissonSolver.cpp:441-467).  A small
 * convolutional network (6 layers of 3x3 convolutions, 32 channels, GELU;
 * bf16 weights and activations, fp32 accumulation on the matrix cores)
 * predicts the correction of the extrapolated guess 2 x_{t-1} - x_{t-2}
 * from three features of the unknown grid: the residual of the extrapolated
 * guess, the last solution increment x_{t-1} - x_{t-2}, and the
 * extrapolated guess itself, each scaled by its RMS.  The output is scaled
 * by RMS(x_{t-1} - x_{t-2}).  The solver still iterates to its stop rule
 * from that guess, so the network changes the iteration count, never the
 * stopping criterion.
 *
 * Parameters are one fp32 array of vvm_net_param_count() values: for each
 * layer l = 1..6, W_l[co][ci][t] (32 x 32 x 9, tap t = 3 (dz+1) + (dx+1))
 * followed by b_l[32].  Layer 1 reads input channels 0..2 only and layer 6
 * writes output channel 0 only; the other weights are held at zero.
 *
 * Return codes as in vvm_dev.h.
 */
#ifndef VVM_NET_H
#define VVM_NET_H

#include <stdint.h>

#include "vvm_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvm_net vvm_net_t;

int vvm_net_create(int device, vvm_net_t **out);
void vvm_net_destroy(vvm_net_t *n);
const char *vvm_net_last_error(const vvm_net_t *n);
int vvm_net_param_count(void);
int vvm_net_set_params(vvm_net_t *n, const float *host);
int vvm_net_get_params(vvm_net_t *n, float *host);
/* gradient of the last vvm_net_grad_host / training step */
int vvm_net_get_grad(vvm_net_t *n, float *host);
/* He-normal weights from the seed (last layer scaled by 0.1), zero biases */
int vvm_net_init_params(vvm_net_t *n, uint64_t seed);

/* Forward pass on host data: input [nb][3][h][w], out [nb][h][w] (fp32).
 * Rows outside [0, h) read as zero; columns wrap when periodic != 0, else
 * read as zero. */
int vvm_net_forward_host(vvm_net_t *n, int nb, int h, int w, int periodic, const float *input, float *out);
/* loss = sum(mask (out - target)^2) / sum(mask) and its gradient with
 * respect to every parameter (grad may be NULL). */
int vvm_net_grad_host(vvm_net_t *n, int nb, int h, int w, int periodic, const float *input, const float *target,
                      const float *mask, float *grad, double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient. */
int vvm_net_adam(vvm_net_t *n, float lr);
/* Tile height (2, 4 or 8 pixel rows; default 2) of the convolution kernels
 * of every network in the process: a launch-shape knob like
 * vvm_dev_set_shape, with identical results for every choice. */
int vvm_net_set_conv_rows(int rows);
/* 1: inference (vvm_net_forward_host, the solver's learned guess) runs
 * hidden layers two at a time in one kernel, the activations between them
 * kept in LDS; 0 (default): one kernel a layer, as training does.
 * Identical results either way (the same bf16 roundings in the same order). */
int vvm_net_set_layer_pairs(int on);

/* ---- training data from the model's own runs ---------------------------
 * A timestep handle records one sample per solve (from its third solve on):
 * the three scaled features before the solve and the scaled target
 * (x - xe) / RMS(x_{t-1} - x_{t-2}) after it.  Call after vvm_dev_upload;
 * capacity 0 stops recording and frees the buffer. */
int vvm_dev_collect(vvm_dev_t *h, int capacity);
int vvm_dev_collected(vvm_dev_t *h);
/* Copy a handle's recorded samples into the network's training pool. */
int vvm_net_add_samples(vvm_net_t *n, vvm_dev_t *h);
int vvm_net_pool_size(vvm_net_t *n);

typedef struct vvm_net_train_opts {
    int32_t steps;        /* Adam steps                                      */
    int32_t batch;        /* crops per step                                  */
    int32_t crop_h, crop_w;
    int32_t margin;       /* crop border excluded from the loss (>= 6)       */
    float lr, lr_final;   /* cosine decay from lr to lr_final                */
    uint64_t seed;        /* crop sampling                                   */
    int32_t log_every;    /* loss read back every log_every steps (0: never) */
} vvm_net_train_opts;

typedef struct vvm_net_train_stats {
    double loss_first;    /* mean loss of the first logged window            */
    double loss_last;     /* mean loss of the last logged window             */
    double seconds;       /* wall time of the training loop                  */
    int32_t steps;
} vvm_net_train_stats;

int vvm_net_train(vvm_net_t *n, const vvm_net_train_opts *o, vvm_net_train_stats *s);

/* Held-out evaluation on a handle's recorded samples (whole periodic
 * domain): err = sum ||pred - target||^2 / sum ||target||^2, the squared
 * error left relative to the extrapolated guess's. */
int vvm_net_eval(vvm_net_t *n, vvm_dev_t *h, double *err);

/* Parameters to / from a file: "VVMNET01", int32 count, count fp32. */
int vvm_net_save(vvm_net_t *n, const char *path);
int vvm_net_load(vvm_net_t *n, const char *path);

/* Attach a network to a timestep handle on the same device (NULL
 * detaches).  With VVM_GUESS_LEARNED the solve then starts from the
 * network's guess; the network must outlive the attachment and must not be
 * trained while a step runs. */
int vvm_dev_set_guess_net(vvm_dev_t *h, vvm_net_t *n);

#ifdef __cplusplus
}
#endif
#endif
