/*
 * vvm_ens.h -- C ABI of the batched ensemble: many independent model
 * members of uneven grid sizes advanced together on one GPU.
 *
 * No counterpart in the reference (one model per process,
 * src/main.cpp:45-135).  Each member is a full timestep handle
 * (vvm_dev.h) with its own HIP stream; host worker threads drive the
 * members (largest first, balanced by grid points) so that the kernels of
 * different members overlap on the device.  Members that share a grid shape
 * share the launch shapes found by vvm_dev_tune.
 *
 * Return codes as in vvm_dev.h.
 */
#ifndef VVM_ENS_H
#define VVM_ENS_H

#include <stdint.h>

#include "vvm_dev.h"
#include "vvm_net.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvm_ens vvm_ens_t;

typedef struct vvm_ens_member {
    vvm_dev_config cfg;       /* grid, constants and solver options        */
    vvm_case_params params;   /* initial state (vvm_init_case)             */
} vvm_ens_member;

typedef struct vvm_ens_stats {
    double seconds;             /* wall time of the call                   */
    double grid_point_steps;    /* sum over members of interior points x steps */
    int64_t solve_iters_total;  /* PCG iterations over all members         */
    int32_t not_converged;      /* solves that missed their stop rule      */
    int32_t members;
} vvm_ens_stats;

/* Create n members on `device`, fill each initial state on the host and
 * upload it.  threads = host worker threads (0: min(n, 8)). */
int vvm_ens_create(const vvm_ens_member *members, int n, int device, int threads, vvm_ens_t **out);
void vvm_ens_destroy(vvm_ens_t *e);
const char *vvm_ens_last_error(const vvm_ens_t *e);
int vvm_ens_size(const vvm_ens_t *e);
/* the timestep handle of one member (owned by the ensemble) */
vvm_dev_t *vvm_ens_handle(vvm_ens_t *e, int i);
/* attach one network to every member (VVM_GUESS_LEARNED members use it) */
int vvm_ens_set_guess_net(vvm_ens_t *e, vvm_net_t *net);
/* tune every distinct grid shape once (vvm_dev_tune on its first member) */
int vvm_ens_tune(vvm_ens_t *e, int steps);
/* advance every member nsteps steps */
int vvm_ens_step(vvm_ens_t *e, int nsteps, vvm_ens_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
