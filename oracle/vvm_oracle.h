/*
 * vvm_oracle.h -- CPU restatement of the reference VVM timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product never links, loads or calls it.
 *
 * It works in place on a vvm_host_view (include/vvm_dev.h): the reference's
 * [i][k] z-fastest layout, FP64, scalar loops in the reference's order.
 */
#ifndef VVM_ORACLE_H
#define VVM_ORACLE_H
#include "vvm_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

/* G-solve flavour used by vo_cal_u. */
#define VO_GSOLVE_CG 0     /* unpreconditioned CG, tol 1e-12, as the reference */
#define VO_GSOLVE_EXACT 1  /* DFT solve of the periodic 1-D problem (mean zero) */

int vo_pzeta_pt(const vvm_dev_config *c, vvm_host_view *v);
int vo_pth_pt(const vvm_dev_config *c, vvm_host_view *v);
int vo_pqv_pt(const vvm_dev_config *c, vvm_host_view *v);
int vo_pqc_pt(const vvm_dev_config *c, vvm_host_view *v);
int vo_pqr_pt(const vvm_dev_config *c, vvm_host_view *v);
int vo_forcing_bc(const vvm_dev_config *c, vvm_host_view *v);
int vo_bc_all(const vvm_dev_config *c, vvm_host_view *v);
int vo_pubartop_pt(const vvm_dev_config *c, vvm_host_view *v);
int vo_cal_w(const vvm_dev_config *c, vvm_host_view *v);
int vo_cal_u(const vvm_dev_config *c, vvm_host_view *v, int gsolve);
int vo_microphysics(const vvm_dev_config *c, vvm_host_view *v);
int vo_negval(const vvm_dev_config *c, double *var);
int vo_turbulence(const vvm_dev_config *c, vvm_host_view *v);
int vo_nudge(const vvm_dev_config *c, vvm_host_view *v);
int vo_update_mean(const vvm_dev_config *c, vvm_host_view *v);
int vo_next_time_step(const vvm_dev_config *c, vvm_host_view *v);
/* One static of class vvm alone (VVM_OP_ADVECTION_ZETA .. VVM_OP_ADD_FORCING). */
int vo_op_single(const vvm_dev_config *c, vvm_host_view *v, int op, int field);

/* Init::RandomPerturbation (src/Init.cpp:488-515): init_th_forcing from
 * std::mt19937(t) and std::normal_distribution (vvm_oracle_rng.cpp). */
int vo_random_perturbation(const vvm_dev_config *c, vvm_host_view *v, int t, double min_range,
                           double max_range, double stddev);

/* nsteps iterations of the TimeMarching loop body starting from v->step.
 * While step*dt <= addforcingtime each step draws a new random theta kick
 * into v->init_th_forcing (src/Iteration.cpp:276-282) before AddForcing. */
int vo_step(const vvm_dev_config *c, vvm_host_view *v, int nsteps, int gsolve);

/* Exact solve of cal_w's system A x = b (b, x in the reference's row order
 * idx = (k-2)(nx-2) + (i-1)).  Returns 0 on success. */
int vo_poisson_exact(const vvm_dev_config *c, const double *rhou, const double *rhow,
                     const double *b, double *x);
/* y = A x with the reference's assembled matrix. */
void vo_poisson_apply(const vvm_dev_config *c, const double *rhou, const double *rhow,
                      const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
