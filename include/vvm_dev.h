/*
 * vvm_dev.h -- C ABI of the MI355X dynamical-core timestep of the 2D vector
 * vorticity model (VVM).
 *
 * The reference swaps its elliptic back-end at compile time behind the
 * static members of class vvm (src/Declare.hpp:707-941; the GPU_POISSON
 * variant src/PoissonSolver_GPU_mpi.cpp keeps the same cal_w / cal_u /
 * pubarTop_pt / InitPoissonMatrix signatures).  This header is the thin C
 * layer below a Config_VVM / vvm compatible C++ API: plain pointers and
 * sizes, no C++ types, no exceptions, no exit().
 *
 * Layout conventions
 *   - nx, nz are the array extents INCLUDING one ghost cell on each side,
 *     exactly as the reference's vvm::nx / vvm::nz (src/Declare.hpp:82-83).
 *   - Host 2-D fields are the reference's contiguous blocks `<name>cont`
 *     (src/Declare.hpp:400-407): element (i,k) at [i*nz + k], k fastest.
 *   - Host AB2 histories are the reference's `***` blocks
 *     (src/Declare.hpp:417-427): element (i,k,slot) at [(i*nz + k)*2 + slot].
 *   - The device keeps every 2-D field x-contiguous ([k][i], padded pitch);
 *     upload / download transpose once, never per step.
 *
 * Return codes: 0 success, > 0 warning (solver did not reach its stop rule;
 * the step continued, as the reference does at src/PoissonSolver.cpp:463-467),
 * < 0 error (see vvm_dev_last_error).
 */
#ifndef VVM_DEV_H
#define VVM_DEV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VVM_OK 0
#define VVM_WARN_NOT_CONVERGED 1
#define VVM_ERR_ARG (-1)
#define VVM_ERR_HIP (-2)
#define VVM_ERR_SIZE (-3)
#define VVM_ERR_STATE (-4)
#define VVM_ERR_NOMEM (-5)

/* Poisson back-ends (north of the reference's single Eigen path). */
#define VVM_SOLVER_FP64 0   /* FP64 multigrid-preconditioned CG            */
#define VVM_SOLVER_MIXED 1  /* FP32 MG-PCG inside FP64 iterative refinement */
/* The reference GPU back-end's solver (src/PoissonSolver_GPU_mpi.cpp:41:
 * AMGX BiCGSTAB + ILU0, tolerance 1e-12, 10000 iterations): BiCGSTAB with a
 * red-black ILU0 preconditioner, for comparison runs; with VVM_STOP_ABS and
 * tol_rel = 1e-12 it stops as AMGX's default ABSOLUTE criterion does.
 * Needs an even interior column count (nx - 2). */
#define VVM_SOLVER_BICGSTAB 2

/* Initial guess for the cal_w unknown x = rhow[k]*w[i][k]. */
#define VVM_GUESS_ZERO 0     /* reference behaviour (Eigen x.setZero())     */
#define VVM_GUESS_PREV 1     /* x from the previous step                    */
#define VVM_GUESS_EXTRAP 2   /* 2 x_{t-1} - x_{t-2}                         */
#define VVM_GUESS_LEARNED 3  /* conv-net correction of the quadratic
                                extrapolation (VVM_GUESS_EXTRAP2)             */
#define VVM_GUESS_EXTRAP2 4  /* 3 (x_{t-1} - x_{t-2}) + x_{t-3}; the linear
                                form until three solutions exist             */

/* Stop rules of the elliptic solve. */
#define VVM_STOP_PARITY 0  /* iterate until the residual stops decreasing   */
#define VVM_STOP_REL 1     /* ||b - A x||_2 / ||b||_2 <= tol_rel            */
#define VVM_STOP_ABS 2     /* ||b - A x||_2 <= tol_rel (BiCGSTAB only)       */

/* Model configuration: the runtime members of the reference's Config_VVM
 * (src/Declare.hpp:27-70) that the timestep reads, plus back-end options. */
typedef struct vvm_dev_config {
    int32_t nx, nz;          /* extents including ghosts                     */
    double dx, dz, dt;
    double GRAVITY, Cp, Cv, Rd, Lv, P0, PSURF;
    double addforcingtime;   /* s; random theta kick while step*dt <= this   */
    double moisture_nudge_time; /* s; 0 disables Nudge_qv                    */
    double tolerance;        /* reference relative tolerance (1e-22)         */
    int32_t CASE;            /* CASE==2 skips Nudge_zeta                     */
    int32_t tropical_forcing;/* 1: thp += dt*Q1LS[k], qvp += dt*Q2LS[k]      */
    int32_t solver_mode;     /* VVM_SOLVER_*                                 */
    int32_t guess_mode;      /* VVM_GUESS_*                                  */
    int32_t stop_mode;       /* VVM_STOP_*                                   */
    int32_t max_iter;        /* cap on PCG iterations per solve              */
    double tol_rel;          /* used by VVM_STOP_REL                         */
} vvm_dev_config;

/* Host view of the whole model state: the `...cont` blocks and 1-D profiles
 * of the reference's class vvm (src/Declare.hpp:490-699).  Any pointer may
 * be NULL on download (that array is skipped); on upload every pointer
 * listed as "state" must be set. */
typedef struct vvm_host_view {
    /* 2-D prognostic triplets (state) */
    double *zeta, *zetam, *zetap;
    double *th, *thm, *thp;
    double *qv, *qvm, *qvp;
    double *qc, *qcm, *qcp;
    double *qr, *qrm, *qrp;
    /* 2-D diagnostics (state: u, w, RKM, RKH; the rest outputs only) */
    double *u, *w, *RKM, *RKH, *init_th_forcing;
    double *evaporation, *accretion, *autoconversion, *condensation;
    /* AB2 histories [i][k][2] (state) */
    double *dth_advect, *dth_buoyancy, *dzeta_advect;
    double *dqv_advect, *dqc_advect, *dqr_advect, *dqr_VT;
    /* nx-long rows */
    double *precip, *xi, *uxi;
    /* nz-long profiles (state) */
    double *rhou, *rhow, *pib, *pb, *thb, *thbm, *thb_init, *thb_zeta;
    double *thvb, *thvbm, *qvb, *qvb0, *qvsb, *z, *z_zeta, *lambda2;
    double *Q1LS, *Q2LS;
    /* scalars */
    double ubarTop, ubarTopp, ubarTopm;
    double dubarTop_advect[2];
    int32_t step;
    int32_t status_for_adding_forcing;
} vvm_host_view;

/* Per-call statistics; per-phase times mirror the reference's t_* arrays
 * (src/Declare.hpp:695-699, filled at src/Iteration.cpp:286,297,405,418,435). */
typedef struct vvm_step_stats {
    int32_t steps;             /* steps taken by this call                   */
    int32_t solve_iters_total; /* PCG iterations summed over the steps      */
    int32_t solve_iters_last;  /* PCG iterations of the last step           */
    int32_t refine_last;       /* outer refinement sweeps of the last step  */
    int32_t not_converged;     /* solves that hit max_iter                   */
    double rel_residual_last;  /* ||b-Ax||/||b|| of the last solve          */
    double t_advection, t_poisson, t_microphysics, t_diffusion, t_all; /* s */
    int32_t guess_fallbacks;   /* solves whose learned guess started worse
                                  than the extrapolation (or not finite) and
                                  were restarted from the extrapolation    */
    int32_t guess_skipped;     /* solves that did not run the network: after
                                  a restart it sits out 1, 4, 16, then 64
                                  solves (while its guesses keep losing;
                                  reset by a kept guess)                     */
} vvm_step_stats;

typedef struct vvm_solve_stats {
    int32_t iters;          /* PCG iterations (inner, summed)               */
    int32_t refine;         /* outer FP64 refinement sweeps (mixed only)    */
    double rel_residual;    /* ||b - A x||_2 / ||b||_2                      */
    double rel_residual0;   /* same, for the initial guess                  */
    double seconds;         /* device time of the solve                     */
} vvm_solve_stats;

typedef struct vvm_dev vvm_dev_t;

/* Create a handle bound to one device and one stream; allocates all device
 * memory for the grid.  Replaces the back-end set-up the reference does in
 * InitPoissonMatrix (src/PoissonSolver.cpp:681-813) / InitAMGX
 * (src/PoissonSolver_GPU_mpi.cpp:8-80). */
int vvm_dev_create(const vvm_dev_config *cfg, int device, vvm_dev_t **out);
void vvm_dev_destroy(vvm_dev_t *h);
const char *vvm_dev_last_error(const vvm_dev_t *h);

/* Copy the full host state to the device ([i][k] -> [k][i]). */
int vvm_dev_upload(vvm_dev_t *h, const vvm_host_view *v);
/* Copy the device state back to the host (non-NULL pointers only). */
int vvm_dev_download(vvm_dev_t *h, vvm_host_view *v);

/* Advance nsteps steps of vvm::Iteration::TimeMarching's loop body
 * (src/Iteration.cpp:227-436) on the device; returns after the stream
 * synchronises.  stats may be NULL. */
int vvm_dev_step(vvm_dev_t *h, int nsteps, vvm_step_stats *stats);
/* Per-step phase timers, as the reference records t_advection, t_poisson,
 * t_microphysics, t_diffusion and t_all for every step
 * (src/Iteration.cpp:286,297,405,418,435): with out set, each later
 * vvm_dev_step writes the device times (s) of its steps s = 0 .. min(nsteps,
 * cap) - 1 to out[5 s + 0..4] in that order (t_all: the whole step on the
 * device).  out must hold 5 * cap doubles; NULL (or cap 0) turns it off. */
int vvm_dev_set_step_timers(vvm_dev_t *h, double *out, int cap);

/* Isolated elliptic solve of the reference's cal_w system A x = b
 * (src/PoissonSolver.cpp:441-467, matrix src/PoissonSolver.cpp:760-796):
 * b and x are host arrays of (nx-2)*(nz-3) doubles in the reference's row
 * order idx = (k-2)*(nx-2) + (i-1).  x is the initial guess on entry. */
int vvm_dev_poisson(vvm_dev_t *h, const double *b, double *x, vvm_solve_stats *stats);
/* Same with device-resident b and x (same layout), for benches. */
int vvm_dev_poisson_device(vvm_dev_t *h, const double *d_b, double *d_x, vvm_solve_stats *stats);

/* Per-operator parity hooks: run one reference operator on the device
 * state (tests).  op is one of VVM_OP_*. */
#define VVM_OP_PZETA 1        /* Iteration::pzeta_pt   src/Iteration.cpp:29-34   */
/* pth_pt, pqv_pt, pqc_pt and pqr_pt (src/Iteration.cpp:36-40, 90-124) act
 * on independent fields and run as one fused operator. */
#define VVM_OP_PSCALARS 2
#define VVM_OP_FORCING_BC 6   /* AddForcing + BoundaryProcess2D_all (:283-285)  */
#define VVM_OP_PUBARTOP 7     /* PoissonSolver::pubarTop_pt :655-676            */
#define VVM_OP_CAL_W 8        /* PoissonSolver::cal_w :439-477                  */
#define VVM_OP_CAL_U 9        /* PoissonSolver::cal_u :580-653                  */
#define VVM_OP_MICROPHYSICS 10/* Kessler + NegativeValueProcess x3 + BC :302-312,404 */
#define VVM_OP_TURBULENCE 11  /* Turbulence::RKM_RKH (+Mparam/Hparam) Turbulent.cpp:5-66, no nudging */
#define VVM_OP_NUDGE 12       /* Nudge_theta/zeta/qv NumericalProcess.cpp:60-96 */
#define VVM_OP_NEXT 13        /* Iteration::nextTimeStep :153-204               */
#define VVM_OP_BC_ALL 14      /* BoundaryProcess2D_all Boundary.cpp:34-52       */
int vvm_dev_op(vvm_dev_t *h, int op);

/* The reference's operators one at a time: each static of class vvm
 * (src/Declare.hpp:707-935) runs alone on the device state and writes what
 * the reference function writes, nothing more (the timestep fuses them).
 * `field` names the array an operator takes as an argument in the
 * reference; it is ignored by the others (pass 0). */
#define VVM_FIELD_ZETA 0
#define VVM_FIELD_TH 1
#define VVM_FIELD_QV 2
#define VVM_FIELD_QC 3
#define VVM_FIELD_QR 4
#define VVM_LEVEL_M 0        /* zetam, thm, ... */
#define VVM_LEVEL_NOW 1      /* zeta, th, ...   */
#define VVM_LEVEL_P 2        /* zetap, thp, ... */
#define VVM_FIELD_AT(fld, level) ((level) * 8 + (fld))
#define VVM_OP_ADVECTION_ZETA 20  /* vvm::Advection_zeta        Advection.cpp:84-149        */
#define VVM_OP_BOUYANCY 21        /* vvm::Bouyancy              Buoyancy.cpp:11-58          */
#define VVM_OP_ADVECTION_THERMO 22/* vvm::Advection_thermo      Advection.cpp:25-81; field TH..QR */
#define VVM_OP_ADVECTION_QRVT 23  /* vvm::Advection_qrVT        Advection.cpp:153-192       */
#define VVM_OP_PTH 24             /* Iteration::pth_pt          Iteration.cpp:36-40         */
#define VVM_OP_PQV 25             /* Iteration::pqv_pt          Iteration.cpp:90-98         */
#define VVM_OP_PQC 26             /* Iteration::pqc_pt          Iteration.cpp:100-107       */
#define VVM_OP_PQR 27             /* Iteration::pqr_pt          Iteration.cpp:109-124       */
#define VVM_OP_UPDATE_MEAN 28     /* Iteration::updateMean      Iteration.cpp:42-86         */
#define VVM_OP_AUTOCONVERSION 29  /* MicroPhysics::autoconversion MicroPhysics.cpp:29-51    */
#define VVM_OP_ACCRETION 30       /* MicroPhysics::accretion    MicroPhysics.cpp:54-77      */
#define VVM_OP_EVAPORATION 31     /* MicroPhysics::evaporation  MicroPhysics.cpp:80-114     */
#define VVM_OP_CONDENSATION 32    /* MicroPhysics::condensation MicroPhysics.cpp:7-26       */
#define VVM_OP_NEGATIVE_VALUE 33  /* MicroPhysics::NegativeValueProcess MicroPhysics.cpp:116-138;
                                     field VVM_FIELD_AT(fld, level), any prognostic array */
#define VVM_OP_MPARAM 34          /* Turbulence::Mparam on zetap Turbulent.cpp:68-90      */
#define VVM_OP_HPARAM 35          /* Turbulence::Hparam         Turbulent.cpp:92-114; field TH..QR (p level) */
#define VVM_OP_NUDGE_THETA 36     /* NumericalProcess::Nudge_theta NumericalProcess.cpp:60-71 */
#define VVM_OP_NUDGE_ZETA 37      /* NumericalProcess::Nudge_zeta  NumericalProcess.cpp:73-85 */
#define VVM_OP_NUDGE_QV 38        /* NumericalProcess::Nudge_qv    NumericalProcess.cpp:87-96 */
#define VVM_OP_ADD_FORCING 39     /* vvm::AddForcing alone      AddForcing.cpp:6-28         */
int vvm_dev_op_field(vvm_dev_t *h, int op, int field);

/* NegativeValueProcess on a caller's device field, allocating nothing.  No
 * counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It runs the kernels of
 * vvm_dev_op_field(h, VVM_OP_NEGATIVE_VALUE, f) on d_var, whose element
 * (i,k) of the reference's var[i][k] is at d_var[k*pitch + i] (ghost cells
 * included, pitch >= nx; only i = 1..nx-2, k = 1..nz-2 are read or written),
 * with their per-workgroup partial sums and the correction ratio in a
 * static device array of 128 KiB that each GPU holds until the process
 * ends.  Limit: 8192 workgroups of 64x4 points, i.e. ceil(nx/64)*ceil(nz/4)
 * <= 8192 (2050 x 992 points); larger grids return VVM_ERR_SIZE with a
 * message on stderr.  Asynchronous on `stream` (a hipStream_t; NULL is the
 * null stream): the caller synchronises.  No hipMalloc/hipFree, no
 * allocating library call.  Calls that overlap on one GPU share the array
 * and race; serialise them (one stream, or events). */
int vvm_negative_value_noalloc(double *d_var, int nx, int nz, int pitch, void *stream);

/* Measurement aid (no counterpart in the reference, whose only
 * instrumentation is the per-phase t_* timers): time every launch of one
 * kernel of vvm_dev_step with HIP events on the handle's stream.
 * vvm_dev_kernel_timer selects the kernel (VVM_KT_*, 0 = off) and resets the
 * totals; vvm_dev_kernel_time returns the mean launch duration so far. */
#define VVM_KT_NONE 0
#define VVM_KT_PZETA 1       /* advection of zeta + buoyancy                 */
#define VVM_KT_THERMO4 2     /* advection of th, qv, qc, qr + rain fall       */
#define VVM_KT_SCALE 3       /* negative-value rescale + forcing + halos      */
#define VVM_KT_KESSLER 4     /* Kessler microphysics                          */
#define VVM_KT_TURB 5        /* RKM/RKH                                       */
#define VVM_KT_DIFFUSION 6   /* in-place K-diffusion sweep (5 fields) + nudge */
#define VVM_KT_MEAN 7        /* horizontal means (updateMean)                 */
#define VVM_KT_CALU 8        /* cal_u (both kernels)                          */
#define VVM_KT_SOLVE 9       /* the whole elliptic solve (many kernels)       */
int vvm_dev_kernel_timer(vvm_dev_t *h, int kernel);
int vvm_dev_kernel_time(vvm_dev_t *h, double *avg_seconds, int *launches);

/* Launch-shape autotuner (no counterpart in the reference): times `steps`
 * whole steps for each candidate launch shape on the live state and keeps
 * the fastest; the model state is restored afterwards.  Candidates: the
 * diffusion sweep's strip width (2, 4 or 8 wavefronts of 64 columns) and
 * the LDS budget (16, 32, 64 KiB) that decides which multigrid levels run in
 * the one-workgroup coarse kernel.  Results are cached per device, grid and
 * solver mode for the life of the process. */
typedef struct vvm_tune_result {
    int32_t diffusion_strip_waves;
    int32_t mg_coarse_lds_kib;
    double ms_per_step;      /* the chosen shape                             */
    double ms_default;       /* the default shape (4 waves, 64 KiB)          */
} vvm_tune_result;
int vvm_dev_tune(vvm_dev_t *h, int steps, vvm_tune_result *out);
/* Set the same launch shapes by hand (0 keeps the current value); takes
 * effect from the next step.  The coarse budget needs an uploaded state. */
int vvm_dev_set_shape(vvm_dev_t *h, int diffusion_strip_waves, int mg_coarse_lds_kib);

/* Host coupling of the radiation scheme (src/Rrtmgp.cpp, src/AddForcing.cpp):
 * heating_rate is the reference's radiation_heating_rate, [nx][nz] in the
 * reference's layout, K/day, which every later step adds to thp as
 * AddForcing.cpp:20 does (thp += dt * rate / 86400); NULL removes it.
 * cooling 1 adds RADIATIONCOOLING's -2 K/day (:21-23).  The host refreshes
 * the rate as often as its radiation runs (every 40 steps in the reference,
 * src/Iteration.cpp:427-431), from the state it downloads. */
int vvm_dev_set_radiation(vvm_dev_t *h, const double *heating_rate, int cooling);

/* Host coupling of a column microphysics scheme in place of Kessler (the
 * reference's P3 call, src/Iteration.cpp:314-402, which runs on the host
 * over [i][k] columns): each step, after cal_u, the handle downloads the
 * step's fields into a host view in the reference's [i][k] layout (every
 * time level, u, w, RKM, RKH, the process rates, precip and the profiles;
 * no AB2 histories), calls fn(user, view), and uploads thp, qvp, qcp, qrp
 * and the four process rates back, then applies the halos (:404).  The
 * scheme replaces Kessler and its NegativeValueProcess; fn returns 0, or
 * a non-zero code that stops the step with VVM_ERR_STATE.  NULL restores
 * the device Kessler. */
typedef int (*vvm_host_microphysics_fn)(void *user, vvm_host_view *view);
int vvm_dev_set_host_microphysics(vvm_dev_t *h, vvm_host_microphysics_fn fn, void *user);

/* Host-side helpers (no GPU use). */

/* Fill the analytic base state and initial fields of one of the benchmark
 * cases (see vvm_case_params) into a caller-allocated host view. */
typedef struct vvm_case_params {
    int32_t sounding;        /* 0: theta = 300 K exp(N^2 z/g), RH ramp       */
    double brunt;            /* N [1/s]                                       */
    double bubble_dT;        /* K, cosine bubble amplitude (0: none)          */
    double bubble_radius;    /* m                                             */
    double bubble_zc;        /* m                                             */
    int32_t noise_kind;      /* 0: none, 1: uniform +-amp, 2: gaussian sigma  */
    double noise_amp;        /* K                                             */
    double noise_ztop;       /* m; noise only below this height               */
    uint64_t seed;
    double sfc_heat_flux;    /* W/m^2 -> Q1LS at k=1                          */
    double sfc_qv_ramp;      /* kg/kg/day at the surface, ramped to 0 at ramp_z */
    double sfc_qv_ramp_z;    /* m                                             */
} vvm_case_params;
int vvm_init_case(const vvm_dev_config *cfg, const vvm_case_params *p, vvm_host_view *v);

/* The reference's own initial state: vvm::Init::Init1d then Init2d
 * (src/Init.cpp:14-266).  sounding_path NULL selects Init1d's analytic
 * sounding (src/Init.cpp:17-116); otherwise the 10-column sounding file of
 * LoadFile (src/Init.cpp:302-370; header line and zero level skipped, the
 * first nz-2 levels used).  cfg->tropical_forcing selects Init2d's
 * TROPICALFORCING branch (random theta kick, Q1LS/Q2LS from the file);
 * otherwise cfg->CASE picks the initial fields (0 mean state, 1 warm
 * bubble, 2 bubble and shear) and cfg->addforcingtime > 0 adds the kick.
 * Every array of the view is written (zeroed first, as a new model).
 * Returns VVM_ERR_SIZE when the file has fewer levels than the grid. */
int vvm_init_reference(const vvm_dev_config *cfg, const char *sounding_path, vvm_host_view *v);
/* The two halves: Init1d (zeroes the view, then the base-state profiles)
 * and Init2d (the initial fields from those profiles). */
int vvm_init_profiles(const vvm_dev_config *cfg, const char *sounding_path, vvm_host_view *v);
int vvm_init_fields(const vvm_dev_config *cfg, vvm_host_view *v);

/* The reference's Init::LoadFromPreviousFile (src/Init.cpp:372-485), the
 * restart from two consecutive per-step netCDF outputs (nc_prev: the step
 * before, nc_now: the restart step; as written by output_nc) and the
 * sounding file.  Writes what the reference writes on top of the view
 * (which should be zero, as a new model); other arrays are left alone, as
 * the reference leaves them.  Reads netCDF classic files (CDF-1 / CDF-2). */
int vvm_init_from_previous(const vvm_dev_config *cfg, const char *sounding_path, const char *nc_prev,
                           const char *nc_now, vvm_host_view *v);

/* The reference's Init::RandomPerturbation (src/Init.cpp:488-515): fills
 * init_th_forcing from std::mt19937(t) and std::normal_distribution. */
int vvm_random_perturbation(const vvm_dev_config *cfg, vvm_host_view *v, int t,
                            double min_range, double max_range, double stddev);

/* Number of doubles a host view needs for each class of array. */
int64_t vvm_field_len(int nx, int nz);   /* nx*nz   */
int64_t vvm_hist_len(int nx, int nz);    /* nx*nz*2 */

#ifdef __cplusplus
}
#endif
#endif
