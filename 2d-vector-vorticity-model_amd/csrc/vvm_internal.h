// vvm_internal.h -- device state of one handle and the kernel launchers
// shared between the translation units of libvvm_amd.so.
#pragma once

#include "dev_common.h"
#include "vvm_dev.h"

namespace vvmd {

// Field slots of the prognostic triplets (rotated instead of copied).
enum { F_ZETA = 0, F_TH, F_QV, F_QC, F_QR, NPROG };
// AB2 histories, two planes each ([slot] = reference's [i][k][slot]).
enum { H_TH_ADV = 0, H_TH_BUOY, H_ZETA_ADV, H_QV_ADV, H_QC_ADV, H_QR_ADV, H_QR_VT, NHIST };

struct Profiles {
    double *rhou, *rhow, *pib, *pb, *thb, *thbm, *thb_init, *thb_zeta, *thvb, *thvbm;
    double *qvb, *qvb0, *qvsb, *z, *z_zeta, *lambda2, *Q1LS, *Q2LS;
    // per-level constants evaluated once on the host with the C library
    double *vtfac;    // pow(rhow[k]/rhow[1], -0.5)            (Advection.cpp:165)
    double *pcoef;    // 380 / (pow(pib[k], Cp/Rd) * P0)       (MicroPhysics.cpp:10,85)
    double *rate_th;  // dt * CGR of Nudge_theta, 0 outside     (NumericalProcess.cpp:63-68)
    double *rate_z;   // dt * CGR of Nudge_zeta, 0 outside      (NumericalProcess.cpp:77-82)
};

// Device scalars (one double each, in one small block).
enum {
    S_UBARTOP = 0,
    S_UBARTOPP,
    S_UBARTOPM,
    S_DUBAR0,
    S_DUBAR1,
    S_NEG_RATIO,       // 3 slots
    S_PRECIP_FAC = 9,  // pow(rhou[1]/rhow[1], -0.5)
    S_DIFF_ERR = 10,   // set to 1 if a diffusion wave gave up waiting
    NSCALAR = 16
};

// Ratios of NegativeValueProcess (3 fields) live at s[S_NEG_RATIO + f]:
// a value < 0 means "no rescale" (the reference's early return).

enum { DC_AM = 0, DC_DM, DC_AH, DC_DH, DC_C0, DC_N = DC_C0 + NPROG };  // DC_C0 + field

// Pointers used by one timestep (the rotation resolved).
struct StepPtrs {
    double *now[NPROG], *past[NPROG], *fut[NPROG];
    double *hist[NHIST][2];
    double *u, *w, *RKM, *RKH, *forcing;
    const double *rad;  // radiation heating rate [K/day] (nullptr: none)
    int radcool;        // RADIATIONCOOLING: -2 K/day on theta
    double *evap, *accr, *autoc, *cond;
    double *precip, *xi, *uxi;
    double *s;  // device scalars
    Profiles P;
    // diffusion-sweep coefficients (new = a yW + c + d yB), formed in
    // parallel before the sequential sweep: a, d of Mparam (zeta) and of
    // Hparam (shared by th, qv, qc, qr), and c of each field.
    double *dcoef[DC_N];
    // carry slots of the multi-workgroup diffusion sweep (uncached device
    // memory, [field][strip][k]); nullptr: the one-workgroup sweep runs
    unsigned long long *dslots;
    int dstrip_cw;  // compute wavefronts per strip of that sweep (0: default)
};

struct Reductions {
    RedSlot negval;  // 6 sums: P,N for qv, qc, qr
    RedSlot solver;  // solver dot products
    RedSlot misc;
};

// ------------------------------------------------------------ launchers
// Dynamics (k_dynamics.hip)
void launch_pzeta(const Grid &G, const StepPtrs &S, int step, hipStream_t st);
void launch_thermo4(const Grid &G, const StepPtrs &S, int step, const RedSlot &rs, hipStream_t st);
void launch_scale_forcing(const Grid &G, const StepPtrs &S, bool forcing, bool tropical, bool random_kick,
                          hipStream_t st, bool halos = true);
void launch_bc_uw(const Grid &G, const StepPtrs &S, hipStream_t st);
void launch_bc_all_prog(const Grid &G, const StepPtrs &S, hipStream_t st);
// Whether a handle runs the microphysics phase on its second stream beside
// the elliptic solve (default on).  The ensemble turns it off: with one
// stream pair per member on a few hardware queues, the cross-stream waits
// block other members' work queued behind them.
void dev_set_phase_overlap(vvm_dev_t *h, bool on);
void launch_pubartop(const Grid &G, const StepPtrs &S, int step, hipStream_t st);
// The reference operators one at a time (vvm_dev_op_field).
void launch_advection_zeta(const Grid &G, const StepPtrs &S, int step, hipStream_t st);
void launch_buoyancy(const Grid &G, const StepPtrs &S, int step, hipStream_t st);
void launch_advection_scalar(const Grid &G, const StepPtrs &S, int step, int field, hipStream_t st);
void launch_advection_qrvt(const Grid &G, const StepPtrs &S, int step, hipStream_t st);
void launch_precip(const Grid &G, const StepPtrs &S, hipStream_t st);
void launch_bc_one(const Grid &G, double *a, bool westdown, hipStream_t st);
// process 0 autoconversion, 1 accretion, 2 evaporation, 3 condensation
void launch_kessler_process(const Grid &G, const StepPtrs &S, int process, hipStream_t st);
void launch_negative_value(const Grid &G, const StepPtrs &S, double *a, const RedSlot &rs, hipStream_t st);

// Physics (k_physics.hip)
void launch_kessler(const Grid &G, const StepPtrs &S, const RedSlot &rs, hipStream_t st);
void launch_turbulence_coeffs(const Grid &G, const StepPtrs &S, hipStream_t st);
bool diffusion_supported(const Grid &G);  // nx - 2 <= 8192
// nudge bit 0 fuses Nudge_theta, bit 1 Nudge_zeta into the sweep's store;
// only_field >= 0 sweeps that one field (Mparam or one Hparam) alone.
void launch_diffusion(const Grid &G, const StepPtrs &S, int nudge, hipStream_t st, int only_field = -1);
// Carry slots for the multi-workgroup sweep on `device` (allocated once per
// process and device, never freed); nullptr when unavailable.
unsigned long long *diffusion_slots(const Grid &G, int device);
void diffusion_slots_release(unsigned long long *p);
bool diffusion_strip_waves_ok(int cw);
void launch_nudge(const Grid &G, const StepPtrs &S, int nudge, hipStream_t st);
void launch_nudge_qv(const Grid &G, const StepPtrs &S, double tau, hipStream_t st);
void launch_update_mean(const Grid &G, const StepPtrs &S, bool shift_ubar, hipStream_t st);

// Elliptic (k_poisson.hip): see poisson.h
}  // namespace vvmd
