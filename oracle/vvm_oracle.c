/*
 * vvm_oracle.c -- scalar FP64 CPU restatement of the reference VVM timestep
 * (Kessler, AB2, radiation off) on the reference's [i][k] layout.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function cites the reference lines it restates.  Arithmetic is kept
 * in the reference's evaluation order (and built with -ffp-contract=off) so
 * that, apart from the elliptic solves, results agree with the compiled
 * reference to the last bit.  The elliptic solve here is an exact direct
 * method (DFT along the periodic x, Thomas along z), which pins the
 * reference's BiCGSTAB/ILUT result to round-off; see vo_poisson_exact.
 *
 * std::max / std::min are restated as MAXS/MINS with the same NaN
 * behaviour ((a<b)?b:a and (b<a)?b:a).
 */
#include "vvm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXS(a, b) (((a) < (b)) ? (b) : (a))
#define MINS(a, b) (((b) < (a)) ? (b) : (a))
#define PI_D 3.14159265358979323846

/* Grid accessors on the [i][k] layout. */
#define AT(a, i, k) (a)[(size_t)(i) * nz + (k)]
#define HS(a, i, k, s) (a)[((size_t)(i) * nz + (k)) * 2 + (s)]

typedef struct {
    int nx, nz;
    double dx, dz, dt, rdx, r2dx, rdz, r2dz, rdx2, rdz2;
} grid_t;

static grid_t mkgrid(const vvm_dev_config *c) {
    /* vvm constructor (src/Declare.hpp:79-84) */
    grid_t g;
    g.nx = c->nx;
    g.nz = c->nz;
    g.dx = c->dx;
    g.dz = c->dz;
    g.dt = c->dt;
    g.rdx = 1.0 / c->dx;
    g.r2dx = g.rdx / 2.0;
    g.rdz = 1.0 / c->dz;
    g.r2dz = g.rdz / 2.0;
    g.rdx2 = g.rdx * g.rdx;
    g.rdz2 = g.rdz * g.rdz;
    return g;
}

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

/* ------------------------------------------------------------------ halos */

/* BoundaryProcess1D_center (src/Boundary.cpp:3-7) */
static void bc1d_center(double *a, int n) {
    a[0] = a[1];
    a[n - 1] = a[n - 2];
}

/* BoundaryProcess2D_center (src/Boundary.cpp:9-19): periodic x columns,
 * then zero-gradient rows (corners follow the column copies). */
static void bc_center(double *a, int nx, int nz) {
    for (int k = 1; k <= nz - 2; k++) {
        AT(a, 0, k) = AT(a, nx - 2, k);
        AT(a, nx - 1, k) = AT(a, 1, k);
    }
    for (int i = 0; i <= nx - 1; i++) {
        AT(a, i, 0) = AT(a, i, 1);
        AT(a, i, nz - 1) = AT(a, i, nz - 2);
    }
}

/* BoundaryProcess2D_westdown (src/Boundary.cpp:21-31): periodic x, then
 * rows 0, 1 and nz-1 zero (w and zeta live on the lower cell edge). */
static void bc_westdown(double *a, int nx, int nz) {
    for (int k = 1; k <= nz - 2; k++) {
        AT(a, 0, k) = AT(a, nx - 2, k);
        AT(a, nx - 1, k) = AT(a, 1, k);
    }
    for (int i = 0; i <= nx - 1; i++) {
        AT(a, i, 0) = 0.;
        AT(a, i, 1) = 0.;
        AT(a, i, nz - 1) = 0.;
    }
}

/* BoundaryProcess2D_all (src/Boundary.cpp:33-52), Kessler build. */
int vo_bc_all(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    bc_westdown(v->zetap, nx, nz);
    bc_center(v->thp, nx, nz);
    bc_westdown(v->w, nx, nz);
    bc_center(v->u, nx, nz);
    bc_center(v->qvp, nx, nz);
    bc_center(v->qcp, nx, nz);
    bc_center(v->qrp, nx, nz);
    return 0;
}

/* -------------------------------------------------------------- advection */

/* Third-order upwind-biased flux correction of Advection.cpp:40-50 for one
 * face.  cp0 = rho*PLUS(vel) at the face, cpm = rho'*PLUS(vel) one face back,
 * cm0 = rho*MINU(vel) at the face, cmp = rho''*MINU(vel) one face ahead,
 * written with the reference's association: sqrt((rho*P0)*rho'*Pm). */
static inline double upwind3(double rho0, double v0, double rhom, double vm, double rhop, double vp,
                             double fm2, double fm1, double f0, double fp1) {
    double p0 = 0.5 * (v0 + fabs(v0)), pm = 0.5 * (vm + fabs(vm));
    double m0 = 0.5 * (v0 - fabs(v0)), mp = 0.5 * (vp - fabs(vp));
    return -1. / 3. *
           (rho0 * p0 * (f0 - fm1) - sqrt(rho0 * p0 * rhom * pm) * (fm1 - fm2) - rho0 * m0 * (f0 - fm1) -
            sqrt(fabs(rho0 * m0 * rhop * mp)) * (fp1 - f0));
}

/* vvm::Advection_thermo (src/Advection.cpp:25-81), AB2 branch. */
static void adv_thermo(const vvm_dev_config *c, vvm_host_view *v, const double *now, double *fut, double *d) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *u = v->u, *w = v->w, *rhou = v->rhou, *rhow = v->rhow;
    double *fu = zalloc((size_t)nx * nz), *fw = zalloc((size_t)nx * nz);
    for (int k = 1; k <= nz - 1; k++) {
        for (int i = 1; i <= nx - 1; i++) {
            AT(fu, i, k) = rhou[k] * AT(u, i, k) * (AT(now, i, k) + AT(now, i - 1, k));
            AT(fw, i, k) = rhow[k] * AT(w, i, k) * (AT(now, i, k) + AT(now, i, k - 1));
            if (i >= 2 && i <= nx - 3 && k >= 2 && k <= nz - 3) {
                AT(fu, i, k) += upwind3(rhou[k], AT(u, i, k), rhou[k], AT(u, i - 1, k), rhou[k], AT(u, i + 1, k),
                                        AT(now, i - 2, k), AT(now, i - 1, k), AT(now, i, k), AT(now, i + 1, k));
                AT(fw, i, k) += upwind3(rhow[k], AT(w, i, k), rhow[k - 1], AT(w, i, k - 1), rhow[k + 1],
                                        AT(w, i, k + 1), AT(now, i, k - 2), AT(now, i, k - 1), AT(now, i, k),
                                        AT(now, i, k + 1));
            }
        }
    }
    bc_center(fu, nx, nz);
    bc_center(fw, nx, nz);
    for (int i = 0; i < nx; i++) AT(fw, i, 0) = AT(fw, i, nz - 1) = 0.;
    const int sn = (v->step + 1) % 2, so = v->step % 2;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double px = (AT(fu, i + 1, k) - AT(fu, i, k)) * g.r2dx / rhou[k];
            double pz = (AT(fw, i, k + 1) - AT(fw, i, k)) * g.r2dz / rhou[k];
            HS(d, i, k, sn) = -px - pz;
            if (v->step == 0) HS(d, i, k, 0) = HS(d, i, k, 1);
            AT(fut, i, k) = AT(now, i, k) + 1.5 * g.dt * HS(d, i, k, sn) - 0.5 * g.dt * HS(d, i, k, so);
        }
    }
    free(fu);
    free(fw);
}

/* vvm::Advection_zeta (src/Advection.cpp:84-149), AB2 branch. */
static void adv_zeta(const vvm_dev_config *c, vvm_host_view *v) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *u = v->u, *w = v->w, *rhou = v->rhou, *rhow = v->rhow, *z = v->zeta;
    double *Uw = zalloc((size_t)nx * nz), *Wu = zalloc((size_t)nx * nz);
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            AT(Uw, i, k) = 0.25 * (rhou[k] * (AT(u, i + 1, k) + AT(u, i, k)) +
                                   rhou[k - 1] * (AT(u, i + 1, k - 1) + AT(u, i, k - 1)));
            AT(Wu, i, k) = 0.25 * (rhow[k + 1] * (AT(w, i, k + 1) + AT(w, i - 1, k + 1)) +
                                   rhow[k] * (AT(w, i, k) + AT(w, i - 1, k)));
        }
    }
    bc_center(Uw, nx, nz);
    bc_center(Wu, nx, nz);
    for (int i = 0; i < nx; i++) AT(Wu, i, 0) = AT(Wu, i, nz - 1) = 0.;

    double *fu = zalloc((size_t)nx * nz), *fw = zalloc((size_t)nx * nz);
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 0; i <= nx - 2; i++) {
            double U = AT(Uw, i, k), W = AT(Wu, i, k);
            AT(fu, i, k) = U * (AT(z, i + 1, k) + AT(z, i, k));
            AT(fw, i, k) = W * (AT(z, i, k + 1) + AT(z, i, k));
            if (i >= 2 && i <= nx - 3 && k >= 2 && k <= nz - 3) {
                /* density-weighted corner velocities: rho factors are 1 */
                double Pu = 0.5 * (U + fabs(U)), Pum = 0.5 * (AT(Uw, i - 1, k) + fabs(AT(Uw, i - 1, k)));
                double Mu = 0.5 * (U - fabs(U)), Mup = 0.5 * (AT(Uw, i + 1, k) - fabs(AT(Uw, i + 1, k)));
                AT(fu, i, k) += -1. / 3. *
                                (Pu * (AT(z, i + 1, k) - AT(z, i, k)) - sqrt(Pu * Pum) * (AT(z, i, k) - AT(z, i - 1, k)) -
                                 Mu * (AT(z, i + 1, k) - AT(z, i, k)) -
                                 sqrt(fabs(Mu * Mup)) * (AT(z, i + 2, k) - AT(z, i + 1, k)));
                double Pw = 0.5 * (W + fabs(W)), Pwm = 0.5 * (AT(Wu, i, k - 1) + fabs(AT(Wu, i, k - 1)));
                double Mw = 0.5 * (W - fabs(W)), Mwp = 0.5 * (AT(Wu, i, k + 1) - fabs(AT(Wu, i, k + 1)));
                AT(fw, i, k) += -1. / 3. *
                                (Pw * (AT(z, i, k + 1) - AT(z, i, k)) - sqrt(Pw * Pwm) * (AT(z, i, k) - AT(z, i, k - 1)) -
                                 Mw * (AT(z, i, k + 1) - AT(z, i, k)) -
                                 sqrt(fabs(Mw * Mwp)) * (AT(z, i, k + 2) - AT(z, i, k + 1)));
            }
        }
    }
    bc_center(fu, nx, nz);
    bc_center(fw, nx, nz);
    for (int i = 0; i < nx; i++) AT(fw, i, 0) = AT(fw, i, nz - 1) = 0.;

    const int sn = (v->step + 1) % 2, so = v->step % 2;
    double *d = v->dzeta_advect;
    for (int k = 2; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double px = (AT(fu, i, k) - AT(fu, i - 1, k)) * g.r2dx / rhow[k];
            double pz = (AT(fw, i, k) - AT(fw, i, k - 1)) * g.r2dz / rhow[k];
            HS(d, i, k, sn) = -px - pz;
            if (v->step == 0) HS(d, i, k, 0) = HS(d, i, k, 1);
            AT(v->zetap, i, k) = AT(z, i, k) + 1.5 * g.dt * HS(d, i, k, sn) - 0.5 * g.dt * HS(d, i, k, so);
        }
    }
    free(Uw);
    free(Wu);
    free(fu);
    free(fw);
}

/* vvm::Bouyancy (src/Buoyancy.cpp:11-58), WATER + AB2. */
static void buoyancy(const vvm_dev_config *c, vvm_host_view *v) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *th = v->th, *qv = v->qv, *qc = v->qc, *qr = v->qr, *rhow = v->rhow, *thvb = v->thvb;
    double *d = v->dth_buoyancy;
    const int sn = (v->step + 1) % 2, so = v->step % 2;
#define THV(i, k) (AT(th, i, k) * (1 + 0.608 * AT(qv, i, k)))
    for (int k = 2; k < nz - 1; k++) {
        for (int i = 1; i < nx - 1; i++) {
            double gth = c->GRAVITY / rhow[k] * 0.5 *
                         ((THV(i, k) - THV(i - 1, k)) / thvb[k] + (THV(i, k - 1) - THV(i - 1, k - 1)) / thvb[k - 1]) *
                         g.rdx;
            HS(d, i, k, sn) = gth;
            if (v->step == 0) HS(d, i, k, 0) = HS(d, i, k, 1);
            double gqc = c->GRAVITY / rhow[k] *
                         (0.5 * (AT(qc, i, k) + AT(qc, i, k - 1)) - 0.5 * (AT(qc, i - 1, k) + AT(qc, i - 1, k - 1))) *
                         g.rdx;
            double gqr = c->GRAVITY / rhow[k] *
                         (0.5 * (AT(qr, i, k) + AT(qr, i, k - 1)) - 0.5 * (AT(qr, i - 1, k) + AT(qr, i - 1, k - 1))) *
                         g.rdx;
            HS(d, i, k, sn) += -gqc - gqr;
            if (v->step == 0) HS(d, i, k, 0) = HS(d, i, k, 1);
            AT(v->zetap, i, k) += (1.5 * g.dt * HS(d, i, k, sn)) - (0.5 * g.dt * HS(d, i, k, so));
        }
    }
#undef THV
}

/* vvm::Advection_qrVT (src/Advection.cpp:153-192): rain sedimentation. */
static void adv_qrvt(const vvm_dev_config *c, vvm_host_view *v) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *qr = v->qr, *rhow = v->rhow, *rhou = v->rhou;
    double *fw = zalloc((size_t)nx * nz);
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double VT = 1E-2 * (3634 * pow(1E-3 * rhow[k] * 0.5 * (AT(qr, i, k) + AT(qr, i, k - 1)), 0.1346) *
                                pow(rhow[k] / rhow[1], -0.5));
            AT(fw, i, k) = rhow[k] * VT * (AT(qr, i, k) + AT(qr, i, k - 1));
        }
    }
    bc_center(fw, nx, nz);
    for (int i = 0; i < nx; i++) AT(fw, i, 0) = AT(fw, i, nz - 1) = 0.;
    const int sn = (v->step + 1) % 2, so = v->step % 2;
    double *d = v->dqr_VT;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double p = (AT(fw, i, k + 1) - AT(fw, i, k)) * g.r2dz / rhou[k];
            HS(d, i, k, sn) = p;
            if (v->step == 0) HS(d, i, k, 0) = HS(d, i, k, 1);
            AT(v->qrp, i, k) += 1.5 * g.dt * HS(d, i, k, sn) - 0.5 * g.dt * HS(d, i, k, so);
        }
    }
    free(fw);
}

/* MicroPhysics::NegativeValueProcess (src/MicroPhysics.cpp:116-138). */
int vo_negval(const vvm_dev_config *c, double *var) {
    const int nx = c->nx, nz = c->nz;
    double positive = 0., negative = 0.;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            if (AT(var, i, k) >= 0.) positive += AT(var, i, k);
            else {
                negative += AT(var, i, k);
                AT(var, i, k) = 0.;
            }
        }
    }
    if (positive == 0. || fabs(negative) > positive) return 0;
    double ratio = 1. - fabs(negative / positive);
    for (int k = 1; k <= nz - 2; k++)
        for (int i = 1; i <= nx - 2; i++)
            if (AT(var, i, k) > 0) AT(var, i, k) = AT(var, i, k) * ratio;
    return 0;
}

/* Iteration::pzeta_pt (src/Iteration.cpp:29-34). */
int vo_pzeta_pt(const vvm_dev_config *c, vvm_host_view *v) {
    adv_zeta(c, v);
    buoyancy(c, v);
    bc_westdown(v->zetap, c->nx, c->nz);
    return 0;
}

/* Iteration::pth_pt (src/Iteration.cpp:36-40). */
int vo_pth_pt(const vvm_dev_config *c, vvm_host_view *v) {
    adv_thermo(c, v, v->th, v->thp, v->dth_advect);
    bc_center(v->thp, c->nx, c->nz);
    return 0;
}

/* Iteration::pqv_pt (src/Iteration.cpp:90-98). */
int vo_pqv_pt(const vvm_dev_config *c, vvm_host_view *v) {
    adv_thermo(c, v, v->qv, v->qvp, v->dqv_advect);
    vo_negval(c, v->qvp);
    bc_center(v->qvp, c->nx, c->nz);
    return 0;
}

/* Iteration::pqc_pt (src/Iteration.cpp:100-107). */
int vo_pqc_pt(const vvm_dev_config *c, vvm_host_view *v) {
    adv_thermo(c, v, v->qc, v->qcp, v->dqc_advect);
    vo_negval(c, v->qcp);
    bc_center(v->qcp, c->nx, c->nz);
    return 0;
}

/* Iteration::pqr_pt (src/Iteration.cpp:109-124), including surface precip. */
int vo_pqr_pt(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    adv_thermo(c, v, v->qr, v->qrp, v->dqr_advect);
    adv_qrvt(c, v);
    vo_negval(c, v->qrp);
    const double *rhou = v->rhou, *rhow = v->rhow;
    for (int i = 1; i <= nx - 2; i++) {
        double VT = 1E-2 * (3634 * pow(1E-3 * rhou[1] * AT(v->qr, i, 1), 0.1346) * pow(rhou[1] / rhow[1], -0.5));
        double rain = -rhou[1] * (0.5 * (AT(v->w, i, 2) + 0.) - VT) * AT(v->qr, i, 1);
        v->precip[i] = (rain < 0) ? 0. : rain;
    }
    bc_center(v->qrp, nx, nz);
    return 0;
}

/* The radiation heating rate [K/day, [nx][nz]] and RADIATIONCOOLING of
 * AddForcing, set by the caller (vo_set_radiation); NULL / 0 by default,
 * which is what the benchmark configurations use. */
static const double *vo_rad_rate = NULL;
static int vo_rad_cooling = 0;
void vo_set_radiation(const double *rate, int cooling) {
    vo_rad_rate = rate;
    vo_rad_cooling = cooling;
}

/* vvm::AddForcing (src/AddForcing.cpp:6-28; Q1LS/Q2LS carry the configs'
 * surface fluxes) followed by BoundaryProcess2D_all (src/Iteration.cpp:
 * 283-285).  The radiation term (:20) and RADIATIONCOOLING (:21-23) apply
 * when vo_set_radiation gave them. */
/* vvm::AddForcing alone (src/AddForcing.cpp:6-28). */
static void add_forcing(const vvm_dev_config *c, vvm_host_view *v);

int vo_forcing_bc(const vvm_dev_config *c, vvm_host_view *v) {
    add_forcing(c, v);
    return vo_bc_all(c, v);
}

static void add_forcing(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    const double dt = c->dt;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            if (c->tropical_forcing) {
                AT(v->thp, i, k) += dt * v->Q1LS[k];
                AT(v->qvp, i, k) += dt * v->Q2LS[k];
            }
            if (vo_rad_rate) AT(v->thp, i, k) += dt * vo_rad_rate[(size_t)i * nz + k] / 86400.;
            if (vo_rad_cooling) AT(v->thp, i, k) += dt * (-2. / 86400.);
            if (v->status_for_adding_forcing) AT(v->thp, i, k) += dt * AT(v->init_th_forcing, i, k);
        }
    }
}

/* ------------------------------------------------------------ elliptic */

/* PoissonSolver::pubarTop_pt (src/PoissonSolver.cpp:655-676). */
int vo_pubartop_pt(const vvm_dev_config *c, vvm_host_view *v) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *u = v->u, *w = v->w, *K = v->RKM, *rhou = v->rhou, *rhow = v->rhow;
    double down = 0., kup = 0., kdown = 0.;
    for (int i = 1; i < nx - 1; i++) {
        down += 0.25 * (rhou[nz - 2] * AT(u, i, nz - 2) + rhou[nz - 3] * AT(u, i, nz - 3)) *
                (AT(w, i, nz - 2) + AT(w, i - 1, nz - 2));
        kup += rhow[nz - 1] * 0.25 * (AT(K, i, nz - 1) + AT(K, i, nz - 2) + AT(K, i - 1, nz - 1) + AT(K, i - 1, nz - 2)) *
               (g.rdz * (AT(u, i, nz - 1) - AT(u, i, nz - 2)));
        kdown += rhow[nz - 2] * 0.25 *
                 (AT(K, i, nz - 2) + AT(K, i, nz - 3) + AT(K, i - 1, nz - 2) + AT(K, i - 1, nz - 3)) *
                 (g.rdx * (AT(w, i, nz - 2) - AT(w, i - 1, nz - 2)) + g.rdz * (AT(u, i, nz - 2) - AT(u, i, nz - 3)));
    }
    down /= ((double)(nx - 2.));
    kup /= ((double)(nx - 2.));
    kdown /= ((double)(nx - 2.));
    double p1 = (-down) * g.rdz / rhou[nz - 2];
    double p2 = (kup - kdown) * g.rdz / rhou[nz - 2];
    const int sn = (v->step + 1) % 2, so = v->step % 2;
    v->dubarTop_advect[sn] = -p1 + p2;
    if (v->step == 0) v->dubarTop_advect[0] = v->dubarTop_advect[1];
    v->ubarTopp = v->ubarTop + 1.5 * g.dt * v->dubarTop_advect[sn] - 0.5 * g.dt * v->dubarTop_advect[so];
    return 0;
}

/* In-place complex DFT of length n (sign -1 forward, +1 backward, no
 * scaling): iterative radix-2 for powers of two, direct sum otherwise. */
static void dft(double *re, double *im, int n, int sign) {
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1;
            for (int s = 0; s < n; s += len) {
                for (int j = 0; j < half; j++) {
                    double ang = sign * 2.0 * PI_D * j / len;
                    double wr = cos(ang), wi = sin(ang);
                    double ar = re[s + j + half], ai = im[s + j + half];
                    double tr = ar * wr - ai * wi, ti = ar * wi + ai * wr;
                    re[s + j + half] = re[s + j] - tr;
                    im[s + j + half] = im[s + j] - ti;
                    re[s + j] += tr;
                    im[s + j] += ti;
                }
            }
        }
        return;
    }
    double *or_ = zalloc((size_t)n), *oi = zalloc((size_t)n);
    for (int m = 0; m < n; m++) {
        double sr = 0., si = 0.;
        for (int j = 0; j < n; j++) {
            double ang = sign * 2.0 * PI_D * (double)(((long long)m * j) % n) / n;
            sr += re[j] * cos(ang) - im[j] * sin(ang);
            si += re[j] * sin(ang) + im[j] * cos(ang);
        }
        or_[m] = sr;
        oi[m] = si;
    }
    memcpy(re, or_, sizeof(double) * n);
    memcpy(im, oi, sizeof(double) * n);
    free(or_);
    free(oi);
}

/* The matrix A of InitPoissonMatrix (src/PoissonSolver.cpp:760-796) has, in
 * row (i,k), k = 2..nz-2: 1 to the periodic x neighbours, E_k = rhow[k]/rhou[k]
 * to (i,k+1) (absent at k = nz-2), F_k = rhow[k]/rhou[k-1] to (i,k-1)
 * (absent at k = 2) and -(2 + E_k + F_k) on the diagonal.  It is separable:
 * a DFT along i diagonalises the x part (eigenvalues 2 cos(2 pi m/N) - 2)
 * and leaves one real tridiagonal system in k per wavenumber. */
void vo_poisson_apply(const vvm_dev_config *c, const double *rhou, const double *rhow, const double *x, double *y) {
    const int N = c->nx - 2, nz = c->nz, M = nz - 3;
    for (int kk = 0; kk < M; kk++) {
        int k = kk + 2;
        double E = rhow[k] / rhou[k], F = rhow[k] / rhou[k - 1];
        double diag = -(2. + E + F);
        for (int ii = 0; ii < N; ii++) {
            size_t idx = (size_t)kk * N + ii;
            double s = diag * x[idx];
            s += x[(size_t)kk * N + (ii + N - 1) % N] + x[(size_t)kk * N + (ii + 1) % N];
            if (kk + 1 < M) s += E * x[idx + N];
            if (kk > 0) s += F * x[idx - N];
            y[idx] = s;
        }
    }
}

int vo_poisson_exact(const vvm_dev_config *c, const double *rhou, const double *rhow, const double *b, double *x) {
    const int N = c->nx - 2, nz = c->nz, M = nz - 3;
    if (N < 1 || M < 1) return -1;
    double *re = zalloc((size_t)N * M), *im = zalloc((size_t)N * M);
    for (int kk = 0; kk < M; kk++) {
        memcpy(re + (size_t)kk * N, b + (size_t)kk * N, sizeof(double) * N);
        dft(re + (size_t)kk * N, im + (size_t)kk * N, N, -1);
    }
    double *cp = zalloc((size_t)M), *dr = zalloc((size_t)M), *di = zalloc((size_t)M);
    for (int m = 0; m < N; m++) {
        /* x neighbours contribute 2 cos(2 pi m/N); the -2 is in the diagonal */
        double lam = 2. * cos(2. * PI_D * m / N);
        /* Thomas algorithm: sub F_k, super E_k, diagonal lam - (2 + E + F). */
        for (int kk = 0; kk < M; kk++) {
            int k = kk + 2;
            double E = rhow[k] / rhou[k], F = rhow[k] / rhou[k - 1];
            double diag = lam - (2. + E + F);
            double sub = (kk > 0) ? F : 0., sup = (kk + 1 < M) ? E : 0.;
            double br = re[(size_t)kk * N + m], bi = im[(size_t)kk * N + m];
            if (kk > 0) {
                double den = diag - sub * cp[kk - 1];
                cp[kk] = sup / den;
                dr[kk] = (br - sub * dr[kk - 1]) / den;
                di[kk] = (bi - sub * di[kk - 1]) / den;
            } else {
                cp[kk] = sup / diag;
                dr[kk] = br / diag;
                di[kk] = bi / diag;
            }
        }
        for (int kk = M - 2; kk >= 0; kk--) {
            dr[kk] -= cp[kk] * dr[kk + 1];
            di[kk] -= cp[kk] * di[kk + 1];
        }
        for (int kk = 0; kk < M; kk++) {
            re[(size_t)kk * N + m] = dr[kk];
            im[(size_t)kk * N + m] = di[kk];
        }
    }
    for (int kk = 0; kk < M; kk++) {
        dft(re + (size_t)kk * N, im + (size_t)kk * N, N, +1);
        for (int ii = 0; ii < N; ii++) x[(size_t)kk * N + ii] = re[(size_t)kk * N + ii] / N;
    }
    free(re); free(im); free(cp); free(dr); free(di);
    return 0;
}

/* PoissonSolver::cal_w (src/PoissonSolver.cpp:439-477), Eigen branch: the
 * right-hand side and the back-substitution w = x / rhow are restated; the
 * solve itself is the exact direct method above. */
int vo_cal_w(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz, N = nx - 2, M = nz - 3;
    const double *rhow = v->rhow;
    double *b = zalloc((size_t)N * M), *x = zalloc((size_t)N * M);
    int count = 0;
    for (int k = 2; k <= nz - 2; k++)
        for (int i = 1; i <= nx - 2; i++)
            b[count++] = rhow[k] * rhow[k] * (AT(v->zetap, i + 1, k) - AT(v->zetap, i, k)) * c->dx;
    int rc = vo_poisson_exact(c, v->rhou, rhow, b, x);
    int cnt = 0;
    for (int k = 2; k <= nz - 2; k++)
        for (int i = 1; i <= nx - 2; i++) AT(v->w, i, k) = x[cnt++] / rhow[k];
    bc_westdown(v->w, nx, nz);
    free(b);
    free(x);
    return rc;
}

/* Eigen's conjugate_gradient (include/Eigen/src/IterativeLinearSolvers/
 * ConjugateGradient.h:27-91) with the identity preconditioner and zero
 * initial guess, on G = [1 -2 1] periodic (src/PoissonSolver.cpp:802-813). */
static void g_cg(const double *h, double *y, int n, double tol, int maxit) {
    double *r = zalloc((size_t)n), *p = zalloc((size_t)n), *t = zalloc((size_t)n);
    for (int i = 0; i < n; i++) { y[i] = 0.; r[i] = h[i]; }
    double rhs2 = 0.;
    for (int i = 0; i < n; i++) rhs2 += h[i] * h[i];
    if (rhs2 == 0) goto done;
    double thr = tol * tol * rhs2;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    double rn2 = rhs2;
    if (rn2 < thr) goto done;
    for (int i = 0; i < n; i++) p[i] = r[i];
    double absnew = 0.;
    for (int i = 0; i < n; i++) absnew += r[i] * p[i];
    for (int it = 0; it < maxit; it++) {
        for (int i = 0; i < n; i++) t[i] = p[(i + n - 1) % n] - 2. * p[i] + p[(i + 1) % n];
        double pt = 0.;
        for (int i = 0; i < n; i++) pt += p[i] * t[i];
        double alpha = absnew / pt;
        for (int i = 0; i < n; i++) { y[i] += alpha * p[i]; r[i] -= alpha * t[i]; }
        rn2 = 0.;
        for (int i = 0; i < n; i++) rn2 += r[i] * r[i];
        if (rn2 < thr) break;
        double absold = absnew;
        absnew = rn2;
        double beta = absnew / absold;
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
    }
done:
    free(r); free(p); free(t);
}

/* Mean-zero solution of G y = h by DFT (G's eigenvalues 2cos(2 pi m/n) - 2,
 * the m = 0 mode, G's null space, set to zero). */
static void g_exact(const double *h, double *y, int n) {
    double *re = zalloc((size_t)n), *im = zalloc((size_t)n);
    memcpy(re, h, sizeof(double) * n);
    dft(re, im, n, -1);
    re[0] = im[0] = 0.;
    for (int m = 1; m < n; m++) {
        double lam = 2. * cos(2. * PI_D * m / n) - 2.;
        re[m] /= lam;
        im[m] /= lam;
    }
    dft(re, im, n, +1);
    for (int i = 0; i < n; i++) y[i] = re[i] / n;
    free(re);
    free(im);
}

/* PoissonSolver::cal_u (src/PoissonSolver.cpp:580-653), Eigen branch. */
int vo_cal_u(const vvm_dev_config *c, vvm_host_view *v, int gsolve) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz, n = nx - 2;
    const double *rhow = v->rhow, *rhou = v->rhou, *w = v->w;
    double *u = v->u;
    double *h = zalloc((size_t)n), *y = zalloc((size_t)n);
    double tmp = 0.;
    for (int i = 1; i <= nx - 2; i++) {
        h[i - 1] = -(0. - rhow[nz - 2] * AT(w, i, nz - 2)) / rhou[nz - 2] * g.dx;
        tmp += h[i - 1];
    }
    tmp /= (nx - 2);
    for (int i = 1; i <= nx - 2; i++) h[i - 1] = h[i - 1] - tmp;
    if (gsolve == VO_GSOLVE_CG) g_cg(h, y, n, 1E-12, 10000);
    else g_exact(h, y, n);
    double *xi = v->xi, *uxi = v->uxi;
    for (int i = 1; i <= nx - 2; i++) xi[i] = y[i - 1];
    xi[0] = xi[nx - 2];
    xi[nx - 1] = xi[1];
    for (int i = 1; i <= nx - 2; i++) uxi[i] = (xi[i] - xi[i - 1]) * g.rdx;
    tmp = 0.;
    for (int i = 1; i < nx - 1; i++) tmp += uxi[i];
    tmp /= (nx - 2);
    for (int i = 1; i <= nx - 2; i++) uxi[i] -= tmp;
    uxi[0] = uxi[nx - 2];
    uxi[nx - 1] = uxi[1];
    for (int i = 1; i <= nx - 2; i++) AT(u, i, nz - 2) = uxi[i] + v->ubarTopp;
    AT(u, 0, nz - 2) = AT(u, nx - 2, nz - 2);
    AT(u, nx - 1, nz - 2) = AT(u, 1, nz - 2);
    for (int i = 1; i <= nx - 2; i++) {
        double area = 0.;
        for (int k = nz - 3; k >= 1; k--) {
            area += ((AT(w, i, k + 1) - AT(w, i - 1, k + 1)) * g.rdx - rhow[k + 1] * AT(v->zetap, i, k + 1)) * -g.dz;
            AT(u, i, k) = area + AT(u, i, nz - 2);
        }
    }
    bc_center(u, nx, nz);
    free(h);
    free(y);
    return 0;
}

/* --------------------------------------------------------- microphysics */

/* MicroPhysics::autoconversion, accretion, evaporation, condensation
 * (src/MicroPhysics.cpp:29-51, 54-77, 80-114, 7-26) in the step's call order
 * (src/Iteration.cpp:303-306), then NegativeValueProcess on qvp, qcp, qrp
 * (:309-311) and BoundaryProcess2D_all (:404). */
/* MicroPhysics::autoconversion (src/MicroPhysics.cpp:29-51). */
static void autoconversion(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    const double dt = c->dt;
    double *qcp = v->qcp, *qrp = v->qrp;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double qcplus = MAXS(0., AT(qcp, i, k));
            double ar = 0.001 * (qcplus - 0.001);
            ar = MAXS(0., ar);
            double arcrdt = MINS(ar * dt, qcplus);
            AT(qcp, i, k) = AT(qcp, i, k) - arcrdt;
            AT(qrp, i, k) = AT(qrp, i, k) + arcrdt;
            AT(v->autoconversion, i, k) = arcrdt;
        }
    }
}

/* MicroPhysics::accretion (src/MicroPhysics.cpp:54-77). */
static void accretion(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    const double dt = c->dt;
    double *qcp = v->qcp, *qrp = v->qrp;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double qcplus = MAXS(0., AT(qcp, i, k));
            double qrplus = MAXS(0., AT(qrp, i, k));
            double cr = 2.2 * qcplus * (pow(qrplus, 0.875));
            double arcrdt = MINS(cr * dt, qcplus);
            AT(qcp, i, k) = AT(qcp, i, k) - arcrdt;
            AT(qrp, i, k) = AT(qrp, i, k) + arcrdt;
            AT(v->accretion, i, k) = arcrdt;
        }
    }
}

/* MicroPhysics::evaporation (src/MicroPhysics.cpp:80-114). */
static void evaporation(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    const double dt = c->dt;
    double *qrp = v->qrp, *qvp = v->qvp, *thp = v->thp;
    const double *pib = v->pib, *rhou = v->rhou, *pb = v->pb;
    for (int k = 1; k <= nz - 2; k++) {
        double pc = 380. / (pow(pib[k], c->Cp / c->Rd) * c->P0);
        for (int i = 1; i <= nx - 2; i++) {
            double qrplus = MAXS(0., AT(qrp, i, k));
            double qvplus = MAXS(0., AT(qvp, i, k));
            double qvs = pc * exp(17.27 * (pib[k] * AT(thp, i, k) - 273.) / (pib[k] * AT(thp, i, k) - 36.));
            double deficit = MAXS((1. - qvplus / qvs), 0.);
            double coef = 1.6 + 124.9 * pow((1E-3 * rhou[k] * qrplus), 0.2046);
            double er = coef * deficit * (pow(1E-3 * rhou[k] * qrplus, 0.525)) /
                        ((5.4E5 + 2.55E6 / (1E-2 * pb[k] * qvs)) * 1E-3 * rhou[k]);
            double erdt = MINS(qrplus, MAXS(0., er * dt));
            AT(qrp, i, k) = AT(qrp, i, k) - erdt;
            AT(qvp, i, k) = AT(qvp, i, k) + erdt;
            AT(thp, i, k) = AT(thp, i, k) - c->Lv * erdt / (c->Cp * pib[k]);
            AT(v->evaporation, i, k) = erdt;
        }
    }
}

/* MicroPhysics::condensation (src/MicroPhysics.cpp:7-26): saturation
 * adjustment with Tetens' formula. */
static void condensation(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    double *qcp = v->qcp, *qvp = v->qvp, *thp = v->thp;
    const double *pib = v->pib;
    for (int k = 1; k <= nz - 2; k++) {
        double pc = 380. / (pow(pib[k], c->Cp / c->Rd) * c->P0);
        for (int i = 1; i <= nx - 2; i++) {
            double qvs = pc * exp(17.27 * (pib[k] * AT(thp, i, k) - 273.) / (pib[k] * AT(thp, i, k) - 36.));
            double phi = qvs * (17.27 * 237. * c->Lv) / (c->Cp * pow(AT(thp, i, k) * pib[k] - 36., 2));
            double C = (AT(qvp, i, k) - qvs) / (1 + phi);
            if (fabs(C) > AT(qcp, i, k) && C < 0) C = -AT(qcp, i, k);
            AT(qvp, i, k) = AT(qvp, i, k) - C;
            AT(qcp, i, k) = AT(qcp, i, k) + C;
            AT(thp, i, k) = AT(thp, i, k) + c->Lv / (c->Cp * pib[k]) * C;
            AT(v->condensation, i, k) = C;
        }
    }
}

int vo_microphysics(const vvm_dev_config *c, vvm_host_view *v) {
    autoconversion(c, v);
    accretion(c, v);
    evaporation(c, v);
    condensation(c, v);
    vo_negval(c, v->qvp);
    vo_negval(c, v->qcp);
    vo_negval(c, v->qrp);
    return vo_bc_all(c, v);
}

/* ------------------------------------------------------------ turbulence */

/* Turbulence::Mparam (src/Turbulent.cpp:68-90): in-place sweep, k outer,
 * i inner, so west and lower neighbours are already updated. */
static void mparam(const vvm_dev_config *c, vvm_host_view *v, double *f) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *K = v->RKM, *rhou = v->rhou, *rhow = v->rhow;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            AT(f, i, k) += g.rdx2 * g.dt *
                               (0.5 * (AT(K, i, k) + AT(K, i, k - 1)) * (AT(f, i + 1, k) - AT(f, i, k)) -
                                0.5 * (AT(K, i - 1, k) + AT(K, i - 1, k - 1)) * (AT(f, i, k) - AT(f, i - 1, k))) +
                           g.rdz2 * g.dt / rhow[k] *
                               (rhou[k] * 0.5 * (AT(K, i, k) + AT(K, i - 1, k)) * (AT(f, i, k + 1) - AT(f, i, k)) -
                                rhou[k - 1] * 0.5 * (AT(K, i, k - 1) + AT(K, i - 1, k - 1)) *
                                    (AT(f, i, k) - AT(f, i, k - 1)));
        }
    }
}

/* Turbulence::Hparam (src/Turbulent.cpp:92-114): same sweep order. */
static void hparam(const vvm_dev_config *c, vvm_host_view *v, double *f) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *K = v->RKH, *rhou = v->rhou, *rhow = v->rhow;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            AT(f, i, k) += g.rdx2 * g.dt *
                               (0.5 * (AT(K, i + 1, k) + AT(K, i, k)) * (AT(f, i + 1, k) - AT(f, i, k)) -
                                0.5 * (AT(K, i, k) + AT(K, i - 1, k)) * (AT(f, i, k) - AT(f, i - 1, k))) +
                           g.rdz2 * g.dt / rhou[k] *
                               (rhow[k + 1] * 0.5 * (AT(K, i, k + 1) + AT(K, i, k)) * (AT(f, i, k + 1) - AT(f, i, k)) -
                                rhow[k] * 0.5 * (AT(K, i, k) + AT(K, i, k - 1)) * (AT(f, i, k) - AT(f, i, k - 1)));
        }
    }
}

/* The coefficient part of Turbulence::RKM_RKH (src/Turbulent.cpp:5-45):
 * K_M and K_H from the Richardson number, clamped, with center halos. */
static int rkm_rkh_coeffs(const vvm_dev_config *c, vvm_host_view *v) {
    const grid_t g = mkgrid(c);
    const int nx = g.nx, nz = g.nz;
    const double *u = v->u, *w = v->w, *thp = v->thp, *l2 = v->lambda2;
    const double kmax = 0.8 * g.dx * g.dz / g.dt;
    for (int k = 1; k <= nz - 2; k++) {
        for (int i = 1; i <= nx - 2; i++) {
            double Rz = 0.5 * ((AT(w, i + 1, k + 1) + AT(w, i + 1, k)) - (AT(w, i - 1, k + 1) + AT(w, i - 1, k))) * g.r2dx +
                        0.5 * ((AT(u, i + 1, k + 1) + AT(u, i, k + 1)) - (AT(u, i + 1, k - 1) + AT(u, i, k - 1))) * g.r2dz;
            double Ro = pow((AT(u, i + 1, k) - AT(u, i, k)) * g.rdx + (AT(w, i, k + 1) - AT(w, i, k)) * g.rdz, 2);
            double Ri = (c->GRAVITY / AT(thp, i, k) * (AT(thp, i, k + 1) - AT(thp, i, k - 1)) * g.r2dz) /
                        (pow(Rz, 2) + 2. * Ro);
            double km, kh;
            if (Ri < 0) {
                km = l2[k] * sqrt(pow(Rz, 2) + 2. * Ro) * sqrt(1. - 16. * Ri);
                kh = l2[k] * sqrt(pow(Rz, 2) + 2. * Ro) * 1.4 * sqrt(1. - 40. * Ri);
            } else if (0 < Ri && Ri < 0.25) {
                km = l2[k] * sqrt(pow(Rz, 2) + 2. * Ro) * pow(1. - 4. * Ri, 4);
                kh = l2[k] * sqrt(pow(Rz, 2) + 2. * Ro) * 1.4 * (1. - 1.2 * Ri) * pow(1. - 4. * Ri, 4);
            } else {
                km = 0.;
                kh = 0.;
            }
            km = MAXS(km, 10.);
            kh = MAXS(kh, 10.);
            km = MINS(km, kmax);
            kh = MINS(kh, kmax);
            AT(v->RKM, i, k) = km;
            AT(v->RKH, i, k) = kh;
        }
    }
    bc_center(v->RKM, nx, nz);
    bc_center(v->RKH, nx, nz);
    return 0;
}

/* Turbulence::RKM_RKH (src/Turbulent.cpp:5-66): the coefficients, then
 * Mparam on zetap and Hparam on thp, qvp, qcp, qrp (:47-52). */
int vo_turbulence(const vvm_dev_config *c, vvm_host_view *v) {
    rkm_rkh_coeffs(c, v);
    mparam(c, v, v->zetap);
    hparam(c, v, v->thp);
    hparam(c, v, v->qvp);
    hparam(c, v, v->qcp);
    hparam(c, v, v->qrp);
    return 0;
}

/* NumericalProcess::Nudge_theta (src/NumericalProcess.cpp:60-71). */
static void nudge_theta(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    const int ks = (int)(12000. / c->dz + 1);
    const int kr = (nz - 1) - ks;
    const double crad = 1. / 3600.;
    if (kr < 0) return;
    for (int k = ks; k < nz - 1; k++) {
        double cgr = crad * (v->z[k] - v->z[kr]) / (v->z[nz - 1] - v->z[kr]);
        for (int i = 1; i < nx - 1; i++) AT(v->thp, i, k) -= c->dt * cgr * (AT(v->thp, i, k) - v->thb_init[k]);
    }
}

/* NumericalProcess::Nudge_zeta (src/NumericalProcess.cpp:73-85). */
static void nudge_zeta(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    const int ks = (int)(12000. / c->dz + 1);
    const int kr = (nz - 1) - ks;
    if (kr < 0) return;
    for (int k = ks; k < nz - 1; k++) {
        double cgr = (1. / 360.) * (v->z_zeta[k] - v->z_zeta[kr]) / (v->z_zeta[nz - 1] - v->z_zeta[kr]);
        for (int i = 1; i < nx - 1; i++) AT(v->zetap, i, k) -= c->dt * cgr * (AT(v->zetap, i, k));
    }
}

/* NumericalProcess::Nudge_qv (src/NumericalProcess.cpp:87-96): every point,
 * ghosts included; off when moisture_nudge_time == 0. */
static void nudge_qv(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    if (c->moisture_nudge_time == 0) return;
    for (int i = 0; i <= nx - 1; i++)
        for (int k = 0; k <= nz - 1; k++)
            AT(v->qvp, i, k) = AT(v->qvp, i, k) + c->dt * (v->qvb0[k] - AT(v->qv, i, k)) / c->moisture_nudge_time;
}

/* The three as called at src/Iteration.cpp:415-417. */
int vo_nudge(const vvm_dev_config *c, vvm_host_view *v) {
    nudge_theta(c, v);
    if (c->CASE != 2) nudge_zeta(c, v);
    nudge_qv(c, v);
    return 0;
}

/* ------------------------------------------------------------- stepping */

/* Iteration::updateMean (src/Iteration.cpp:42-86), WATER + AB2. */
int vo_update_mean(const vvm_dev_config *c, vvm_host_view *v) {
    const int nx = c->nx, nz = c->nz;
    for (int k = 0; k < nz; k++) {
        v->thbm[k] = v->thb[k];
        v->thvbm[k] = v->thvb[k];
    }
    for (int k = 1; k < nz - 1; k++) {
        double tb = 0., qb = 0.;
        for (int i = 1; i < nx - 1; i++) {
            tb += AT(v->thp, i, k);
            qb += AT(v->qvp, i, k);
        }
        v->thb[k] = tb / (double)(nx - 2.);
        v->qvb[k] = qb / (double)(nx - 2.);
        v->thvb[k] = v->thb[k] * (1 + 0.608 * v->qvb[k]);
    }
    bc1d_center(v->thb, nz);
    bc1d_center(v->thvb, nz);
    bc1d_center(v->qvb, nz);
    for (int k = 1; k < nz - 1; k++) v->thb_zeta[k] = 0.5 * (v->thb[k - 1] + v->thb[k]);
    v->thb_zeta[0] = v->thb_zeta[1];
    v->thb_zeta[nz - 1] = v->thb_zeta[nz - 2];
    return 0;
}

/* Iteration::nextTimeStep (src/Iteration.cpp:153-204): m <- now <- p for the
 * five prognostics, ghosts included; the ubarTop shift inside the copy loop
 * leaves ubarTopm = ubarTop = ubarTopp. */
int vo_next_time_step(const vvm_dev_config *c, vvm_host_view *v) {
    const size_t n = (size_t)c->nx * c->nz;
    vo_update_mean(c, v);
    double *trip[5][3] = {{v->zetam, v->zeta, v->zetap}, {v->thm, v->th, v->thp}, {v->qvm, v->qv, v->qvp},
                          {v->qcm, v->qc, v->qcp},       {v->qrm, v->qr, v->qrp}};
    for (int f = 0; f < 5; f++) {
        memcpy(trip[f][0], trip[f][1], n * sizeof(double));
        memcpy(trip[f][1], trip[f][2], n * sizeof(double));
    }
    if (n >= 2) v->ubarTopm = v->ubarTopp;
    else v->ubarTopm = v->ubarTop;
    v->ubarTop = v->ubarTopp;
    return 0;
}

/* One TimeMarching loop body (src/Iteration.cpp:246-435), Kessler build,
 * radiation off, per call. */
int vo_step(const vvm_dev_config *c, vvm_host_view *v, int nsteps, int gsolve) {
    for (int s = 0; s < nsteps; s++) {
        v->step++;
        vo_pzeta_pt(c, v);
        vo_pth_pt(c, v);
        vo_pqv_pt(c, v);
        vo_pqc_pt(c, v);
        vo_pqr_pt(c, v);
        v->status_for_adding_forcing = (v->step * c->dt <= c->addforcingtime) ? 1 : 0;
        /* a new random theta kick every step it is applied
         * (src/Iteration.cpp:276-282) */
        if (v->status_for_adding_forcing && vo_random_perturbation(c, v, v->step + 0, -0.001, 0.001, 1.) != 0)
            return -1;
        vo_forcing_bc(c, v);
        vo_pubartop_pt(c, v);
        if (vo_cal_w(c, v) != 0) return -1;
        vo_cal_u(c, v, gsolve);
        vo_bc_all(c, v);
        vo_microphysics(c, v);
        vo_update_mean(c, v);
        vo_turbulence(c, v);
        vo_nudge(c, v);
        vo_next_time_step(c, v);
    }
    return 0;
}

/* ------------------------------------------------------ single operators */

/* The array a field id (VVM_FIELD_AT) names. */
static double *field_array(vvm_host_view *v, int id) {
    double *arr[3][5] = {{v->zetam, v->thm, v->qvm, v->qcm, v->qrm},
                         {v->zeta, v->th, v->qv, v->qc, v->qr},
                         {v->zetap, v->thp, v->qvp, v->qcp, v->qrp}};
    const int fld = id % 8, lev = id / 8;
    if (fld < 0 || fld > 4 || lev < 0 || lev > 2) return NULL;
    return arr[lev][fld];
}

/* One static of class vvm (src/Declare.hpp:707-935) alone, by its VVM_OP_*
 * id (include/vvm_dev.h); `field` as the device entry vvm_dev_op_field
 * takes it.  Returns -1 for an unknown op or field. */
int vo_op_single(const vvm_dev_config *c, vvm_host_view *v, int op, int field) {
    double *now[5] = {v->zeta, v->th, v->qv, v->qc, v->qr};
    double *fut[5] = {v->zetap, v->thp, v->qvp, v->qcp, v->qrp};
    double *hist[5] = {v->dzeta_advect, v->dth_advect, v->dqv_advect, v->dqc_advect, v->dqr_advect};
    const int scalar = (field >= VVM_FIELD_TH && field <= VVM_FIELD_QR);
    switch (op) {
        case VVM_OP_ADVECTION_ZETA: adv_zeta(c, v); return 0;
        case VVM_OP_BOUYANCY: buoyancy(c, v); return 0;
        case VVM_OP_ADVECTION_THERMO:
            if (!scalar) return -1;
            adv_thermo(c, v, now[field], fut[field], hist[field]);
            return 0;
        case VVM_OP_ADVECTION_QRVT: adv_qrvt(c, v); return 0;
        case VVM_OP_PTH: return vo_pth_pt(c, v);
        case VVM_OP_PQV: return vo_pqv_pt(c, v);
        case VVM_OP_PQC: return vo_pqc_pt(c, v);
        case VVM_OP_PQR: return vo_pqr_pt(c, v);
        case VVM_OP_UPDATE_MEAN: return vo_update_mean(c, v);
        case VVM_OP_AUTOCONVERSION: autoconversion(c, v); return 0;
        case VVM_OP_ACCRETION: accretion(c, v); return 0;
        case VVM_OP_EVAPORATION: evaporation(c, v); return 0;
        case VVM_OP_CONDENSATION: condensation(c, v); return 0;
        case VVM_OP_NEGATIVE_VALUE: {
            double *a = field_array(v, field);
            if (!a) return -1;
            return vo_negval(c, a);
        }
        case VVM_OP_MPARAM: mparam(c, v, v->zetap); return 0;
        case VVM_OP_HPARAM:
            if (!scalar) return -1;
            hparam(c, v, fut[field]);
            return 0;
        case VVM_OP_NUDGE_THETA: nudge_theta(c, v); return 0;
        case VVM_OP_NUDGE_ZETA: nudge_zeta(c, v); return 0;
        case VVM_OP_NUDGE_QV: nudge_qv(c, v); return 0;
        case VVM_OP_ADD_FORCING: add_forcing(c, v); return 0;
        default: return -1;
    }
}
