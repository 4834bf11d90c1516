// host_init.cpp -- analytic base state and initial fields of the benchmark
// cases, plus the reference's random theta kick.  Host code only.
//
// The base-state derivation follows vvm::Init::Init1d's analytic branch
// (src/Init.cpp:14-135): hydrostatic Exner function from the virtual
// potential temperature, rho_u = P0 pi^(Cv/Rd) / (Rd thv), rho_w as the
// mean of adjacent rho_u with a linear extrapolation at k=1, Tetens q_vs.
// The sounding differs: theta(z) = 300 K exp(N^2 z / g) and q_v = RH(z) q_vs
// with RH = max(0.8 (1 - z/12 km), 0.1).  q_vs depends on p, which depends on
// q_v through theta_v; this is resolved by a fixed two-pass scheme (dry
// hydrostatic pass, then one moist pass).
#include "vvm_dev.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

inline double &at(double *a, int nz, int i, int k) { return a[size_t(i) * nz + k]; }

void bc1d_center(double *a, int n) {
    a[0] = a[1];
    a[n - 1] = a[n - 2];
}

void bc2d_center(double *a, int nx, int nz) {
    for (int k = 1; k <= nz - 2; k++) {
        at(a, nz, 0, k) = at(a, nz, nx - 2, k);
        at(a, nz, nx - 1, k) = at(a, nz, 1, k);
    }
    for (int i = 0; i < nx; i++) {
        at(a, nz, i, 0) = at(a, nz, i, 1);
        at(a, nz, i, nz - 1) = at(a, nz, i, nz - 2);
    }
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Counter-based uniform in [0,1): one independent draw per (seed, i, k, stream).
double uniform01(uint64_t seed, int i, int k, uint64_t stream) {
    uint64_t key = seed * 0x9E3779B97F4A7C15ull + uint64_t(i) * 0xD1B54A32D192ED03ull +
                   uint64_t(k) * 0x8CB92BA72F3D8DD7ull + stream * 0xF1357AEA2E62A9C5ull;
    return double(splitmix64(key) >> 11) * 0x1.0p-53;
}

double tetens_qvs(double p, double T) { return (380. / p) * std::exp((17.27 * (T - 273.)) / (T - 36.)); }

}  // namespace

extern "C" int64_t vvm_field_len(int nx, int nz) { return int64_t(nx) * nz; }
extern "C" int64_t vvm_hist_len(int nx, int nz) { return int64_t(nx) * nz * 2; }

extern "C" int vvm_init_case(const vvm_dev_config *c, const vvm_case_params *p, vvm_host_view *v) {
    if (!c || !p || !v) return VVM_ERR_ARG;
    const int nx = c->nx, nz = c->nz;
    if (nx < 6 || nz < 6) return VVM_ERR_SIZE;
    const double dx = c->dx, dz = c->dz, g = c->GRAVITY;
    const double N2 = p->brunt * p->brunt;

    std::vector<double> thb(nz), qvb(nz, 0.), thvb(nz), pib(nz), rhou(nz), rhow(nz), pb(nz), qvsb(nz), thbz(nz);
    for (int k = 1; k <= nz - 2; k++) thb[k] = 300. * std::exp(N2 * (k - 0.5) * dz / g);
    bc1d_center(thb.data(), nz);

    const double pisfc = std::pow(c->PSURF / c->P0, c->Rd / c->Cp);
    for (int pass = 0; pass < 2; pass++) {
        for (int k = 1; k <= nz - 2; k++) thvb[k] = thb[k] * (1. + 0.61 * qvb[k]);
        bc1d_center(thvb.data(), nz);
        for (int k = 1; k <= nz - 2; k++) {
            if (k == 1) pib[k] = pisfc - g * 0.5 * dz / (c->Cp * thvb[k]);
            else pib[k] = pib[k - 1] - g * dz / (c->Cp * 0.5 * (thvb[k] + thvb[k - 1]));
        }
        bc1d_center(pib.data(), nz);
        for (int k = 1; k <= nz - 2; k++) {
            pb[k] = c->P0 * std::pow(pib[k], c->Cp / c->Rd);
            const double z = (k - 0.5) * dz;
            const double rh = std::fmax(0.8 * (1. - z / 12000.), 0.1);
            qvsb[k] = tetens_qvs(pb[k], thb[k] * pib[k]);
            qvb[k] = rh * qvsb[k];
        }
        bc1d_center(pb.data(), nz);
        bc1d_center(qvsb.data(), nz);
        bc1d_center(qvb.data(), nz);
    }
    // Density from the second pass's theta_v (src/Init.cpp:64-82).
    for (int k = 1; k <= nz - 2; k++) rhou[k] = c->P0 * std::pow(pib[k], c->Cv / c->Rd) / (c->Rd * thvb[k]);
    bc1d_center(rhou.data(), nz);
    for (int k = 2; k <= nz - 1; k++) {
        thbz[k] = 0.5 * (thb[k] + thb[k - 1]);
        rhow[k] = 0.5 * (rhou[k] + rhou[k - 1]);
    }
    thbz[1] = thbz[2] - (thbz[3] - thbz[2]);
    rhow[1] = rhow[2] - (rhow[3] - rhow[2]);
    thbz[0] = thbz[1];
    rhow[0] = rhow[1];
    rhou[0] = rhow[0];

    for (int k = 0; k < nz; k++) {
        v->thb[k] = thb[k];
        v->thb_init[k] = thb[k];
        v->thbm[k] = thb[k];
        v->qvb[k] = qvb[k];
        v->qvb0[k] = qvb[k];
        v->qvsb[k] = qvsb[k];
        v->pib[k] = pib[k];
        v->pb[k] = pb[k];
        v->rhou[k] = rhou[k];
        v->rhow[k] = rhow[k];
        v->thb_zeta[k] = thbz[k];
        // src/Init.cpp:119 sets thvb = thb + 0.61 qvb after the derivation.
        v->thvb[k] = v->thvbm[k] = thb[k] + 0.61 * qvb[k];
        v->Q1LS[k] = 0.;
        v->Q2LS[k] = 0.;
    }
    for (int k = 1; k < nz - 1; k++) {
        v->z[k] = (k - 0.5) * dz;
        v->z_zeta[k] = (k - 1) * dz;
        v->lambda2[k] = 1. / (1. / std::pow(0.23 * std::sqrt(dx * dz), 2) + 1. / std::pow(0.4 * 0.4 * v->z[k], 2));
    }
    v->z[0] = v->z[1];
    v->z[nz - 1] = v->z[nz - 2];
    v->z_zeta[0] = v->z_zeta[1];
    v->z_zeta[nz - 1] = v->z_zeta[nz - 2];
    v->lambda2[0] = v->lambda2[nz - 1] = 0.;

    // Surface fluxes as level-1 tendencies in the TROPICALFORCING slot
    // (src/AddForcing.cpp:15-18): dtheta/dt = H / (rho Cp pi dz) at k=1 and a
    // moisture tendency ramping linearly from sfc_qv_ramp at z=0 to 0 at
    // sfc_qv_ramp_z.
    if (p->sfc_heat_flux != 0.) v->Q1LS[1] = p->sfc_heat_flux / (rhou[1] * c->Cp * pib[1] * dz);
    if (p->sfc_qv_ramp != 0. && p->sfc_qv_ramp_z > 0.) {
        for (int k = 1; k <= nz - 2; k++) {
            const double z = (k - 0.5) * dz;
            const double f = 1. - z / p->sfc_qv_ramp_z;
            if (f > 0.) v->Q2LS[k] = p->sfc_qv_ramp / 86400. * f;
        }
    }

    const size_t n2 = size_t(nx) * nz;
    double *zero2[] = {v->zeta, v->zetam, v->zetap, v->thp, v->qvp, v->qc, v->qcm, v->qcp, v->qr, v->qrm,
                       v->qrp,  v->u,     v->w,     v->RKM, v->RKH, v->init_th_forcing, v->evaporation,
                       v->accretion, v->autoconversion, v->condensation};
    for (double *a : zero2)
        if (a) std::memset(a, 0, n2 * sizeof(double));
    double *zero3[] = {v->dth_advect, v->dth_buoyancy, v->dzeta_advect, v->dqv_advect,
                       v->dqc_advect, v->dqr_advect,   v->dqr_VT};
    for (double *a : zero3)
        if (a) std::memset(a, 0, 2 * n2 * sizeof(double));
    for (double *a : {v->precip, v->xi, v->uxi})
        if (a) std::memset(a, 0, size_t(nx) * sizeof(double));

    // theta: base state + cosine bubble (src/Init.cpp:277-290 shape, centred
    // at XRANGE/2 with XRANGE including the ghost cells) + seeded noise.
    const double xc = nx * dx / 2.;
    for (int i = 1; i <= nx - 2; i++) {
        for (int k = 1; k <= nz - 2; k++) {
            const double x = (i - 0.5) * dx, z = (k - 0.5) * dz;
            double th = thb[k];
            if (p->bubble_dT != 0. && p->bubble_radius > 0.) {
                const double r = std::sqrt(std::pow((x - xc) / p->bubble_radius, 2) +
                                           std::pow((z - p->bubble_zc) / p->bubble_radius, 2));
                if (r <= 1.) th += 0.5 * p->bubble_dT * (std::cos(M_PI * r) + 1.);
            }
            if (p->noise_kind != 0 && z < p->noise_ztop) {
                if (p->noise_kind == 1) {
                    th += p->noise_amp * (2. * uniform01(p->seed, i, k, 0) - 1.);
                } else {
                    const double u1 = 1. - uniform01(p->seed, i, k, 0), u2 = uniform01(p->seed, i, k, 1);
                    th += p->noise_amp * std::sqrt(-2. * std::log(u1)) * std::cos(2. * M_PI * u2);
                }
            }
            at(v->th, nz, i, k) = th;
        }
    }
    bc2d_center(v->th, nx, nz);
    std::memcpy(v->thm, v->th, n2 * sizeof(double));
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) at(v->qv, nz, i, k) = at(v->qvm, nz, i, k) = qvb[k];

    v->ubarTop = v->ubarTopp = v->ubarTopm = 0.;
    v->dubarTop_advect[0] = v->dubarTop_advect[1] = 0.;
    v->step = 0;
    v->status_for_adding_forcing = 0;
    return VVM_OK;
}

// vvm::Init::RandomPerturbation (src/Init.cpp:488-515): the same standard
// library engine and distribution, drawn in the same (k outer, i inner)
// order with rejection outside [min_range, max_range], only below 200 m.
extern "C" int vvm_random_perturbation(const vvm_dev_config *c, vvm_host_view *v, int t, double min_range,
                                       double max_range, double stddev) {
    if (!c || !v || !v->init_th_forcing) return VVM_ERR_ARG;
    const int nx = c->nx, nz = c->nz;
    std::mt19937 gen(t);
    gen.seed(t);
    std::normal_distribution<> dist(0.0, 1.0);
    for (int k = 1; k < nz - 1; k++) {
        const double z = (k - 0.5) * c->dz;
        for (int i = 1; i < nx - 1; i++) {
            double val = 0.;
            if (z < 200) {
                do {
                    val = 0. + stddev * dist(gen);
                } while (val < min_range || val > max_range);
            }
            at(v->init_th_forcing, nz, i, k) = val;
        }
    }
    bc2d_center(v->init_th_forcing, nx, nz);
    return VVM_OK;
}
