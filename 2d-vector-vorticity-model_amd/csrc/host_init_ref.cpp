// host_init_ref.cpp -- the reference's own initial state: vvm::Init::Init1d
// (analytic sounding, src/Init.cpp:14-135, or the 10-column sounding file
// of LoadFile, src/Init.cpp:302-370) followed by vvm::Init::Init2d
// (src/Init.cpp:137-266).  Host code only; the statements follow the
// reference's order so that every value rounds the same way.
//
// Omitted: pib_lev, pb_lev and pb_lev_all, which only P3 and RRTMGP read.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "vvm_dev.h"

namespace {

inline double &at(double *a, int nz, int i, int k) { return a[size_t(i) * nz + k]; }

void bc1d(double *a, int n) {  // BoundaryProcess1D_center (src/Boundary.cpp:3-7)
    a[0] = a[1];
    a[n - 1] = a[n - 2];
}

void bc2d_center(double *a, int nx, int nz) {  // src/Boundary.cpp:9-19
    for (int k = 1; k <= nz - 2; k++) {
        at(a, nz, 0, k) = at(a, nz, nx - 2, k);
        at(a, nz, nx - 1, k) = at(a, nz, 1, k);
    }
    for (int i = 0; i < nx; i++) {
        at(a, nz, i, 0) = at(a, nz, i, 1);
        at(a, nz, i, nz - 1) = at(a, nz, i, nz - 2);
    }
}

void bc2d_westdown(double *a, int nx, int nz) {  // src/Boundary.cpp:21-31
    for (int k = 1; k <= nz - 2; k++) {
        at(a, nz, 0, k) = at(a, nz, nx - 2, k);
        at(a, nz, nx - 1, k) = at(a, nz, 1, k);
    }
    for (int i = 0; i < nx; i++) at(a, nz, i, 0) = at(a, nz, i, 1) = at(a, nz, i, nz - 1) = 0.;
}

// src/Init.cpp:272-277
double get_tb(int k, const vvm_dev_config &c) {
    const double z_top = 12000., T_top = 213., tb_top = 343.;
    const double z_t = c.dz * (k - 0.5);
    if (z_t <= z_top) return 300. + 43. * std::pow(z_t / z_top, 1.25);
    return tb_top * std::exp(c.GRAVITY * (z_t - z_top) / (c.Cp * T_top));
}

// src/Init.cpp:294-299; the reference passes dz as an int
double get_qvb(int k, int dz) {
    const double z_t = (k - 0.5) * dz;
    if (z_t <= 4000) return 0.0161 - 0.000003375 * z_t;
    if (4000 < z_t && z_t <= 8000) return 0.0026 - 0.00000065 * (z_t - 4000);
    return 0.;
}

// src/Init.cpp:279-292 (XRANGE of the reference includes the ghost cells)
double get_th(int i, int k, const vvm_dev_config &c) {
    const double XC = (c.nx * c.dx) / 2., XR = 4000.;
    const double ZC = 2500., ZR = 2000.;
    const double x = (i - 0.5) * c.dx, z = (k - 0.5) * c.dz;
    const double rad = std::sqrt(std::pow((x - XC) / XR, 2) + std::pow((z - ZC) / ZR, 2));
    const double delta = 3.;
    if (rad <= 1) return 0.5 * delta * (std::cos(M_PI * rad) + 1);
    return 0.;
}

void init1d_analytic(const vvm_dev_config &c, vvm_host_view &v) {
    const int nz = c.nz;
    v.thb[1] = 300.;
    for (int k = 2; k <= nz - 2; k++) {
        v.thb[k] = get_tb(k, c);
        v.thb_init[k] = v.thb[k];
    }
    bc1d(v.thb, nz);
    bc1d(v.thb_init, nz);
    for (int k = 1; k <= nz - 2; k++) {
        v.qvb[k] = get_qvb(k, int(c.dz));
        v.thvb[k] = v.thb[k] * (1. + 0.61 * v.qvb[k]);
    }
    bc1d(v.qvb, nz);
    bc1d(v.qvb0, nz);
    bc1d(v.thvb, nz);
    const double pisfc = std::pow((c.PSURF / c.P0), c.Rd / c.Cp);
    for (int k = 1; k <= nz - 2; k++) {
        if (k == 1) {
            v.pib[k] = pisfc - c.GRAVITY * 0.5 * c.dz / (c.Cp * v.thvb[k]);
        } else {
            const double tvbavg = 0.5 * (v.thvb[k] + v.thvb[k - 1]);
            v.pib[k] = v.pib[k - 1] - c.GRAVITY * c.dz / (c.Cp * tvbavg);
        }
    }
    bc1d(v.pib, nz);
    for (int k = 1; k <= nz - 2; k++) v.rhou[k] = c.P0 * std::pow(v.pib[k], c.Cv / c.Rd) / (c.Rd * v.thvb[k]);
    bc1d(v.rhou, nz);
    for (int k = 2; k <= nz - 1; k++) {
        v.thb_zeta[k] = 0.5 * (v.thb[k] + v.thb[k - 1]);
        v.rhow[k] = 0.5 * (v.rhou[k] + v.rhou[k - 1]);
    }
    v.thb_zeta[1] = v.thb_zeta[2] - (v.thb_zeta[3] - v.thb_zeta[2]);
    v.rhow[1] = v.rhow[2] - (v.rhow[3] - v.rhow[2]);
    v.thb_zeta[0] = v.thb_zeta[1];
    v.rhow[0] = v.rhow[1];
    v.rhou[0] = v.rhow[0];
    for (int k = 1; k <= nz - 2; k++) {
        v.pb[k] = c.P0 * std::pow(v.pib[k], c.Cp / c.Rd);
        const double Tbar = v.thb[k] * v.pib[k];
        v.qvsb[k] = (380. / v.pb[k]) * std::exp((17.27 * (Tbar - 273.)) / (Tbar - 36.));
    }
    bc1d(v.pb, nz);
    bc1d(v.qvsb, nz);
    for (int k = 1; k <= nz - 2; k++) v.qvb[k] = get_qvb(k, int(c.dz));
    bc1d(v.qvb, nz);
}

int load_file(const vvm_dev_config &c, const char *path, vvm_host_view &v) {
    const int nz = c.nz;
    std::ifstream in(path);
    if (!in) return VVM_ERR_ARG;
    std::string line;
    std::getline(in, line);  // header
    std::getline(in, line);  // the zero level is skipped
    double ZZ, ZT, RHO, THBAR, PBAR, PIBAR, QVBAR, Q1, Q2, RHOZ;
    int i = 1;
    // the reference reads every row of the file; rows beyond the grid's
    // nz-2 levels would fall outside its arrays, so they are not read here
    while (i <= nz - 2 && (in >> ZZ >> ZT >> RHO >> THBAR >> PBAR >> PIBAR >> QVBAR >> Q1 >> Q2 >> RHOZ)) {
        v.thb[i] = THBAR;
        v.thb_init[i] = v.thb[i];
        v.qvb[i] = QVBAR;
        v.pib[i] = PIBAR;
        v.pb[i] = PBAR;
        v.rhou[i] = RHOZ;
        v.rhow[i] = RHO;
        if (c.tropical_forcing) {
            v.Q1LS[i] = Q1 * 6.;
            v.Q2LS[i] = Q2 * 6.;
        }
        v.thvb[i] = v.thb[i] * (1. + 0.61 * v.qvb[i]);
        v.qvsb[i] = (380. / v.pb[i]) *
                    std::exp((17.27 * (v.thb[i] * v.pib[i] - 273.)) / (v.thb[i] * v.pib[i] - 36.));
        i++;
    }
    if (i <= nz - 2) return VVM_ERR_SIZE;  // fewer levels in the file than the grid has
    double *bcs[] = {v.thb, v.thb_init, v.qvb, v.pib, v.pb, v.rhou, v.rhow, v.thvb, v.qvsb};
    for (double *a : bcs) bc1d(a, nz);
    v.rhow[nz - 1] = v.rhou[nz - 2];
    for (int k = 0; k < nz - 1; k++) v.pib[k] = std::pow(v.pb[k] / 100000., c.Rd / c.Cp);
    for (int k = 1; k <= nz - 2; k++) v.thb_zeta[k] = 0.5 * (v.thb[k] + v.thb[k - 1]);
    v.thb_zeta[1] = v.thb_zeta[2] - (v.thb_zeta[3] - v.thb_zeta[2]);
    bc1d(v.thb_zeta, nz);
    v.thb_zeta[nz - 1] = v.thb[nz - 2];
    return VVM_OK;
}

// the part of Init1d after either sounding (src/Init.cpp:118-134)
void init1d_tail(const vvm_dev_config &c, vvm_host_view &v) {
    const int nz = c.nz;
    for (int k = 0; k < nz; k++) {
        v.thbm[k] = v.thb[k];
        v.qvb0[k] = v.qvb[k];
        v.thvb[k] = v.thvbm[k] = v.thb[k] + 0.61 * v.qvb[k];
    }
    for (int k = 1; k < nz - 1; k++) {
        v.z[k] = (k - 0.5) * c.dz;
        v.z_zeta[k] = (k - 1) * c.dz;
        v.lambda2[k] = 1. / (1. / std::pow(0.23 * std::sqrt(c.dx * c.dz), 2) + 1. / std::pow(0.4 * 0.4 * v.z[k], 2));
    }
    v.z[0] = v.z[1];
    v.z[nz - 1] = v.z[nz - 2];
    v.z_zeta[0] = v.z_zeta[1];
    v.z_zeta[nz - 1] = v.z_zeta[nz - 2];
}

int init2d(const vvm_dev_config &c, vvm_host_view &v) {
    const int nx = c.nx, nz = c.nz;
    for (int k = 0; k <= nz - 1; k++) {
        for (int i = 0; i <= nx - 1; i++) {
            at(v.qv, nz, i, k) = at(v.qvm, nz, i, k) = v.qvb[k];
            at(v.qvp, nz, i, k) = 0.;
            at(v.qc, nz, i, k) = at(v.qcp, nz, i, k) = at(v.qcm, nz, i, k) = 0.;
            at(v.qr, nz, i, k) = at(v.qrp, nz, i, k) = at(v.qrm, nz, i, k) = 0.;
        }
        v.qvb0[k] = v.qvb[k];
    }
    const double rdx = 1. / c.dx, rdz = 1. / c.dz;
    if (c.tropical_forcing) {
        int rc = vvm_random_perturbation(&c, &v, 0, -0.25, 0.25, 1.);
        if (rc) return rc;
        for (int i = 1; i <= nx - 2; i++)
            for (int k = 1; k <= nz - 2; k++) {
                at(v.th, nz, i, k) = v.thb[k] + at(v.init_th_forcing, nz, i, k);
                at(v.thm, nz, i, k) = at(v.th, nz, i, k);
                at(v.qv, nz, i, k) = v.qvb[k];
                at(v.qvm, nz, i, k) = at(v.qv, nz, i, k);
            }
        bc2d_center(v.th, nx, nz);
        bc2d_center(v.thm, nx, nz);
        bc2d_center(v.qv, nx, nz);
        bc2d_center(v.qvm, nx, nz);
    } else {
        // the reference regenerates the same kick (seed 0) at every point
        // (src/Init.cpp:193-196); once gives the same field
        if (c.addforcingtime > 0) {
            int rc = vvm_random_perturbation(&c, &v, 0, -0.25, 0.25, 1.);
            if (rc) return rc;
        }
        for (int i = 1; i <= nx - 2; i++)
            for (int k = 1; k <= nz - 2; k++) {
                double th = at(v.th, nz, i, k);  // other CASE values leave it as allocated
                if (c.CASE == 0) th = v.thb[k];
                else if (c.CASE == 1 || c.CASE == 2) th = v.thb[k] + get_th(i, k, c);
                if (c.addforcingtime > 0) th += at(v.init_th_forcing, nz, i, k);
                at(v.th, nz, i, k) = th;
                at(v.thm, nz, i, k) = th;
                at(v.u, nz, i, k) = 0.;
                at(v.w, nz, i, k) = 0.;
            }
        bc2d_center(v.th, nx, nz);
        bc2d_center(v.thm, nx, nz);
        for (int k = 0; k <= nz - 1; k++) {
            for (int i = 0; i <= nx - 1; i++) at(v.qv, nz, i, k) = at(v.qvm, nz, i, k) = v.qvb[k];
            v.qvb0[k] = v.qvb[k];
        }
        if (c.CASE == 2) {  // shear: -8 m/s at the surface to 0 at 2 km, constant above
            double u_tmp = 0.;
            for (int k = 1; k <= nz - 2; k++) {
                const double z_u = (k - 1) * c.dz;
                if (z_u <= 2000.) u_tmp = -8. + (8. / 2000.0) * z_u;
                for (int i = 1; i <= nx - 2; i++) at(v.u, nz, i, k) = u_tmp;
            }
            bc2d_center(v.u, nx, nz);
        }
    }
    for (int i = 1; i <= nx - 2; i++)
        for (int k = 1; k <= nz - 2; k++) {
            const double pw_px = (at(v.w, nz, i, k) - at(v.w, nz, i - 1, k)) * rdx;
            const double pu_pz = (at(v.u, nz, i, k) - at(v.u, nz, i, k - 1)) * rdz;
            at(v.zeta, nz, i, k) = (pw_px - pu_pz) / v.rhow[k];
            at(v.zetam, nz, i, k) = at(v.zeta, nz, i, k);
        }
    bc2d_westdown(v.zeta, nx, nz);
    bc2d_westdown(v.zetam, nx, nz);
    double ubar = 0.;
    for (int i = 1; i < nx - 1; i++) ubar += at(v.u, nz, i, nz - 2);
    ubar /= double(nx - 2.);
    v.ubarTopm = ubar;
    v.ubarTop = v.ubarTopm;
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            at(v.zetap, nz, i, k) = 0.;
            at(v.thp, nz, i, k) = 0.;
        }
    return VVM_OK;
}

}  // namespace

extern "C" int vvm_init_profiles(const vvm_dev_config *c, const char *sounding_path, vvm_host_view *v) {
    if (!c || !v || c->nx < 4 || c->nz < 5) return VVM_ERR_ARG;
    const size_t f2 = size_t(c->nx) * c->nz;
    // a freshly constructed model: every array zero (src/Declare.hpp:102-398)
    double *fields[] = {v->zeta, v->zetam, v->zetap, v->th, v->thm, v->thp, v->qv, v->qvm, v->qvp, v->qc,
                        v->qcm, v->qcp, v->qr, v->qrm, v->qrp, v->u, v->w, v->RKM, v->RKH, v->init_th_forcing};
    for (double *p : fields) {
        if (!p) return VVM_ERR_ARG;
        std::memset(p, 0, f2 * sizeof(double));
    }
    double *opt2[] = {v->evaporation, v->accretion, v->autoconversion, v->condensation};
    for (double *p : opt2)
        if (p) std::memset(p, 0, f2 * sizeof(double));
    double *hists[] = {v->dth_advect, v->dth_buoyancy, v->dzeta_advect, v->dqv_advect,
                       v->dqc_advect, v->dqr_advect,   v->dqr_VT};
    for (double *p : hists)
        if (p) std::memset(p, 0, 2 * f2 * sizeof(double));
    double *rows[] = {v->precip, v->xi, v->uxi};
    for (double *p : rows)
        if (p) std::memset(p, 0, size_t(c->nx) * sizeof(double));
    double *profs[] = {v->rhou, v->rhow, v->pib, v->pb, v->thb, v->thbm, v->thb_init, v->thb_zeta, v->thvb,
                       v->thvbm, v->qvb, v->qvb0, v->qvsb, v->z, v->z_zeta, v->lambda2, v->Q1LS, v->Q2LS};
    for (double *p : profs) {
        if (!p) return VVM_ERR_ARG;
        std::memset(p, 0, size_t(c->nz) * sizeof(double));
    }
    v->ubarTop = v->ubarTopp = v->ubarTopm = 0.;
    v->dubarTop_advect[0] = v->dubarTop_advect[1] = 0.;
    v->step = 0;
    v->status_for_adding_forcing = 1;
    if (sounding_path) {
        const int rc = load_file(*c, sounding_path, *v);
        if (rc) return rc;
    } else {
        init1d_analytic(*c, *v);
    }
    init1d_tail(*c, *v);
    return VVM_OK;
}

extern "C" int vvm_init_fields(const vvm_dev_config *c, vvm_host_view *v) {
    if (!c || !v || c->nx < 4 || c->nz < 5 || !v->qv || !v->th || !v->u || !v->w || !v->zeta || !v->init_th_forcing)
        return VVM_ERR_ARG;
    return init2d(*c, *v);
}

extern "C" int vvm_init_reference(const vvm_dev_config *c, const char *sounding_path, vvm_host_view *v) {
    const int rc = vvm_init_profiles(c, sounding_path, v);
    return rc ? rc : vvm_init_fields(c, v);
}
