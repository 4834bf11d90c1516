// vvm.cpp -- the C++ state/config API (include/vvm.hpp) over the C ABI of
// include/vvm_dev.h.  Host arrays keep the reference's layout and names
// (src/Declare.hpp); the timestep itself runs on the device.
#include "vvm.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "ReadConfig.hpp"

Config_VVM::Config_VVM(double dt, double dx, double dz, int XRANGE, int ZRANGE, double TIMEEND, int TIMEROUTPUTSIZE,
                       std::string outputpath, int OUTPUTSTEP, double Kx, double Kz, double TIMETS, double tolerance,
                       double GRAVITY, double Cp, double Cv, double Rd, double Lv, double P0, double PSURF,
                       double addforcingtime, int CASE, double mositure_nudge_time, int year, int month, int day,
                       double hour, double minute, double second, double lon, double lat)
    : dt(dt), dx(dx), dz(dz),
      // the reference adds one ghost cell on each side (src/Declare.hpp:33)
      XRANGE(int(XRANGE + 2 * dx)), ZRANGE(int(ZRANGE + 2 * dz)),
      TIMEEND(TIMEEND), TIMEROUTPUTSIZE(TIMEROUTPUTSIZE), outputpath(std::move(outputpath)), OUTPUTSTEP(OUTPUTSTEP),
      Kx(Kx), Kz(Kz), TIMETS(TIMETS), tolerance(tolerance), GRAVITY(GRAVITY), Cp(Cp), Cv(Cv), Rd(Rd), Lv(Lv), P0(P0),
      PSURF(PSURF), addforcingtime(addforcingtime), CASE(CASE), mositure_nudge_time(mositure_nudge_time), year(year),
      month(month), day(day), hour(hour), minute(minute), second(second), lon(lon), lat(lat) {}

namespace {

[[noreturn]] void die(const vvm &m, int rc, const char *what) {
    std::fprintf(stderr, "vvm: %s failed (%d): %s\n", what, rc, m.dev ? vvm_dev_last_error(m.dev) : "no device");
    std::exit(1);
}

void check(const vvm &m, int rc, const char *what) {
    if (rc < 0) die(m, rc, what);
}

void ensure_device(vvm &m) {
    if (!m.dev) vvm::PoissonSolver::InitPoissonMatrix(m);
    m.sync_to_device();
}

// One operator on the device with the host arrays current before and after.
// The host state is uploaded first, so edits made to the host arrays between
// two operator calls are seen, as the reference's operators see them.
void run_op_field(vvm &m, int op, int field, const char *name) {
    if (m.dev) m.dev_stale = true;
    ensure_device(m);
    int rc = vvm_dev_op_field(m.dev, op, field);
    if (rc > 0) std::printf("%s: the elliptic solve did not reach its stop rule\n", name);
    check(m, rc, name);
    m.host_stale = true;
    m.sync_to_host();
}

void run_op(vvm &m, int op, const char *name) { run_op_field(m, op, 0, name); }

// Live models, so that the reference's NegativeValueProcess(double **, nx,
// nz), which takes no model, can find the model (and device) that owns an
// array.
std::mutex g_models_mu;
std::vector<vvm *> g_models;

// VVM_FIELD_AT id of one of a model's fifteen prognostic arrays, or -1.
int prognostic_id(const vvm &m, double **var) {
    double **lev[3][5] = {{m.zetam, m.thm, m.qvm, m.qcm, m.qrm},
                          {m.zeta, m.th, m.qv, m.qc, m.qr},
                          {m.zetap, m.thp, m.qvp, m.qcp, m.qrp}};
    for (int l = 0; l < 3; l++)
        for (int f = 0; f < 5; f++)
            if (lev[l][f] == var) return VVM_FIELD_AT(f, l);
    return -1;
}

// Field id of a p-level array (zetap .. qrp), or -1.
int future_field(const vvm &m, double **var) {
    const int id = prognostic_id(m, var);
    return (id >= 0 && id / 8 == VVM_LEVEL_P) ? id % 8 : -1;
}

[[noreturn]] void not_model_array(const char *what) {
    std::fprintf(stderr, "vvm: %s: the array is not one of the model's prognostic arrays (the operators run on "
                         "the device copy of the model state)\n", what);
    std::exit(1);
}

bool is_output_step(const vvm &m, int step) {
    // src/Iteration.cpp:231
    return step % m.OUTPUTSTEP == 0 || step == m.TIMEEND - 1 || step == m.TIMEEND - 2 || step == 550001;
}

}  // namespace

vvm::vvm(const Config_VVM &c)
    : rdx(1.0 / c.dx), r2dx(rdx / 2.0), rdz(1.0 / c.dz), r2dz(rdz / 2.0), rdx2(rdx * rdx), rdz2(rdz * rdz),
      nx(int(c.XRANGE / c.dx)), nz(int(c.ZRANGE / c.dz)), dt(c.dt), d2t(2.0 * c.dt), dx(c.dx), dz(c.dz),
      XRANGE(c.XRANGE), ZRANGE(c.ZRANGE), TIMEEND(c.TIMEEND), TIMEROUTPUTSIZE(c.TIMEROUTPUTSIZE),
      outputpath(c.outputpath), OUTPUTSTEP(c.OUTPUTSTEP), Kx(c.Kx), Kz(c.Kz), TIMETS(c.TIMETS),
      tolerance(c.tolerance), GRAVITY(c.GRAVITY), Cp(c.Cp), Cv(c.Cv), Rd(c.Rd), Lv(c.Lv), P0(c.P0), PSURF(c.PSURF),
      addforcingtime(c.addforcingtime), CASE(c.CASE), year(c.year), month(c.month), day(c.day), hour(c.hour),
      minute(c.minute), second(c.second), lon(c.lon), lat(c.lat), moisture_nudge_time(c.mositure_nudge_time) {
    dev_config.nx = nx;
    dev_config.nz = nz;
    dev_config.dx = dx;
    dev_config.dz = dz;
    dev_config.dt = dt;
    dev_config.GRAVITY = GRAVITY;
    dev_config.Cp = Cp;
    dev_config.Cv = Cv;
    dev_config.Rd = Rd;
    dev_config.Lv = Lv;
    dev_config.P0 = P0;
    dev_config.PSURF = PSURF;
    dev_config.addforcingtime = addforcingtime;
    dev_config.moisture_nudge_time = moisture_nudge_time;
    dev_config.tolerance = tolerance;
    dev_config.CASE = CASE;
    dev_config.tropical_forcing = c.tropical_forcing;
    dev_config.solver_mode = c.solver_mode;
    dev_config.guess_mode = c.guess_mode;
    dev_config.stop_mode = c.stop_mode;
    dev_config.max_iter = c.max_iter;
    dev_config.tol_rel = c.tol_rel;
    gpu_id = c.gpu_id;
    allocateMemory();
    std::lock_guard<std::mutex> lk(g_models_mu);
    g_models.push_back(this);
}

vvm::~vvm() {
    {
        std::lock_guard<std::mutex> lk(g_models_mu);
        g_models.erase(std::remove(g_models.begin(), g_models.end(), this), g_models.end());
    }
    if (dev) vvm_dev_destroy(dev);
    dev = nullptr;
    deallocateMemory();
}

double **vvm::allocate2DContinuousArray(int rows, int cols, double *&contMemory) {
    double **array = new double *[rows]();
    contMemory = new double[size_t(rows) * cols]();
    for (int i = 0; i < rows; ++i) array[i] = &contMemory[size_t(i) * cols];
    return array;
}

void vvm::deallocate2DContinuousArray(double **array, double *contMemory) {
    if (array != nullptr) {
        delete[] contMemory;
        delete[] array;
    }
}

double ***vvm::allocate3DContinuousArray(int dim1, int dim2, int dim3, double *&contMemory) {
    double ***array = new double **[dim1]();
    contMemory = new double[size_t(dim1) * dim2 * dim3]();
    for (int i = 0; i < dim1; ++i) {
        array[i] = new double *[dim2]();
        for (int j = 0; j < dim2; ++j) array[i][j] = &contMemory[(size_t(i) * dim2 + j) * dim3];
    }
    return array;
}

void vvm::deallocate3DContinuousArray(double ***array, double *contMemory) {
    if (array != nullptr) {
        delete[] contMemory;
        for (int i = 0; i < nx; ++i) delete[] array[i];
        delete[] array;
    }
}

void vvm::allocateMemory() {
    double **profiles[] = {&thb, &thb_init, &thbm, &thb_zeta, &rhou, &rhow, &pib, &qvb, &qvb0,
                           &qvsb, &pb, &thvb, &thvbm, &z, &z_zeta, &lambda2, &Q1LS, &Q2LS};
    for (double **p : profiles) *p = new double[nz]();
    xi = new double[nx]();
    uxi = new double[nx]();
    precip = new double[nx]();
    std::pair<double ***, double **> f2[] = {
        {&zetap, &zetapcont}, {&zeta, &zetacont}, {&zetam, &zetamcont}, {&thp, &thpcont}, {&th, &thcont},
        {&thm, &thmcont}, {&qvp, &qvpcont}, {&qv, &qvcont}, {&qvm, &qvmcont}, {&qcp, &qcpcont},
        {&qc, &qccont}, {&qcm, &qcmcont}, {&qrp, &qrpcont}, {&qr, &qrcont}, {&qrm, &qrmcont},
        {&u, &ucont}, {&w, &wcont}, {&RKM, &RKMcont}, {&RKH, &RKHcont}, {&evaporation, &evaporationcont},
        {&accretion, &accretioncont}, {&autoconversion, &autoconversioncont},
        {&condensation, &condensationcont}, {&init_th_forcing, &init_th_forcingcont},
        {&radiation_heating_rate, &radiation_heating_ratecont}};
    for (auto &f : f2) *f.first = allocate2DContinuousArray(nx, nz, *f.second);
    std::pair<double ****, double **> f3[] = {
        {&dth_advect, &dth_advectcont}, {&dth_buoyancy, &dth_buoyancycont}, {&dzeta_advect, &dzeta_advectcont},
        {&dqv_advect, &dqv_advectcont}, {&dqc_advect, &dqc_advectcont}, {&dqr_advect, &dqr_advectcont},
        {&dqr_VT, &dqr_VTcont}};
    for (auto &f : f3) *f.first = allocate3DContinuousArray(nx, nz, 2, *f.second);
    double **timers[] = {&t_advection, &t_poisson, &t_diffusion, &t_microphysics, &t_all};
    for (double **t : timers) *t = new double[TIMEROUTPUTSIZE > 0 ? TIMEROUTPUTSIZE : 1]();
}

void vvm::deallocateMemory() {
    double **profiles[] = {&thb, &thb_init, &thbm, &thb_zeta, &rhou, &rhow, &pib, &qvb, &qvb0, &qvsb, &pb,
                           &thvb, &thvbm, &z, &z_zeta, &lambda2, &Q1LS, &Q2LS, &xi, &uxi, &precip,
                           &t_advection, &t_poisson, &t_diffusion, &t_microphysics, &t_all};
    for (double **p : profiles) {
        delete[] *p;
        *p = nullptr;
    }
    std::pair<double ***, double **> f2[] = {
        {&zetap, &zetapcont}, {&zeta, &zetacont}, {&zetam, &zetamcont}, {&thp, &thpcont}, {&th, &thcont},
        {&thm, &thmcont}, {&qvp, &qvpcont}, {&qv, &qvcont}, {&qvm, &qvmcont}, {&qcp, &qcpcont},
        {&qc, &qccont}, {&qcm, &qcmcont}, {&qrp, &qrpcont}, {&qr, &qrcont}, {&qrm, &qrmcont},
        {&u, &ucont}, {&w, &wcont}, {&RKM, &RKMcont}, {&RKH, &RKHcont}, {&evaporation, &evaporationcont},
        {&accretion, &accretioncont}, {&autoconversion, &autoconversioncont},
        {&condensation, &condensationcont}, {&init_th_forcing, &init_th_forcingcont},
        {&radiation_heating_rate, &radiation_heating_ratecont}};
    for (auto &f : f2) {
        deallocate2DContinuousArray(*f.first, *f.second);
        *f.first = nullptr;
        *f.second = nullptr;
    }
    std::pair<double ****, double **> f3[] = {
        {&dth_advect, &dth_advectcont}, {&dth_buoyancy, &dth_buoyancycont}, {&dzeta_advect, &dzeta_advectcont},
        {&dqv_advect, &dqv_advectcont}, {&dqc_advect, &dqc_advectcont}, {&dqr_advect, &dqr_advectcont},
        {&dqr_VT, &dqr_VTcont}};
    for (auto &f : f3) {
        deallocate3DContinuousArray(*f.first, *f.second);
        *f.first = nullptr;
        *f.second = nullptr;
    }
}

vvm_host_view vvm::host_view() {
    vvm_host_view v{};
    v.zeta = zetacont, v.zetam = zetamcont, v.zetap = zetapcont;
    v.th = thcont, v.thm = thmcont, v.thp = thpcont;
    v.qv = qvcont, v.qvm = qvmcont, v.qvp = qvpcont;
    v.qc = qccont, v.qcm = qcmcont, v.qcp = qcpcont;
    v.qr = qrcont, v.qrm = qrmcont, v.qrp = qrpcont;
    v.u = ucont, v.w = wcont, v.RKM = RKMcont, v.RKH = RKHcont, v.init_th_forcing = init_th_forcingcont;
    v.evaporation = evaporationcont, v.accretion = accretioncont, v.autoconversion = autoconversioncont;
    v.condensation = condensationcont;
    v.dth_advect = dth_advectcont, v.dth_buoyancy = dth_buoyancycont, v.dzeta_advect = dzeta_advectcont;
    v.dqv_advect = dqv_advectcont, v.dqc_advect = dqc_advectcont, v.dqr_advect = dqr_advectcont;
    v.dqr_VT = dqr_VTcont;
    v.precip = precip, v.xi = xi, v.uxi = uxi;
    v.rhou = rhou, v.rhow = rhow, v.pib = pib, v.pb = pb, v.thb = thb, v.thbm = thbm, v.thb_init = thb_init;
    v.thb_zeta = thb_zeta, v.thvb = thvb, v.thvbm = thvbm, v.qvb = qvb, v.qvb0 = qvb0, v.qvsb = qvsb;
    v.z = z, v.z_zeta = z_zeta, v.lambda2 = lambda2, v.Q1LS = Q1LS, v.Q2LS = Q2LS;
    v.ubarTop = ubarTop, v.ubarTopp = ubarTopp, v.ubarTopm = ubarTopm;
    v.dubarTop_advect[0] = dubarTop_advect[0], v.dubarTop_advect[1] = dubarTop_advect[1];
    v.step = step;
    v.status_for_adding_forcing = status_for_adding_forcing ? 1 : 0;
    return v;
}

void vvm::sync_to_device() {
    if (!dev_stale) return;
    if (host_stale) {
        std::fprintf(stderr, "vvm: host and device copies both modified\n");
        std::exit(1);
    }
    vvm_host_view v = host_view();
    check(*this, vvm_dev_upload(dev, &v), "vvm_dev_upload");
    dev_stale = false;
}

void vvm::sync_to_host() {
    if (!host_stale) return;
    vvm_host_view v = host_view();
    check(*this, vvm_dev_download(dev, &v), "vvm_dev_download");
    ubarTop = v.ubarTop, ubarTopp = v.ubarTopp, ubarTopm = v.ubarTopm;
    dubarTop_advect[0] = v.dubarTop_advect[0], dubarTop_advect[1] = v.dubarTop_advect[1];
    step = v.step;
    status_for_adding_forcing = v.status_for_adding_forcing != 0;
    host_stale = false;
}

// src/Boundary.cpp:3-31
void vvm::BoundaryProcess1D_center(double *var, int nz) {
    var[0] = var[1];
    var[nz - 1] = var[nz - 2];
}

void vvm::BoundaryProcess2D_center(double **var, int nx, int nz) {
    for (int k = 1; k <= nz - 2; k++) {
        var[0][k] = var[nx - 2][k];
        var[nx - 1][k] = var[1][k];
    }
    for (int i = 0; i < nx; i++) {
        var[i][0] = var[i][1];
        var[i][nz - 1] = var[i][nz - 2];
    }
}

void vvm::BoundaryProcess2D_westdown(double **var, int nx, int nz) {
    for (int k = 1; k <= nz - 2; k++) {
        var[0][k] = var[nx - 2][k];
        var[nx - 1][k] = var[1][k];
    }
    for (int i = 0; i < nx; i++) var[i][0] = var[i][1] = var[i][nz - 1] = 0.;
}

void vvm::BoundaryProcess2D_all(vvm &model) { run_op(model, VVM_OP_BC_ALL, "BoundaryProcess2D_all"); }

void vvm::AddForcing(vvm &model) {
    // the heating rate the reference's AddForcing reads (AddForcing.cpp:20)
    ensure_device(model);
    check(model, vvm_dev_set_radiation(model.dev, model.radiation_heating_ratecont, model.radiation_cooling ? 1 : 0),
          "vvm_dev_set_radiation");
    run_op(model, VVM_OP_ADD_FORCING, "AddForcing");
}

void vvm::PoissonSolver::InitPoissonMatrix(vvm &model) {
    if (model.dev) return;
    int rc = vvm_dev_create(&model.dev_config, model.gpu_id, &model.dev);
    if (rc < 0) {
        std::fprintf(stderr, "vvm: vvm_dev_create failed (%d)%s%s\n", rc, model.dev ? ": " : "",
                     model.dev ? vvm_dev_last_error(model.dev) : "");
        std::exit(1);
    }
    model.dev_stale = true;
    model.host_stale = false;
}

void vvm::PoissonSolver::cal_w(vvm &model, int, int, int) { run_op(model, VVM_OP_CAL_W, "cal_w"); }
void vvm::PoissonSolver::cal_u(vvm &model) { run_op(model, VVM_OP_CAL_U, "cal_u"); }
void vvm::PoissonSolver::pubarTop_pt(vvm &model) { run_op(model, VVM_OP_PUBARTOP, "pubarTop_pt"); }
void vvm::Turbulence::RKM_RKH(vvm &model) { run_op(model, VVM_OP_TURBULENCE, "RKM_RKH"); }
void vvm::NumericalProcess::Nudge(vvm &model) { run_op(model, VVM_OP_NUDGE, "Nudge"); }
void vvm::MicroPhysics::Kessler(vvm &model) { run_op(model, VVM_OP_MICROPHYSICS, "Kessler"); }
void vvm::Iteration::pzeta_pt(vvm &model) { run_op(model, VVM_OP_PZETA, "pzeta_pt"); }
void vvm::Iteration::pthermo_pt(vvm &model) { run_op(model, VVM_OP_PSCALARS, "pthermo_pt"); }
void vvm::Iteration::nextTimeStep(vvm &model) { run_op(model, VVM_OP_NEXT, "nextTimeStep"); }
void vvm::Iteration::pth_pt(vvm &model) { run_op(model, VVM_OP_PTH, "pth_pt"); }
void vvm::Iteration::pqv_pt(vvm &model) { run_op(model, VVM_OP_PQV, "pqv_pt"); }
void vvm::Iteration::pqc_pt(vvm &model) { run_op(model, VVM_OP_PQC, "pqc_pt"); }
void vvm::Iteration::pqr_pt(vvm &model) { run_op(model, VVM_OP_PQR, "pqr_pt"); }
void vvm::Iteration::updateMean(vvm &model) { run_op(model, VVM_OP_UPDATE_MEAN, "updateMean"); }

void vvm::Advection_zeta(vvm &model) { run_op(model, VVM_OP_ADVECTION_ZETA, "Advection_zeta"); }
void vvm::Bouyancy(vvm &model) { run_op(model, VVM_OP_BOUYANCY, "Bouyancy"); }
void vvm::Advection_qrVT(vvm &model) { run_op(model, VVM_OP_ADVECTION_QRVT, "Advection_qrVT"); }
void vvm::Advection_thermo(double **past, double **now, double **future, double ***dvar, vvm &model) {
    // the reference's four calls (src/Iteration.cpp:37,91,101,110)
    const int f = future_field(model, future);
    double ***hist[5] = {nullptr, model.dth_advect, model.dqv_advect, model.dqc_advect, model.dqr_advect};
    double **pasts[5] = {nullptr, model.thm, model.qvm, model.qcm, model.qrm};
    double **nows[5] = {nullptr, model.th, model.qv, model.qc, model.qr};
    if (f < VVM_FIELD_TH || past != pasts[f] || now != nows[f] || dvar != hist[f]) not_model_array("Advection_thermo");
    run_op_field(model, VVM_OP_ADVECTION_THERMO, f, "Advection_thermo");
}

void vvm::Turbulence::Mparam(vvm &model, double **, double **var_future) {
    if (var_future != model.zetap) not_model_array("Mparam");
    run_op(model, VVM_OP_MPARAM, "Mparam");
}
void vvm::Turbulence::Hparam(vvm &model, double **, double **var_future) {
    const int f = future_field(model, var_future);
    if (f < VVM_FIELD_TH) not_model_array("Hparam");
    run_op_field(model, VVM_OP_HPARAM, f, "Hparam");
}
void vvm::NumericalProcess::Nudge_theta(vvm &model) { run_op(model, VVM_OP_NUDGE_THETA, "Nudge_theta"); }
void vvm::NumericalProcess::Nudge_zeta(vvm &model) { run_op(model, VVM_OP_NUDGE_ZETA, "Nudge_zeta"); }
void vvm::NumericalProcess::Nudge_qv(vvm &model) { run_op(model, VVM_OP_NUDGE_QV, "Nudge_qv"); }
void vvm::MicroPhysics::condensation(vvm &model) { run_op(model, VVM_OP_CONDENSATION, "condensation"); }
void vvm::MicroPhysics::autoconversion(vvm &model) { run_op(model, VVM_OP_AUTOCONVERSION, "autoconversion"); }
void vvm::MicroPhysics::accretion(vvm &model) { run_op(model, VVM_OP_ACCRETION, "accretion"); }
void vvm::MicroPhysics::evaporation(vvm &model) { run_op(model, VVM_OP_EVAPORATION, "evaporation"); }
void vvm::MicroPhysics::NegativeValueProcess(double **var, int nx, int nz) {
    vvm *owner = nullptr;
    int id = -1;
    {
        std::lock_guard<std::mutex> lk(g_models_mu);
        for (vvm *m : g_models)
            if ((id = prognostic_id(*m, var)) >= 0) {
                owner = m;
                break;
            }
    }
    if (!owner || owner->nx != nx || owner->nz != nz) not_model_array("NegativeValueProcess");
    run_op_field(*owner, VVM_OP_NEGATIVE_VALUE, id, "NegativeValueProcess");
}

void vvm::Init::Init1d(vvm &model) {
    model.sync_to_host();
    vvm_host_view v = model.host_view();
    const int rc = vvm_init_profiles(&model.dev_config, model.sounding_path.empty() ? nullptr : model.sounding_path.c_str(), &v);
    if (rc < 0) {
        std::fprintf(stderr, "vvm: Init1d failed (%d)%s%s\n", rc, model.sounding_path.empty() ? "" : " reading ",
                     model.sounding_path.c_str());
        std::exit(1);
    }
    model.ubarTop = v.ubarTop, model.ubarTopp = v.ubarTopp, model.ubarTopm = v.ubarTopm;
    model.step = v.step;
    model.dev_stale = true;
}

void vvm::Init::Init2d(vvm &model) {
    model.sync_to_host();
    vvm_host_view v = model.host_view();
    const int rc = vvm_init_fields(&model.dev_config, &v);
    if (rc < 0) {
        std::fprintf(stderr, "vvm: Init2d failed (%d)\n", rc);
        std::exit(1);
    }
    model.ubarTop = v.ubarTop, model.ubarTopm = v.ubarTopm;
    model.dev_stale = true;
}

void vvm::Init::LoadFromPreviousFile(vvm &model) {
    model.sync_to_host();
    vvm_host_view v = model.host_view();
    const int rc = vvm_init_from_previous(&model.dev_config, model.sounding_path.c_str(),
                                          model.restart_prev_nc.c_str(), model.restart_now_nc.c_str(), &v);
    if (rc < 0) {
        std::fprintf(stderr, "vvm: LoadFromPreviousFile failed (%d): %s, %s, %s\n", rc, model.sounding_path.c_str(),
                     model.restart_prev_nc.c_str(), model.restart_now_nc.c_str());
        std::exit(1);
    }
    model.ubarTopm = v.ubarTopm;
    model.dev_stale = true;
}

void vvm::Init::InitCase(vvm &model, const vvm_case_params &p) {
    model.sync_to_host();
    vvm_host_view v = model.host_view();
    int rc = vvm_init_case(&model.dev_config, &p, &v);
    if (rc < 0) {
        std::fprintf(stderr, "vvm: vvm_init_case failed (%d)\n", rc);
        std::exit(1);
    }
    model.ubarTop = v.ubarTop, model.ubarTopp = v.ubarTopp, model.ubarTopm = v.ubarTopm;
    model.step = v.step;
    model.dev_stale = true;
}

void vvm::Init::RandomPerturbation(vvm &model, int t, double min_range, double max_range,
                                   double standard_deviation) {
    model.sync_to_host();
    vvm_host_view v = model.host_view();
    int rc = vvm_random_perturbation(&model.dev_config, &v, t, min_range, max_range, standard_deviation);
    if (rc < 0) {
        std::fprintf(stderr, "vvm: vvm_random_perturbation failed (%d)\n", rc);
        std::exit(1);
    }
    model.dev_stale = true;
}

namespace {

// netCDF classic (CDF-1) writer: big-endian header of dimensions and
// variables without attributes, then each variable's doubles at its offset.
// A dimension of length 0 is the record (unlimited) dimension: variables
// whose first dimension it is are stored after the fixed ones, one slab
// per record, records interleaved.
struct NcVar {
    std::string name;
    std::vector<int> dims;  // dimension ids
    const double *data;
    size_t count;           // values per record for a record variable
};

void put_be32(std::vector<unsigned char> &b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back(static_cast<unsigned char>(v >> s));
}

void put_name(std::vector<unsigned char> &b, const std::string &s) {
    put_be32(b, uint32_t(s.size()));
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
}

void put_be64_doubles(std::vector<unsigned char> &buf, const double *p, size_t n) {
    buf.resize(n * 8);
    for (size_t i = 0; i < n; i++) {
        uint64_t u;
        std::memcpy(&u, &p[i], 8);
        for (int b = 0; b < 8; b++) buf[i * 8 + b] = static_cast<unsigned char>(u >> (56 - 8 * b));
    }
}

bool is_record_var(const std::vector<std::pair<std::string, int>> &dims, const NcVar &v) {
    return !v.dims.empty() && dims[size_t(v.dims[0])].second == 0;
}

// A new file holding the fixed variables and, when there are record
// variables, their first record.
bool write_nc_classic(const std::string &path, const std::vector<std::pair<std::string, int>> &dims,
                      const std::vector<NcVar> &vars) {
    const uint32_t NC_DIMENSION = 10, NC_VARIABLE = 11, NC_DOUBLE = 6;
    bool has_rec = false;
    for (auto &v : vars) has_rec = has_rec || is_record_var(dims, v);
    // header size first (offsets depend on it)
    auto header = [&](const std::vector<uint32_t> &begins) {
        std::vector<unsigned char> h = {'C', 'D', 'F', 1};
        put_be32(h, has_rec ? 1 : 0);  // numrecs
        put_be32(h, NC_DIMENSION);
        put_be32(h, uint32_t(dims.size()));
        for (auto &d : dims) {
            put_name(h, d.first);
            put_be32(h, uint32_t(d.second));
        }
        put_be32(h, 0), put_be32(h, 0);  // no global attributes
        put_be32(h, NC_VARIABLE);
        put_be32(h, uint32_t(vars.size()));
        for (size_t v = 0; v < vars.size(); v++) {
            put_name(h, vars[v].name);
            put_be32(h, uint32_t(vars[v].dims.size()));
            for (int id : vars[v].dims) put_be32(h, uint32_t(id));
            put_be32(h, 0), put_be32(h, 0);  // no attributes
            put_be32(h, NC_DOUBLE);
            put_be32(h, uint32_t(vars[v].count * 8));
            put_be32(h, begins.empty() ? 0u : begins[v]);
        }
        return h;
    };
    const size_t hsize = header({}).size();
    std::vector<uint32_t> begins(vars.size());
    size_t off = hsize;
    for (int pass = 0; pass < 2; pass++)  // fixed variables, then the record
        for (size_t v = 0; v < vars.size(); v++) {
            if (is_record_var(dims, vars[v]) != (pass == 1)) continue;
            if (off > 0xffffffffull) return false;  // CDF-1 offsets are 32-bit
            begins[v] = uint32_t(off);
            off += vars[v].count * 8;
        }
    std::vector<unsigned char> h = header(begins);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    std::vector<unsigned char> buf;
    for (int pass = 0; pass < 2; pass++)
        for (auto &v : vars) {
            if (is_record_var(dims, v) != (pass == 1)) continue;
            put_be64_doubles(buf, v.data, v.count);
            ok = ok && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
        }
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

// What the appender needs of a classic (CDF-1 / CDF-2) header written by
// any netCDF library: the record count, and per variable its type, slab
// size, offset and whether it is a record variable.
struct NcHeaderVar {
    std::string name;
    bool record = false;
    uint32_t type = 0;
    uint64_t vsize = 0, begin = 0;
};

struct NcHeader {
    int version = 0;
    uint32_t numrecs = 0;
    std::vector<NcHeaderVar> vars;
};

bool read_nc_header(FILE *f, NcHeader &h) {
    auto u32 = [&](uint32_t &v) {
        unsigned char b[4];
        if (std::fread(b, 1, 4, f) != 4) return false;
        v = (uint32_t(b[0]) << 24) | (uint32_t(b[1]) << 16) | (uint32_t(b[2]) << 8) | b[3];
        return true;
    };
    auto u64 = [&](uint64_t &v) {
        uint32_t hi, lo;
        if (!u32(hi) || !u32(lo)) return false;
        v = (uint64_t(hi) << 32) | lo;
        return true;
    };
    // n bytes of data padded to a multiple of 4
    auto skip = [&](uint64_t n) { return std::fseek(f, long((n + 3) & ~uint64_t(3)), SEEK_CUR) == 0; };
    auto name = [&](std::string &s) {
        uint32_t n;
        if (!u32(n) || n > 4096) return false;
        s.assign(n, '\0');
        if (n && std::fread(&s[0], 1, n, f) != n) return false;
        return std::fseek(f, long((4 - n % 4) % 4), SEEK_CUR) == 0;
    };
    static const int type_size[7] = {0, 1, 1, 2, 4, 4, 8};
    auto atts = [&]() {
        uint32_t tag, n;
        if (!u32(tag) || !u32(n)) return false;
        if (tag == 0 && n == 0) return true;
        if (tag != 12) return false;  // NC_ATTRIBUTE
        for (uint32_t a = 0; a < n; a++) {
            std::string s;
            uint32_t type, nelems;
            if (!name(s) || !u32(type) || !u32(nelems) || type < 1 || type > 6) return false;
            if (!skip(uint64_t(nelems) * type_size[type])) return false;
        }
        return true;
    };
    unsigned char magic[4];
    if (std::fread(magic, 1, 4, f) != 4 || magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F') return false;
    h.version = magic[3];
    if (h.version != 1 && h.version != 2) return false;
    uint32_t tag, n;
    if (!u32(h.numrecs) || !u32(tag) || !u32(n)) return false;
    std::vector<uint32_t> dimlen;
    if (tag == 10) {
        for (uint32_t d = 0; d < n; d++) {
            std::string s;
            uint32_t len;
            if (!name(s) || !u32(len)) return false;
            dimlen.push_back(len);
        }
    } else if (tag != 0 || n != 0) {
        return false;
    }
    if (!atts()) return false;
    if (!u32(tag) || !u32(n)) return false;
    if (tag != 11 && !(tag == 0 && n == 0)) return false;
    for (uint32_t v = 0; tag == 11 && v < n; v++) {
        NcHeaderVar var;
        uint32_t nd;
        if (!name(var.name) || !u32(nd) || nd > 64) return false;
        for (uint32_t d = 0; d < nd; d++) {
            uint32_t id;
            if (!u32(id) || id >= dimlen.size()) return false;
            if (d == 0 && dimlen[id] == 0) var.record = true;
        }
        uint32_t vs;
        if (!atts() || !u32(var.type) || !u32(vs)) return false;
        var.vsize = vs;
        if (h.version == 2) {
            if (!u64(var.begin)) return false;
        } else {
            uint32_t b;
            if (!u32(b)) return false;
            var.begin = b;
        }
        h.vars.push_back(var);
    }
    return true;
}

// One more record of every record variable in `vars` (all of them must be
// record variables of type double with the file's slab size), then the
// record count.
bool append_nc_record(const std::string &path, const std::vector<NcVar> &vars) {
    FILE *f = std::fopen(path.c_str(), "r+b");
    if (!f) return false;
    NcHeader h;
    bool ok = read_nc_header(f, h);
    uint64_t recsize = 0;
    int nrec = 0;
    for (auto &v : h.vars)
        if (v.record) recsize += v.vsize, nrec++;
    if (nrec == 1)  // a lone record variable is not padded
        for (auto &v : h.vars)
            if (v.record) recsize = v.vsize;
    std::vector<unsigned char> buf;
    for (auto &w : vars) {
        if (!ok) break;
        const NcHeaderVar *hv = nullptr;
        for (auto &v : h.vars)
            if (v.name == w.name) hv = &v;
        ok = hv && hv->record && hv->type == 6 && hv->vsize == w.count * 8;
        if (!ok) break;
        put_be64_doubles(buf, w.data, w.count);
        const uint64_t at = hv->begin + uint64_t(h.numrecs) * recsize;
        ok = std::fseek(f, long(at), SEEK_SET) == 0 && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    if (ok) {
        std::vector<unsigned char> nr;
        put_be32(nr, h.numrecs + 1);
        ok = std::fseek(f, 4, SEEK_SET) == 0 && std::fwrite(nr.data(), 1, 4, f) == 4;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

}  // namespace

namespace {

// src/Output.cpp:106-285 (OUTPUTNCSAMEFILE) with Kessler: file number
// n / 1200000, record n * dt; a file that exists gets one more record.
void write_same_file_record(int n, vvm &model) {
    const size_t f2 = size_t(model.nx) * model.nz;
    const std::vector<std::pair<std::string, int>> dims = {{"t", 0}, {"x", model.nx}, {"z", model.nz}};
    const std::vector<int> txz = {0, 1, 2}, tx = {0, 1}, t = {0};
    const double tval = n * model.dt;
    std::vector<NcVar> vars = {
        {"t", t, &tval, 1},
        {"th", txz, model.thcont, f2},
        {"zeta", txz, model.zetacont, f2},
        {"u", txz, model.ucont, f2},
        {"w", txz, model.wcont, f2},
        {"ubarTop", t, &model.ubarTopp, 1},
        {"qv", txz, model.qvcont, f2},
        {"qc", txz, model.qccont, f2},
        {"qr", txz, model.qrcont, f2},
        {"precip", tx, model.precip, size_t(model.nx)},
    };
    if (model.output_microphysics) {
        vars.push_back({"accretion", txz, model.accretioncont, f2});
        vars.push_back({"autoconversion", txz, model.autoconversioncont, f2});
        vars.push_back({"evaporation", txz, model.evaporationcont, f2});
        vars.push_back({"condensation", txz, model.condensationcont, f2});
    }
    const std::string path = model.outputpath + "/nc/" + std::to_string(n / 1200000) + ".nc";
    struct stat sb;
    const bool exists = stat(path.c_str(), &sb) == 0;
    if (!(exists ? append_nc_record(path, vars) : write_nc_classic(path, dims, vars))) {
        std::fprintf(stderr, "vvm: cannot %s %s\n", exists ? "append a record to" : "write", path.c_str());
        std::exit(1);
    }
}

}  // namespace

void vvm::Output::output_nc(int n, vvm &model) {
    model.sync_to_host();
    if (model.output_nc_same_file) {
        write_same_file_record(n, model);
        return;
    }
    const size_t f2 = size_t(model.nx) * model.nz;
    const std::vector<std::pair<std::string, int>> dims = {{"x", model.nx}, {"z", model.nz}};
    const std::vector<int> xz = {0, 1}, x = {0};
    // the reference stores ubarTopp in "ubarTop" (src/Output.cpp:87)
    const std::vector<NcVar> vars = {
        {"th", xz, model.thcont, f2},
        {"zeta", xz, model.zetacont, f2},
        {"u", xz, model.ucont, f2},
        {"w", xz, model.wcont, f2},
        {"ubarTop", {}, &model.ubarTopp, 1},
        {"qv", xz, model.qvcont, f2},
        {"qc", xz, model.qccont, f2},
        {"qr", xz, model.qrcont, f2},
        {"accretion", xz, model.accretioncont, f2},
        {"autoconversion", xz, model.autoconversioncont, f2},
        {"evaporation", xz, model.evaporationcont, f2},
        {"condensation", xz, model.condensationcont, f2},
        {"precip", x, model.precip, size_t(model.nx)},
    };
    const std::string path = model.outputpath + "/nc/" + std::to_string(n) + ".nc";
    if (!write_nc_classic(path, dims, vars)) {
        std::fprintf(stderr, "vvm: cannot write %s\n", path.c_str());
        std::exit(1);
    }
}

// src/Output.cpp:288-318: the five phase timers as double [x] variables
// advection, poisson, diffusion, microphysics, all.  The reference sizes
// the dimension nx while the arrays hold TIMEROUTPUTSIZE values; here the
// dimension is the arrays' length, so nothing is read past them.
void vvm::Output::output_time_nc(int n, vvm &model) {
    const int T = model.TIMEROUTPUTSIZE > 0 ? model.TIMEROUTPUTSIZE : 1;
    const std::vector<std::pair<std::string, int>> dims = {{"x", T}};
    const std::vector<int> x = {0};
    const std::vector<NcVar> vars = {
        {"advection", x, model.t_advection, size_t(T)},
        {"poisson", x, model.t_poisson, size_t(T)},
        {"diffusion", x, model.t_diffusion, size_t(T)},
        {"microphysics", x, model.t_microphysics, size_t(T)},
        {"all", x, model.t_all, size_t(T)},
    };
    const std::string path = model.outputpath + "/timer/" + std::to_string(n) + ".nc";
    if (!write_nc_classic(path, dims, vars)) {
        std::fprintf(stderr, "vvm: cannot write %s\n", path.c_str());
        std::exit(1);
    }
}

void vvm::Output::create_directory(std::string path) {
    // every prefix of the path, as mkdir -p
    bool ok = !path.empty();
    for (size_t pos = 1; ok && pos <= path.size(); pos++) {
        if (pos < path.size() && path[pos] != '/') continue;
        const std::string prefix = path.substr(0, pos);
        if (mkdir(prefix.c_str(), 0755) != 0 && errno != EEXIST) ok = false;
    }
    if (!ok) std::cout << "Error on creating directory!\n" << std::endl;
}

void vvm::Output::printInit(vvm &model) {
    std::cout << "z             Pbar          thb        thb_zeta     rhou       rhow       qvb   \t RH      pib"
              << std::endl;
    for (int k = 0; k <= model.nz - 1; k++) {
        const double z = double(k - 0.5) * model.dz;
        std::cout << std::fixed << std::setprecision(4) << z << "    " << model.pb[k] << "     " << model.thb[k] << "    "
                  << model.thb_zeta[k] << "    " << model.rhou[k] << "     " << model.rhow[k] << "    "
                  << model.qvb[k] * 1000 << "    " << model.qvb[k] / model.qvsb[k] << "    " << model.pib[k] << std::endl;
    }
    std::ofstream out(model.outputpath + "init.txt");
    for (int k = 1; k <= model.nz - 2; k++) {
        const double z = double(k - 0.5) * model.dz;
        out << z << "    " << model.thb[k] << "    " << model.rhou[k] << "     " << model.rhow[k] << "   \t "
            << model.qvb[k] << "    " << model.qvsb[k] << "    " << model.qvb[k] / model.qvsb[k] << "    "
            << model.pib[k] << "    " << model.pb[k] << std::endl;
    }
}

void vvm::Output::create_all_directory(vvm &model) {
    if (model.outputpath.empty()) return;
    std::string dirs[] = {model.outputpath, model.outputpath + "/state", model.outputpath + "/nc",
                          model.outputpath + "/timer"};
    for (const auto &d : dirs)
        if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) {
            std::fprintf(stderr, "vvm: cannot create %s\n", d.c_str());
            std::exit(1);
        }
}

// VVMSTATE record file: "VVMSTATE", int32 version 1, int32 record count,
// then per record a 32-byte name, int32 ndim (3), int32 dims[3] and the
// little-endian doubles ([i][k] arrays keep k fastest).
void vvm::Output::output_state(int step, vvm &model) {
    write_state(model.outputpath + "/state/state_" + std::to_string(step) + ".bin", model);
}

void vvm::Output::write_state(const std::string &path, vvm &model) {
    model.sync_to_host();
    std::vector<std::pair<std::string, std::vector<double>>> recs;
    auto one = [&](const char *n, double x) { recs.push_back({n, {x}}); };
    one("cfg.nx", model.nx), one("cfg.nz", model.nz), one("cfg.dx", model.dx), one("cfg.dz", model.dz);
    one("cfg.dt", model.dt), one("cfg.GRAVITY", model.GRAVITY), one("cfg.Cp", model.Cp), one("cfg.Cv", model.Cv);
    one("cfg.Rd", model.Rd), one("cfg.Lv", model.Lv), one("cfg.P0", model.P0), one("cfg.PSURF", model.PSURF);
    one("cfg.addforcingtime", model.addforcingtime), one("cfg.moisture_nudge_time", model.moisture_nudge_time);
    one("cfg.tolerance", model.tolerance), one("cfg.CASE", model.CASE);
    struct A {
        const char *n;
        const double *p;
        int d0, d1, d2;
    };
    const int nx = model.nx, nz = model.nz;
    const A arrs[] = {
        {"zeta", model.zetacont, nx, nz, 1}, {"zetam", model.zetamcont, nx, nz, 1},
        {"zetap", model.zetapcont, nx, nz, 1}, {"th", model.thcont, nx, nz, 1}, {"thm", model.thmcont, nx, nz, 1},
        {"thp", model.thpcont, nx, nz, 1}, {"qv", model.qvcont, nx, nz, 1}, {"qvm", model.qvmcont, nx, nz, 1},
        {"qvp", model.qvpcont, nx, nz, 1}, {"qc", model.qccont, nx, nz, 1}, {"qcm", model.qcmcont, nx, nz, 1},
        {"qcp", model.qcpcont, nx, nz, 1}, {"qr", model.qrcont, nx, nz, 1}, {"qrm", model.qrmcont, nx, nz, 1},
        {"qrp", model.qrpcont, nx, nz, 1}, {"u", model.ucont, nx, nz, 1}, {"w", model.wcont, nx, nz, 1},
        {"RKM", model.RKMcont, nx, nz, 1}, {"RKH", model.RKHcont, nx, nz, 1},
        {"init_th_forcing", model.init_th_forcingcont, nx, nz, 1},
        {"evaporation", model.evaporationcont, nx, nz, 1}, {"accretion", model.accretioncont, nx, nz, 1},
        {"autoconversion", model.autoconversioncont, nx, nz, 1},
        {"condensation", model.condensationcont, nx, nz, 1}, {"dth_advect", model.dth_advectcont, nx, nz, 2},
        {"dth_buoyancy", model.dth_buoyancycont, nx, nz, 2}, {"dzeta_advect", model.dzeta_advectcont, nx, nz, 2},
        {"dqv_advect", model.dqv_advectcont, nx, nz, 2}, {"dqc_advect", model.dqc_advectcont, nx, nz, 2},
        {"dqr_advect", model.dqr_advectcont, nx, nz, 2}, {"dqr_VT", model.dqr_VTcont, nx, nz, 2},
        {"precip", model.precip, nx, 1, 1}, {"xi", model.xi, nx, 1, 1}, {"uxi", model.uxi, nx, 1, 1},
        {"rhou", model.rhou, nz, 1, 1}, {"rhow", model.rhow, nz, 1, 1}, {"pib", model.pib, nz, 1, 1},
        {"pb", model.pb, nz, 1, 1}, {"thb", model.thb, nz, 1, 1}, {"thbm", model.thbm, nz, 1, 1},
        {"thb_init", model.thb_init, nz, 1, 1}, {"thb_zeta", model.thb_zeta, nz, 1, 1},
        {"thvb", model.thvb, nz, 1, 1}, {"thvbm", model.thvbm, nz, 1, 1}, {"qvb", model.qvb, nz, 1, 1},
        {"qvb0", model.qvb0, nz, 1, 1}, {"qvsb", model.qvsb, nz, 1, 1}, {"z", model.z, nz, 1, 1},
        {"z_zeta", model.z_zeta, nz, 1, 1}, {"lambda2", model.lambda2, nz, 1, 1}, {"Q1LS", model.Q1LS, nz, 1, 1},
        {"Q2LS", model.Q2LS, nz, 1, 1}};
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "vvm: cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::vector<std::pair<std::string, std::vector<double>>> tail;
    tail.push_back({"ubarTop", {model.ubarTop}});
    tail.push_back({"ubarTopp", {model.ubarTopp}});
    tail.push_back({"ubarTopm", {model.ubarTopm}});
    tail.push_back({"dubarTop_advect0", {model.dubarTop_advect[0]}});
    tail.push_back({"dubarTop_advect1", {model.dubarTop_advect[1]}});
    tail.push_back({"step", {double(model.step)}});
    tail.push_back({"status_for_adding_forcing", {model.status_for_adding_forcing ? 1.0 : 0.0}});
    int32_t head[2] = {1, int32_t(recs.size() + sizeof(arrs) / sizeof(arrs[0]) + tail.size())};
    std::fwrite("VVMSTATE", 1, 8, f);
    std::fwrite(head, sizeof(int32_t), 2, f);
    auto rec = [&](const std::string &n, const double *p, int d0, int d1, int d2) {
        char name[32] = {0};
        std::strncpy(name, n.c_str(), 31);
        int32_t dims[4] = {3, d0, d1, d2};
        std::fwrite(name, 1, 32, f);
        std::fwrite(dims, sizeof(int32_t), 4, f);
        std::fwrite(p, sizeof(double), size_t(d0) * d1 * d2, f);
    };
    for (auto &r : recs) rec(r.first, r.second.data(), 1, 1, 1);
    for (auto &a : arrs) rec(a.n, a.p, a.d0, a.d1, a.d2);
    for (auto &r : tail) rec(r.first, r.second.data(), 1, 1, 1);
    if (std::fclose(f) != 0) {
        std::fprintf(stderr, "vvm: write error on %s\n", path.c_str());
        std::exit(1);
    }
}

// src/Iteration.cpp:208-438.  The device runs the loop body; the host
// takes the state back only on output steps (src/Iteration.cpp:231) and at
// the end.  Steps between two output steps go to the device in one call;
// t_all holds the mean wall time per step of that call and the phase
// timers the device time of its last step.
void vvm::Iteration::TimeMarching(vvm &model) {
    model.sync_to_host();
    model.step = model.start_step;  // 0, or TIMENOW after a restart (src/Iteration.cpp:212-219)
    model.dev_stale = true;
    const int nmax = int(model.TIMEEND / model.dt);
    ensure_device(model);
    const bool radiation = static_cast<bool>(model.solve_radiation) && model.radiation_every > 0;
    if (model.host_microphysics) {
        auto tramp = [](void *user, vvm_host_view *view) -> int {
            return static_cast<vvm *>(user)->host_microphysics(*view);
        };
        check(model, vvm_dev_set_host_microphysics(model.dev, tramp, &model), "vvm_dev_set_host_microphysics");
    } else {
        check(model, vvm_dev_set_host_microphysics(model.dev, nullptr, nullptr), "vvm_dev_set_host_microphysics");
    }
    auto push_radiation = [&]() {
        const int rc = vvm_dev_set_radiation(model.dev, model.radiation_heating_ratecont, model.radiation_cooling ? 1 : 0);
        check(model, rc, "vvm_dev_set_radiation");
    };
    if (radiation || model.radiation_cooling) push_radiation();  // the rate as it stands (zero at a cold start)
    while (model.step < nmax) {
        if (model.OUTPUTSTEP > 0 && is_output_step(model, model.step) && !model.outputpath.empty()) {
            if (model.output_nc_files) vvm::Output::output_nc(model.step, model);
            if (model.output_state_files) vvm::Output::output_state(model.step, model);
        }
        int next = model.step + 1;
        while (next < nmax && !(model.OUTPUTSTEP > 0 && is_output_step(model, next)) &&
               !(radiation && next % model.radiation_every == 0))
            next++;
        const int n = next - model.step;
        vvm_step_stats st{};
        std::vector<double> times(size_t(n) * 5, 0.);
        check(model, vvm_dev_set_step_timers(model.dev, times.data(), n), "vvm_dev_set_step_timers");
        int rc = vvm_dev_step(model.dev, n, &st);
        check(model, rc, "vvm_dev_step");
        check(model, vvm_dev_set_step_timers(model.dev, nullptr, 0), "vvm_dev_set_step_timers");
        if (st.not_converged > 0)
            std::printf("The Poisson solver did not converge in %d of %d steps (last residual %.3e)\n",
                        st.not_converged, n, st.rel_residual_last);
        model.solve_iters_total += st.solve_iters_total;
        model.host_stale = true;
        const int T = model.TIMEROUTPUTSIZE > 0 ? model.TIMEROUTPUTSIZE : 1;
        for (int s = model.step; s < next; s++) {
            if (model.print_steps) std::cout << s << "\n";
            // each step's own phase times (device events of that step)
            const int slot = s % T;
            const double *t = &times[size_t(s - model.step) * 5];
            model.t_advection[slot] = t[0];
            model.t_poisson[slot] = t[1];
            model.t_microphysics[slot] = t[2];
            model.t_diffusion[slot] = t[3];
            model.t_all[slot] = t[4];
        }
        model.step = next;
        if (radiation && next % model.radiation_every == 0) {
            // the reference calls solve_radiation before nextTimeStep, on the
            // step's current level, which the device step has since rotated
            // into the m arrays
            model.sync_to_host();
            std::swap(model.th, model.thm);
            std::swap(model.qv, model.qvm);
            std::swap(model.qc, model.qcm);
            std::swap(model.qr, model.qrm);
            std::swap(model.zeta, model.zetam);
            model.solve_radiation(model);
            std::swap(model.th, model.thm);
            std::swap(model.qv, model.qvm);
            std::swap(model.qc, model.qcm);
            std::swap(model.qr, model.qrm);
            std::swap(model.zeta, model.zetam);
            push_radiation();
        }
    }
    if (model.print_steps) std::cout.flush();
    model.sync_to_host();
}

// key=value configuration file, '#' starting a comment, blanks around keys
// and values dropped (the reference's vvm_read_config, src/ReadConfig.cpp).
std::map<std::string, std::string> vvm_read_config(const std::string &filename) {
    std::map<std::string, std::string> out;
    std::ifstream in(filename);
    std::string line;
    auto trim = [](std::string t) {
        const size_t a = t.find_first_not_of(" \t"), b = t.find_last_not_of(" \t");
        return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
    };
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        line = trim(line);
        const size_t eq = line.find('=');
        // a key needs an '=' with something after it ("KEY=" is skipped)
        if (eq == std::string::npos || eq + 1 >= line.size()) continue;
        out[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    return out;
}
