// ref_harness.cpp -- drives the reference model compiled from its own
// sources (see build_ref.sh) on a state file, for golden vectors and the
// CPU baseline.  Test infrastructure only: never linked into the product.
//
// Usage:
//   vvm_ref step <nsteps> <in.state> <out.state>     TimeMarching (src/Iteration.cpp:208-438)
//   vvm_ref op <name> <in.state> <out.state>         one reference operator
//   vvm_ref solve <tol> <in.state> <out.state>       cal_w's Eigen solve at a given tolerance
//   vvm_ref init - <in.state> <out.state>            Init1d + Init2d on the grid of in.state
//
// State file: "VVMSTATE", int32 version, int32 nrec, then records
//   { char name[32]; int32 ndim; int32 dims[3]; double data[prod(dims)] }.
// 2-D fields are [i][k] (k fastest), AB2 histories [i][k][2], as in the
// reference's contiguous blocks (src/Declare.hpp:400-427).

#include "Declare.hpp"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Rec {
    std::vector<int32_t> dims;
    std::vector<double> data;
};
using RecMap = std::map<std::string, Rec>;

RecMap read_state(const char *path) {
    RecMap m;
    std::ifstream f(path, std::ios::binary);
    if (!f) { std::fprintf(stderr, "cannot open %s\n", path); std::exit(2); }
    char magic[8];
    f.read(magic, 8);
    if (std::memcmp(magic, "VVMSTATE", 8) != 0) { std::fprintf(stderr, "bad magic\n"); std::exit(2); }
    int32_t ver = 0, nrec = 0;
    f.read(reinterpret_cast<char *>(&ver), 4);
    f.read(reinterpret_cast<char *>(&nrec), 4);
    for (int r = 0; r < nrec; ++r) {
        char name[33] = {0};
        int32_t ndim = 0, dims[3] = {1, 1, 1};
        f.read(name, 32);
        f.read(reinterpret_cast<char *>(&ndim), 4);
        f.read(reinterpret_cast<char *>(dims), 12);
        size_t n = size_t(dims[0]) * dims[1] * dims[2];
        Rec rec;
        rec.dims.assign(dims, dims + 3);
        rec.data.resize(n);
        f.read(reinterpret_cast<char *>(rec.data.data()), std::streamsize(n * 8));
        if (!f) { std::fprintf(stderr, "truncated state file\n"); std::exit(2); }
        m[name] = std::move(rec);
    }
    return m;
}

struct Writer {
    std::vector<std::pair<std::string, Rec>> recs;
    void add(const std::string &name, const double *p, int d0, int d1 = 1, int d2 = 1) {
        Rec r;
        r.dims = {d0, d1, d2};
        r.data.assign(p, p + size_t(d0) * d1 * d2);
        recs.emplace_back(name, std::move(r));
    }
    void scalar(const std::string &name, double v) { add(name, &v, 1); }
    void save(const char *path) {
        std::ofstream f(path, std::ios::binary);
        f.write("VVMSTATE", 8);
        int32_t ver = 1, nrec = int32_t(recs.size());
        f.write(reinterpret_cast<char *>(&ver), 4);
        f.write(reinterpret_cast<char *>(&nrec), 4);
        for (auto &kv : recs) {
            char name[32] = {0};
            std::strncpy(name, kv.first.c_str(), 31);
            int32_t ndim = 3;
            f.write(name, 32);
            f.write(reinterpret_cast<char *>(&ndim), 4);
            f.write(reinterpret_cast<const char *>(kv.second.dims.data()), 12);
            f.write(reinterpret_cast<const char *>(kv.second.data.data()),
                    std::streamsize(kv.second.data.size() * 8));
        }
    }
};

double get1(const RecMap &m, const char *name, double dflt) {
    auto it = m.find(name);
    return it == m.end() ? dflt : it->second.data[0];
}

// Name -> storage of every array of the Kessler / AB2 build of class vvm.
struct Binding {
    const char *name;
    double *ptr;
    int d0, d1, d2;
};

std::vector<Binding> bindings(vvm &m) {
    const int nx = m.nx, nz = m.nz;
    return {
        {"zeta", m.zetacont, nx, nz, 1}, {"zetam", m.zetamcont, nx, nz, 1}, {"zetap", m.zetapcont, nx, nz, 1},
        {"th", m.thcont, nx, nz, 1},     {"thm", m.thmcont, nx, nz, 1},     {"thp", m.thpcont, nx, nz, 1},
        {"qv", m.qvcont, nx, nz, 1},     {"qvm", m.qvmcont, nx, nz, 1},     {"qvp", m.qvpcont, nx, nz, 1},
        {"qc", m.qccont, nx, nz, 1},     {"qcm", m.qcmcont, nx, nz, 1},     {"qcp", m.qcpcont, nx, nz, 1},
        {"qr", m.qrcont, nx, nz, 1},     {"qrm", m.qrmcont, nx, nz, 1},     {"qrp", m.qrpcont, nx, nz, 1},
        {"u", m.ucont, nx, nz, 1},       {"w", m.wcont, nx, nz, 1},
        {"RKM", m.RKMcont, nx, nz, 1},   {"RKH", m.RKHcont, nx, nz, 1},
        {"init_th_forcing", m.init_th_forcingcont, nx, nz, 1},
        {"evaporation", m.evaporationcont, nx, nz, 1}, {"accretion", m.accretioncont, nx, nz, 1},
        {"autoconversion", m.autoconversioncont, nx, nz, 1}, {"condensation", m.condensationcont, nx, nz, 1},
        {"dth_advect", m.dth_advectcont, nx, nz, 2},     {"dth_buoyancy", m.dth_buoyancycont, nx, nz, 2},
        {"dzeta_advect", m.dzeta_advectcont, nx, nz, 2}, {"dqv_advect", m.dqv_advectcont, nx, nz, 2},
        {"dqc_advect", m.dqc_advectcont, nx, nz, 2},     {"dqr_advect", m.dqr_advectcont, nx, nz, 2},
        {"dqr_VT", m.dqr_VTcont, nx, nz, 2},
        {"precip", m.precip, nx, 1, 1}, {"xi", m.xi, nx, 1, 1}, {"uxi", m.uxi, nx, 1, 1},
        {"rhou", m.rhou, nz, 1, 1},     {"rhow", m.rhow, nz, 1, 1},   {"pib", m.pib, nz, 1, 1},
        {"pb", m.pb, nz, 1, 1},         {"thb", m.thb, nz, 1, 1},     {"thbm", m.thbm, nz, 1, 1},
        {"thb_init", m.thb_init, nz, 1, 1}, {"thb_zeta", m.thb_zeta, nz, 1, 1},
        {"thvb", m.thvb, nz, 1, 1},     {"thvbm", m.thvbm, nz, 1, 1}, {"qvb", m.qvb, nz, 1, 1},
        {"qvb0", m.qvb0, nz, 1, 1},     {"qvsb", m.qvsb, nz, 1, 1},   {"z", m.z, nz, 1, 1},
        {"z_zeta", m.z_zeta, nz, 1, 1}, {"lambda2", m.lambda2, nz, 1, 1},
#if defined(TROPICALFORCING)
        {"Q1LS", m.Q1LS, nz, 1, 1},     {"Q2LS", m.Q2LS, nz, 1, 1},
#endif
    };
}

void load_into(vvm &m, const RecMap &recs) {
    for (auto &b : bindings(m)) {
        auto it = recs.find(b.name);
        if (it == recs.end()) continue;
        size_t n = size_t(b.d0) * b.d1 * b.d2;
        if (it->second.data.size() != n) {
            std::fprintf(stderr, "size mismatch for %s\n", b.name);
            std::exit(2);
        }
        std::memcpy(b.ptr, it->second.data.data(), n * 8);
    }
    m.ubarTop = get1(recs, "ubarTop", 0.);
    m.ubarTopp = get1(recs, "ubarTopp", 0.);
    m.ubarTopm = get1(recs, "ubarTopm", 0.);
    m.dubarTop_advect[0] = get1(recs, "dubarTop_advect0", 0.);
    m.dubarTop_advect[1] = get1(recs, "dubarTop_advect1", 0.);
    m.step = int(get1(recs, "step", 0.));
    m.status_for_adding_forcing = get1(recs, "status_for_adding_forcing", 0.) != 0.;
}

void save_from(vvm &m, Writer &w) {
    for (auto &b : bindings(m)) w.add(b.name, b.ptr, b.d0, b.d1, b.d2);
    w.scalar("ubarTop", m.ubarTop);
    w.scalar("ubarTopp", m.ubarTopp);
    w.scalar("ubarTopm", m.ubarTopm);
    w.scalar("dubarTop_advect0", m.dubarTop_advect[0]);
    w.scalar("dubarTop_advect1", m.dubarTop_advect[1]);
    w.scalar("step", m.step);
    w.scalar("status_for_adding_forcing", m.status_for_adding_forcing ? 1. : 0.);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s step|op|solve <arg> <in.state> <out.state>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const std::string arg = argv[2];
    RecMap recs = read_state(argv[3]);

    const int nx = int(get1(recs, "cfg.nx", 0));
    const int nz = int(get1(recs, "cfg.nz", 0));
    const double dx = get1(recs, "cfg.dx", 200.), dz = get1(recs, "cfg.dz", 200.);
    const double dt = get1(recs, "cfg.dt", 1.);
    const int xrange = int(std::lround((nx - 2) * dx));
    const int zrange = int(std::lround((nz - 2) * dz));
    int nsteps = (mode == "step") ? std::atoi(arg.c_str()) : 1;
    // The constants main() passes to Config_VVM (src/main.cpp:68-72), with
    // addforcingtime, CASE and the nudge time taken from the state file.
    Config_VVM config(dt, dx, dz, xrange, zrange, nsteps * dt, 10000, "/dev/null/", 1 << 30,
                      10., 10., 0.01, get1(recs, "cfg.tolerance", 1E-22), get1(recs, "cfg.GRAVITY", 9.80665),
                      get1(recs, "cfg.Cp", 1003.5), get1(recs, "cfg.Cv", 716.5), get1(recs, "cfg.Rd", 287.),
                      get1(recs, "cfg.Lv", 2.5E6), get1(recs, "cfg.P0", 1E5), get1(recs, "cfg.PSURF", 96500.),
                      get1(recs, "cfg.addforcingtime", 0.), int(get1(recs, "cfg.CASE", 1)),
                      get1(recs, "cfg.moisture_nudge_time", 0.), 2025, 3, 21, 12., 0., 0., 0., 0.);
    vvm model(config);
    if (model.nx != nx || model.nz != nz) {
        std::fprintf(stderr, "grid mismatch: model %dx%d state %dx%d\n", model.nx, model.nz, nx, nz);
        return 2;
    }
    if (mode == "init") {
        // the reference's own initial state (Init1d + Init2d as main() calls
        // them, src/main.cpp:78-83); the input file supplies the grid only
        vvm::Init::Init1d(model);
        vvm::Init::Init2d(model);
        Writer w;
        save_from(model, w);
        w.save(argv[4]);
        return 0;
    }
    load_into(model, recs);
    vvm::PoissonSolver::InitPoissonMatrix(model);

    // The reference prints progress and solver warnings on stdout; keep them.
    std::ostringstream captured;
    std::streambuf *old = std::cout.rdbuf(captured.rdbuf());
    auto t0 = std::chrono::steady_clock::now();
    double extra_iters = -1., extra_err = -1.;
    if (mode == "step") {
        vvm::Iteration::TimeMarching(model);
    } else if (mode == "op") {
        if (arg == "pzeta_pt") vvm::Iteration::pzeta_pt(model);
        else if (arg == "pth_pt") vvm::Iteration::pth_pt(model);
        else if (arg == "pqv_pt") vvm::Iteration::pqv_pt(model);
        else if (arg == "pqc_pt") vvm::Iteration::pqc_pt(model);
        else if (arg == "pqr_pt") vvm::Iteration::pqr_pt(model);
        else if (arg == "forcing_bc") { model.AddForcing(model); vvm::BoundaryProcess2D_all(model); }
        else if (arg == "bc_all") vvm::BoundaryProcess2D_all(model);
        else if (arg == "pubarTop_pt") vvm::PoissonSolver::pubarTop_pt(model);
        else if (arg == "cal_w") vvm::PoissonSolver::cal_w(model);
        else if (arg == "cal_u") vvm::PoissonSolver::cal_u(model);
        else if (arg == "microphysics") {
            vvm::MicroPhysics::autoconversion(model);
            vvm::MicroPhysics::accretion(model);
            vvm::MicroPhysics::evaporation(model);
            vvm::MicroPhysics::condensation(model);
            vvm::MicroPhysics::NegativeValueProcess(model.qvp, model.nx, model.nz);
            vvm::MicroPhysics::NegativeValueProcess(model.qcp, model.nx, model.nz);
            vvm::MicroPhysics::NegativeValueProcess(model.qrp, model.nx, model.nz);
            vvm::BoundaryProcess2D_all(model);
        }
        else if (arg == "negval_qvp") vvm::MicroPhysics::NegativeValueProcess(model.qvp, model.nx, model.nz);
        else if (arg == "turbulence") vvm::Turbulence::RKM_RKH(model);
        else if (arg == "mparam") vvm::Turbulence::Mparam(model, model.zeta, model.zetap);
        else if (arg == "hparam_th") vvm::Turbulence::Hparam(model, model.th, model.thp);
        else if (arg == "nudge") {
            vvm::NumericalProcess::Nudge_theta(model);
            if (model.CASE != 2) vvm::NumericalProcess::Nudge_zeta(model);
            vvm::NumericalProcess::Nudge_qv(model);
        }
        else if (arg == "updateMean") vvm::Iteration::updateMean(model);
        else if (arg == "nextTimeStep") vvm::Iteration::nextTimeStep(model);
        else if (arg == "random_perturbation") vvm::Init::RandomPerturbation(model, model.step, -0.001, 0.001, 1.);
        else { std::cout.rdbuf(old); std::fprintf(stderr, "unknown op %s\n", arg.c_str()); return 2; }
    } else if (mode == "solve") {
        // cal_w's right-hand side and solver (src/PoissonSolver.cpp:441-462)
        // with the tolerance given on the command line.
        const int n = (model.nx - 2) * (model.nz - 3);
        Eigen::VectorXd x(n), b(n);
        int count = 0;
        for (int k = 2; k <= model.nz - 2; k++)
            for (int i = 1; i <= model.nx - 2; i++)
                b(count++) = model.rhow[k] * model.rhow[k] * (model.zetap[i + 1][k] - model.zetap[i][k]) * model.dx;
        Eigen::BiCGSTAB<Eigen::SparseMatrix<double>, Eigen::IncompleteLUT<double>> solver;
        solver.setTolerance(std::atof(arg.c_str()));
        solver.setMaxIterations(10000);
        Eigen::SparseMatrix<double> A = model.A;
        solver.compute(A);
        x = solver.solve(b);
        extra_iters = double(solver.iterations());
        extra_err = (b - model.A * x).norm() / b.norm();
        int cnt = 0;
        for (int k = 2; k <= model.nz - 2; k++)
            for (int i = 1; i <= model.nx - 2; i++) model.w[i][k] = x[cnt++] / model.rhow[k];
    } else {
        std::cout.rdbuf(old);
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout.rdbuf(old);
    const std::string log = captured.str();
    int warnings = 0;
    for (size_t pos = log.find("Warning"); pos != std::string::npos; pos = log.find("Warning", pos + 1)) ++warnings;

    Writer w;
    save_from(model, w);
    w.scalar("wall_seconds", wall);
    w.scalar("warnings", warnings);
    if (extra_iters >= 0) {
        w.scalar("solve_iterations", extra_iters);
        w.scalar("solve_rel_residual", extra_err);
    }
    if (mode == "step") {
        const int n = nsteps < model.TIMEROUTPUTSIZE ? nsteps : model.TIMEROUTPUTSIZE;
        w.add("t_advection", model.t_advection, n);
        w.add("t_poisson", model.t_poisson, n);
        w.add("t_microphysics", model.t_microphysics, n);
        w.add("t_diffusion", model.t_diffusion, n);
        w.add("t_all", model.t_all, n);
    }
    w.save(argv[4]);
    return 0;
}
