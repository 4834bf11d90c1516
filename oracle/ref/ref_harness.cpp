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

#include "state_io.hpp"


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
        // the statics of class vvm one at a time (src/Declare.hpp:707-935)
        else if (arg == "advection_zeta") model.Advection_zeta(model);
        else if (arg == "bouyancy") model.Bouyancy(model);
        else if (arg == "advection_thermo_th") model.Advection_thermo(model.thm, model.th, model.thp, model.dth_advect, model);
        else if (arg == "advection_thermo_qv") model.Advection_thermo(model.qvm, model.qv, model.qvp, model.dqv_advect, model);
        else if (arg == "advection_thermo_qc") model.Advection_thermo(model.qcm, model.qc, model.qcp, model.dqc_advect, model);
        else if (arg == "advection_thermo_qr") model.Advection_thermo(model.qrm, model.qr, model.qrp, model.dqr_advect, model);
        else if (arg == "advection_qrvt") model.Advection_qrVT(model);
        else if (arg == "autoconversion") vvm::MicroPhysics::autoconversion(model);
        else if (arg == "accretion") vvm::MicroPhysics::accretion(model);
        else if (arg == "evaporation") vvm::MicroPhysics::evaporation(model);
        else if (arg == "condensation") vvm::MicroPhysics::condensation(model);
        else if (arg == "negval_qcp") vvm::MicroPhysics::NegativeValueProcess(model.qcp, model.nx, model.nz);
        else if (arg == "negval_qc") vvm::MicroPhysics::NegativeValueProcess(model.qc, model.nx, model.nz);
        else if (arg == "hparam_qv") vvm::Turbulence::Hparam(model, model.qv, model.qvp);
        else if (arg == "hparam_qc") vvm::Turbulence::Hparam(model, model.qc, model.qcp);
        else if (arg == "hparam_qr") vvm::Turbulence::Hparam(model, model.qr, model.qrp);
        else if (arg == "nudge_theta") vvm::NumericalProcess::Nudge_theta(model);
        else if (arg == "nudge_zeta") vvm::NumericalProcess::Nudge_zeta(model);
        else if (arg == "nudge_qv") vvm::NumericalProcess::Nudge_qv(model);
        else if (arg == "add_forcing") model.AddForcing(model);
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
