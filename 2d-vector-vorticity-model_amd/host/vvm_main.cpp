// vvm_main.cpp -- the model driver of src/main.cpp on the device timestep.
//
//   vvm_main <config.txt> [--dry-run]
//
// The config file has the reference's KEY=value lines (vvm_config.txt,
// read by vvm_read_config; `#` starts a comment).  The keys of the
// reference are honoured with the same meaning:
//   VVM_OUTPUTPATH VVM_XRANGE VVM_ZRANGE VVM_DT VVM_DX VVM_DZ VVM_TIMEEND
//   VVM_OUTPUTSTEP VVM_CASE VVM_MOISTURE_NUDGE_TIME VVM_YEAR VVM_MONTH
//   VVM_DAY VVM_HOUR VVM_MINUTE VVM_SECOND VVM_LON VVM_LAT
// and the constants main() passes to Config_VVM (src/main.cpp:68-72) are
// the defaults.  Further keys choose the initial state and the back-end:
//   VVM_INIT (case|reference|previous): the benchmark cases below, the
//     reference's own Init1d/Init2d with VVM_SOUNDING_FILE (the
//     LOADINITPATH sounding; empty for the analytic sounding), or its
//     LoadFromPreviousFile restart with VVM_RESTART_PREV, VVM_RESTART_NOW
//     (LOADPATH1/2: two consecutive nc outputs) and VVM_TIMENOW
//   VVM_ADDFORCINGTIME (10, as main() passes), VVM_TROPICAL_FORCING (0|1)
//   VVM_BRUNT VVM_BUBBLE_DT VVM_BUBBLE_RADIUS VVM_BUBBLE_ZC
//   VVM_NOISE (none|uniform|gaussian) VVM_NOISE_AMP VVM_NOISE_ZTOP VVM_SEED
//   VVM_SFC_HEAT_FLUX VVM_SFC_QV_RAMP VVM_SFC_QV_RAMP_Z
//   VVM_SOLVER (fp64|mixed|bicgstab) VVM_GUESS (zero|prev|extrap)
//   VVM_STOP (parity|rel|abs; abs: bicgstab only, AMGX's absolute criterion)
//   VVM_TOL_REL VVM_MAX_ITER VVM_GPU VVM_PRINT_STEPS (0|1)
//   VVM_FINAL_STATE (path of a VVMSTATE file written after the run)
//   VVM_OUTPUT_NC (1: <outputpath>/nc/<step>.nc on output steps, as the
//     reference) VVM_OUTPUT_NC_SAMEFILE (1: the OUTPUTNCSAMEFILE layout, one
//     file with a record per output step) VVM_OUTPUT_MICROPHYSICS (1: the
//     Kessler rates in that file) VVM_OUTPUT_STATE (1: also
//     <outputpath>/state/state_<step>.bin)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <string>

#include "ReadConfig.hpp"
#include "vvm.hpp"

namespace {


std::map<std::string, std::string> read_config(const std::string &path) {
    if (!std::ifstream(path)) {
        std::fprintf(stderr, "vvm_main: cannot open %s\n", path.c_str());
        std::exit(1);
    }
    return vvm_read_config(path);
}

struct Keys {
    std::map<std::string, std::string> m;
    double d(const char *k, double def) const {
        auto it = m.find(k);
        return it == m.end() ? def : std::stod(it->second);
    }
    int i(const char *k, int def) const {
        auto it = m.find(k);
        return it == m.end() ? def : std::stoi(it->second);
    }
    std::string s(const char *k, const std::string &def) const {
        auto it = m.find(k);
        return it == m.end() ? def : it->second;
    }
    double need(const char *k) const {
        auto it = m.find(k);
        if (it == m.end()) {
            std::fprintf(stderr, "vvm_main: missing key %s\n", k);
            std::exit(1);
        }
        return std::stod(it->second);
    }
};

int pick(const std::string &v, std::initializer_list<std::pair<const char *, int>> opts, const char *key) {
    for (auto &o : opts)
        if (v == o.first) return o.second;
    std::fprintf(stderr, "vvm_main: bad value '%s' for %s\n", v.c_str(), key);
    std::exit(1);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <config.txt> [--dry-run]\n", argv[0]);
        return 2;
    }
    const bool dry = argc > 2 && std::string(argv[2]) == "--dry-run";
    auto t_start = std::chrono::steady_clock::now();
    Keys K{read_config(argv[1])};
    std::string out = K.s("VVM_OUTPUTPATH", "");
    if (!out.empty() && out.back() != '/') out += "/";

    Config_VVM config(K.need("VVM_DT"), K.need("VVM_DX"), K.need("VVM_DZ"), int(K.need("VVM_XRANGE")),
                      int(K.need("VVM_ZRANGE")), K.need("VVM_TIMEEND"), 10000, out, K.i("VVM_OUTPUTSTEP", 0), 10.,
                      10., 0.01, 1E-22, 9.80665, 1003.5, 716.5, 287., 2.5E6, 1E5, 96500.,
                      K.d("VVM_ADDFORCINGTIME", 10.), K.i("VVM_CASE", 0), K.d("VVM_MOISTURE_NUDGE_TIME", 0.),
                      K.i("VVM_YEAR", 2025), K.i("VVM_MONTH", 3), K.i("VVM_DAY", 21), K.d("VVM_HOUR", 12.),
                      K.d("VVM_MINUTE", 0.), K.d("VVM_SECOND", 0.), K.d("VVM_LON", 0.), K.d("VVM_LAT", 0.));
    config.gpu_id = K.i("VVM_GPU", 0);
    config.solver_mode =
        pick(K.s("VVM_SOLVER", "fp64"),
             {{"fp64", VVM_SOLVER_FP64}, {"mixed", VVM_SOLVER_MIXED}, {"bicgstab", VVM_SOLVER_BICGSTAB}}, "VVM_SOLVER");
    config.guess_mode = pick(K.s("VVM_GUESS", "zero"),
                             {{"zero", VVM_GUESS_ZERO}, {"prev", VVM_GUESS_PREV}, {"extrap", VVM_GUESS_EXTRAP},
                              {"extrap2", VVM_GUESS_EXTRAP2}},
                             "VVM_GUESS");
    config.stop_mode =
        pick(K.s("VVM_STOP", "parity"), {{"parity", VVM_STOP_PARITY}, {"rel", VVM_STOP_REL}, {"abs", VVM_STOP_ABS}},
             "VVM_STOP");
    config.tol_rel = K.d("VVM_TOL_REL", 1e-10);
    config.max_iter = K.i("VVM_MAX_ITER", 2000);
    config.tropical_forcing = K.i("VVM_TROPICAL_FORCING", 0);

    vvm_case_params p{};
    p.sounding = 0;
    p.brunt = K.d("VVM_BRUNT", 0.01);
    p.bubble_dT = K.d("VVM_BUBBLE_DT", 0.);
    p.bubble_radius = K.d("VVM_BUBBLE_RADIUS", 2000.);
    p.bubble_zc = K.d("VVM_BUBBLE_ZC", 2000.);
    p.noise_kind = pick(K.s("VVM_NOISE", "none"), {{"none", 0}, {"uniform", 1}, {"gaussian", 2}}, "VVM_NOISE");
    p.noise_amp = K.d("VVM_NOISE_AMP", 0.);
    p.noise_ztop = K.d("VVM_NOISE_ZTOP", 0.);
    p.seed = uint64_t(K.d("VVM_SEED", 1));
    p.sfc_heat_flux = K.d("VVM_SFC_HEAT_FLUX", 0.);
    p.sfc_qv_ramp = K.d("VVM_SFC_QV_RAMP", 0.);
    p.sfc_qv_ramp_z = K.d("VVM_SFC_QV_RAMP_Z", 500.);

    vvm model(config);
    model.print_steps = K.i("VVM_PRINT_STEPS", 1) != 0;
    model.output_nc_files = K.i("VVM_OUTPUT_NC", 1) != 0;
    model.output_nc_same_file = K.i("VVM_OUTPUT_NC_SAMEFILE", 0) != 0;
    model.output_microphysics = K.i("VVM_OUTPUT_MICROPHYSICS", 0) != 0;
    model.output_state_files = K.i("VVM_OUTPUT_STATE", 0) != 0;
    std::printf("nx=%d nz=%d dx=%g dz=%g dt=%g steps=%d solver=%s\n", model.nx, model.nz, model.dx, model.dz,
                model.dt, int(model.TIMEEND / model.dt), config.solver_mode == VVM_SOLVER_MIXED      ? "mixed"
                : config.solver_mode == VVM_SOLVER_BICGSTAB ? "bicgstab"
                                                            : "fp64");
    if (dry) return 0;

    if (K.s("VVM_INIT", "case") == "previous") {
        // restart from two outputs of an earlier run (src/main.cpp:75-76)
        model.sounding_path = K.s("VVM_SOUNDING_FILE", "");
        model.restart_prev_nc = K.s("VVM_RESTART_PREV", "");
        model.restart_now_nc = K.s("VVM_RESTART_NOW", "");
        model.start_step = K.i("VVM_TIMENOW", 1);
        vvm::Init::LoadFromPreviousFile(model);
    } else if (K.s("VVM_INIT", "case") == "reference") {
        // the reference's own Init1d / Init2d (src/main.cpp:78-83)
        model.sounding_path = K.s("VVM_SOUNDING_FILE", "");
        vvm::Init::Init1d(model);
        vvm::Init::Init2d(model);
    } else {
        vvm::Init::InitCase(model, p);
    }
    vvm::PoissonSolver::InitPoissonMatrix(model);
    vvm::Output::create_all_directory(model);
    vvm::Iteration::TimeMarching(model);

    std::string fin = K.s("VVM_FINAL_STATE", "");
    if (!fin.empty()) vvm::Output::write_state(fin, model);
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::printf("Total time: %.6f s, Poisson iterations: %lld\n", wall, model.solve_iters_total);
    return 0;
}
