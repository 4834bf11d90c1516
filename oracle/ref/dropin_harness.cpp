// dropin_harness.cpp -- the reference's own TimeMarching (src/Iteration.cpp,
// compiled unchanged from the reference tree) driving the drop-in C++ API
// (include/Declare.hpp -> vvm.hpp, libvvm.so): every operator call of the
// reference's step lands on a device-backed static of this build.  Test
// infrastructure: a check that the per-operator API is a drop-in for the
// reference's own call sequence; see build_ref.sh.
//
// Usage: vvm_dropin <nsteps> <in.state> <out.state>
// The state file is ref_harness.cpp's; the configuration constants are the
// ones src/main.cpp passes (src/main.cpp:68-72), with addforcingtime, CASE
// and the nudge time from the file, as ref_harness.cpp does.
#include "Declare.hpp"

#include <cmath>
#include <cstdio>

#include "state_io.hpp"

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <nsteps> <in.state> <out.state>\n", argv[0]);
        return 2;
    }
    const int nsteps = std::atoi(argv[1]);
    RecMap recs = read_state(argv[2]);
    const int nx = int(get1(recs, "cfg.nx", 0)), nz = int(get1(recs, "cfg.nz", 0));
    const double dx = get1(recs, "cfg.dx", 200.), dz = get1(recs, "cfg.dz", 200.), dt = get1(recs, "cfg.dt", 1.);
    const int xrange = int(std::lround((nx - 2) * dx)), zrange = int(std::lround((nz - 2) * dz));
    Config_VVM config(dt, dx, dz, xrange, zrange, nsteps * dt, 10000, "/dev/null/", 1 << 30, 10., 10., 0.01,
                      get1(recs, "cfg.tolerance", 1E-22), get1(recs, "cfg.GRAVITY", 9.80665),
                      get1(recs, "cfg.Cp", 1003.5), get1(recs, "cfg.Cv", 716.5), get1(recs, "cfg.Rd", 287.),
                      get1(recs, "cfg.Lv", 2.5E6), get1(recs, "cfg.P0", 1E5), get1(recs, "cfg.PSURF", 96500.),
                      get1(recs, "cfg.addforcingtime", 0.), int(get1(recs, "cfg.CASE", 1)),
                      get1(recs, "cfg.moisture_nudge_time", 0.), 2025, 3, 21, 12., 0., 0., 0., 0.);
    config.tropical_forcing = 1;  // the reference build's TROPICALFORCING (build_ref.sh)
    vvm model(config);
    if (model.nx != nx || model.nz != nz) {
        std::fprintf(stderr, "grid mismatch: model %dx%d state %dx%d\n", model.nx, model.nz, nx, nz);
        return 2;
    }
    model.print_steps = false;
    load_into(model, recs);
    vvm::PoissonSolver::InitPoissonMatrix(model);
    vvm::Iteration::TimeMarching(model);  // the reference's, from src/Iteration.cpp
    Writer w;
    save_from(model, w);
    w.save(argv[3]);
    return 0;
}
