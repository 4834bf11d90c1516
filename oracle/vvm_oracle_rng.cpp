/*
 * vvm_oracle_rng.cpp -- the oracle's restatement of the reference's
 * vvm::Init::RandomPerturbation (src/Init.cpp:488-515).
 *
 * TEST INFRASTRUCTURE ONLY: part of the checker (oracle/liboracle.so).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the product never links or calls it.
 *
 * The reference draws its theta kick from libstdc++'s std::mt19937 seeded
 * with the step number and std::normal_distribution<double>(0, 1), rejecting
 * draws outside [min_range, max_range], k outer and i inner, below z = 200 m
 * only, then applies BoundaryProcess2D_center (src/Boundary.cpp:17-31).  No
 * closed form reproduces that stream, so this file calls the same standard
 * library engine and distribution (this is C++ for that reason alone; the
 * rest of the oracle is C).  Pinned bit for bit against the compiled
 * reference's own RandomPerturbation (tests/golden/misc.npz rp/1, rp/7).
 */
#include <random>

#include "vvm_oracle.h"

extern "C" int vo_random_perturbation(const vvm_dev_config *c, vvm_host_view *v, int t, double min_range,
                                      double max_range, double stddev) {
    if (!c || !v || !v->init_th_forcing) return -1;
    const int nx = c->nx, nz = c->nz;
    double *f = v->init_th_forcing;
    std::mt19937 engine(t);
    engine.seed(t);
    std::normal_distribution<> normal(0.0, 1.0);
    for (int k = 1; k < nz - 1; k++) {
        const double z = (k - 0.5) * c->dz;
        for (int i = 1; i < nx - 1; i++) {
            double draw = 0.;
            if (z < 200) {
                do {
                    draw = 0. + stddev * normal(engine);
                } while (draw < min_range || draw > max_range);
            }
            f[(size_t)i * nz + k] = draw;
        }
    }
    /* BoundaryProcess2D_center: periodic x over the interior rows, then the
     * z ghosts copy their neighbours (corners follow from the x pass). */
    for (int k = 1; k <= nz - 2; k++) {
        f[(size_t)0 * nz + k] = f[(size_t)(nx - 2) * nz + k];
        f[(size_t)(nx - 1) * nz + k] = f[(size_t)1 * nz + k];
    }
    for (int i = 0; i < nx; i++) {
        f[(size_t)i * nz + 0] = f[(size_t)i * nz + 1];
        f[(size_t)i * nz + nz - 1] = f[(size_t)i * nz + nz - 2];
    }
    return 0;
}
