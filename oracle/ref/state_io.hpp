// state_io.hpp -- VVMSTATE record files <-> the arrays of a model whose
// class has the reference's member names (src/Declare.hpp): the reference
// itself (ref_harness.cpp) or the drop-in C++ API (dropin_harness.cpp).
// Test infrastructure only.
//
// State file: "VVMSTATE", int32 version, int32 nrec, then records
//   { char name[32]; int32 ndim; int32 dims[3]; double data[prod(dims)] }.
// 2-D fields are [i][k] (k fastest), AB2 histories [i][k][2], as in the
// reference's contiguous blocks (src/Declare.hpp:400-427).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
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
