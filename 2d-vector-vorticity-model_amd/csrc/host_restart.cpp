// host_restart.cpp -- vvm::Init::LoadFromPreviousFile (src/Init.cpp:372-485):
// the sounding file's profiles plus the fields of two consecutive netCDF
// outputs of an earlier run (the previous step's th, zeta, qv, qc, qr and
// ubarTop; the current step's th, zeta, qv, qc, qr, u, w), then the
// horizontal means.  Host code; it includes a reader for netCDF classic
// files (CDF-1 and CDF-2 64-bit offsets) sufficient for the reference's
// per-step output (double variables).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "vvm_dev.h"

namespace {

inline double &at(double *a, int nz, int i, int k) { return a[size_t(i) * nz + k]; }

void bc1d(double *a, int n) {
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

// netCDF classic reader: variables of type double, no record dimension.
class NcFile {
   public:
    bool open(const std::string &path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) return false;
        buf_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        pos_ = 0;
        if (buf_.size() < 8 || buf_[0] != 'C' || buf_[1] != 'D' || buf_[2] != 'F') return false;
        version_ = buf_[3];
        if (version_ != 1 && version_ != 2) return false;
        pos_ = 4;
        u32();  // numrecs
        uint32_t tag = u32(), n = u32();
        if (tag != 0 && tag != 10) return false;
        for (uint32_t d = 0; d < n; d++) {
            name();
            dims_.push_back(u32());
        }
        if (!skip_atts()) return false;
        tag = u32(), n = u32();
        if (tag != 0 && tag != 11) return false;
        for (uint32_t v = 0; v < n; v++) {
            Var var;
            const std::string nm = name();
            const uint32_t nd = u32();
            var.count = 1;
            for (uint32_t d = 0; d < nd; d++) {
                const uint32_t id = u32();
                if (id >= dims_.size()) return false;
                var.count *= dims_[id];
            }
            if (!skip_atts()) return false;
            var.type = u32();
            u32();  // vsize
            var.begin = (version_ == 2) ? u64() : u32();
            vars_[nm] = var;
        }
        return pos_ <= buf_.size();
    }
    // the variable's values (doubles) or an empty vector
    std::vector<double> get(const std::string &nm) const {
        auto it = vars_.find(nm);
        if (it == vars_.end() || it->second.type != 6) return {};
        const Var &v = it->second;
        if (v.begin + v.count * 8 > buf_.size()) return {};
        std::vector<double> out(v.count);
        for (size_t i = 0; i < v.count; i++) {
            uint64_t u = 0;
            for (int b = 0; b < 8; b++) u = (u << 8) | static_cast<unsigned char>(buf_[v.begin + i * 8 + b]);
            std::memcpy(&out[i], &u, 8);
        }
        return out;
    }

   private:
    struct Var {
        size_t count = 0, begin = 0;
        uint32_t type = 0;
    };
    uint32_t u32() {
        if (pos_ + 4 > buf_.size()) {
            pos_ = buf_.size() + 1;
            return 0;
        }
        uint32_t v = 0;
        for (int b = 0; b < 4; b++) v = (v << 8) | static_cast<unsigned char>(buf_[pos_ + b]);
        pos_ += 4;
        return v;
    }
    uint64_t u64() {
        const uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    std::string name() {
        const uint32_t len = u32();
        if (pos_ + len > buf_.size()) {
            pos_ = buf_.size() + 1;
            return "";
        }
        std::string s(buf_.data() + pos_, len);
        pos_ += (len + 3) / 4 * 4;
        return s;
    }
    bool skip_atts() {
        const uint32_t tag = u32(), n = u32();
        if (tag != 0 && tag != 12) return false;
        for (uint32_t a = 0; a < n; a++) {
            name();
            const uint32_t type = u32(), cnt = u32();
            static const int size[] = {0, 1, 1, 2, 4, 4, 8};
            if (type < 1 || type > 6) return false;
            pos_ += (size_t(cnt) * size[type] + 3) / 4 * 4;
        }
        return pos_ <= buf_.size();
    }
    std::vector<char> buf_;
    size_t pos_ = 0;
    int version_ = 0;
    std::vector<uint32_t> dims_;
    std::map<std::string, Var> vars_;
};

double get_th(int i, int k, const vvm_dev_config &c) {  // src/Init.cpp:279-292
    const double XC = (c.nx * c.dx) / 2., XR = 4000.;
    const double ZC = 2500., ZR = 2000.;
    const double x = (i - 0.5) * c.dx, z = (k - 0.5) * c.dz;
    const double rad = std::sqrt(std::pow((x - XC) / XR, 2) + std::pow((z - ZC) / ZR, 2));
    if (rad <= 1) return 0.5 * 3. * (std::cos(M_PI * rad) + 1);
    return 0.;
}

}  // namespace

extern "C" int vvm_init_from_previous(const vvm_dev_config *c, const char *sounding_path, const char *nc_prev,
                                      const char *nc_now, vvm_host_view *v) {
    if (!c || !v || !sounding_path || !nc_prev || !nc_now) return VVM_ERR_ARG;
    const int nx = c->nx, nz = c->nz;
    // the profile part of LoadFromPreviousFile (src/Init.cpp:375-401)
    std::ifstream in(sounding_path);
    if (!in) return VVM_ERR_ARG;
    std::string line;
    std::getline(in, line);
    std::getline(in, line);
    double ZZ, ZT, RHO, THBAR, PBAR, PIBAR, QVBAR, Q1, Q2, RHOZ;
    int i = 1;
    while (i <= nz - 2 && (in >> ZZ >> ZT >> RHO >> THBAR >> PBAR >> PIBAR >> QVBAR >> Q1 >> Q2 >> RHOZ)) {
        v->thb[i] = THBAR;
        v->qvb[i] = QVBAR;
        v->pib[i] = PIBAR;
        v->pb[i] = PBAR;
        v->rhou[i] = RHOZ;
        v->rhow[i] = RHO;
        if (c->tropical_forcing && v->Q1LS && v->Q2LS) {
            v->Q1LS[i] = Q1 * 6.;
            v->Q2LS[i] = Q2 * 6.;
        }
        i++;
    }
    if (i <= nz - 2) return VVM_ERR_SIZE;
    bc1d(v->pib, nz);
    bc1d(v->pb, nz);
    bc1d(v->rhou, nz);
    bc1d(v->rhow, nz);
    v->rhow[nz - 1] = v->rhou[nz - 2];
    // the fields of the two outputs (src/Init.cpp:403-435)
    NcFile fm, fn;
    if (!fm.open(nc_prev) || !fn.open(nc_now)) return VVM_ERR_ARG;
    const size_t f2 = size_t(nx) * nz;
    struct {
        const NcFile *f;
        const char *var;
        double *dst;
        size_t n;
    } loads[] = {{&fm, "th", v->thm, f2}, {&fn, "th", v->th, f2},   {&fm, "zeta", v->zetam, f2},
                 {&fn, "zeta", v->zeta, f2}, {&fm, "qv", v->qvm, f2}, {&fn, "qv", v->qv, f2},
                 {&fm, "qc", v->qcm, f2},  {&fn, "qc", v->qc, f2},   {&fm, "qr", v->qrm, f2},
                 {&fn, "qr", v->qr, f2},   {&fn, "u", v->u, f2},     {&fn, "w", v->w, f2}};
    for (auto &l : loads) {
        if (!l.dst) return VVM_ERR_ARG;
        const std::vector<double> d = l.f->get(l.var);
        if (d.size() != l.n) return VVM_ERR_SIZE;
        std::memcpy(l.dst, d.data(), l.n * sizeof(double));
    }
    const std::vector<double> ub = fm.get("ubarTop");
    if (ub.size() != 1) return VVM_ERR_SIZE;
    v->ubarTopm = ub[0];
    if (c->CASE == 1 || c->CASE == 2) {
        for (int k = 1; k < nz - 1; k++)
            for (int ii = 1; ii < nx - 1; ii++) {
                at(v->thm, nz, ii, k) += get_th(ii, k, *c);
                at(v->th, nz, ii, k) += get_th(ii, k, *c);
            }
    }
    bc2d_center(v->thm, nx, nz);
    bc2d_center(v->th, nx, nz);
    // horizontal means (src/Init.cpp:448-484)
    for (int k = 1; k < nz - 1; k++) {
        double tb = 0., qb = 0.;
        for (int ii = 1; ii < nx - 1; ii++) {
            tb += at(v->th, nz, ii, k);
            qb += at(v->qv, nz, ii, k);
        }
        v->thb[k] = tb / double(nx - 2.);
        v->qvb[k] = qb / double(nx - 2.);
        v->thvb[k] = v->thb[k] * (1 + 0.61 * v->qvb[k]);
        v->qvsb[k] = (380. / v->pb[k]) *
                     std::exp((17.27 * (v->thb[k] * v->pib[k] - 273.)) / (v->thb[k] * v->pib[k] - 36.));
    }
    bc1d(v->thb, nz);
    bc1d(v->thvb, nz);
    bc1d(v->qvb, nz);
    bc1d(v->qvsb, nz);
    for (int k = 1; k < nz - 1; k++) v->thb_zeta[k] = 0.5 * (v->thb[k - 1] + v->thb[k]);
    bc1d(v->thb_zeta, nz);
    return VVM_OK;
}
