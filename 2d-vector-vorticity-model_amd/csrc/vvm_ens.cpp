// vvm_ens.cpp -- the batched ensemble (include/vvm_ens.h): member handles,
// host worker threads, shape tuning shared by members of equal grids.
#include "vvm_ens.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace vvmd {
// (vvm_dev.hip) members run their microphysics phase on their main stream
void dev_set_phase_overlap(vvm_dev_t *h, bool on);
}  // namespace vvmd

struct vvm_ens {
    int device = 0, threads = 1;
    std::vector<vvm_dev_t *> dev;
    std::vector<long> points;           // interior points per member
    std::vector<std::vector<int>> work;  // members per worker thread
    std::string err;
};

namespace {

int efail(vvm_ens *e, int rc, const std::string &msg) {
    if (e) e->err = msg;
    return rc;
}

// host buffers of one member's initial state
struct HostState {
    std::vector<std::vector<double>> store;
    vvm_host_view v{};
    HostState(int nx, int nz) {
        const size_t f2 = size_t(nx) * nz;
        auto buf = [&](size_t n) {
            store.emplace_back(n, 0.);
            return store.back().data();
        };
        store.reserve(64);
        double **fields[] = {&v.zeta, &v.zetam, &v.zetap, &v.th, &v.thm, &v.thp, &v.qv, &v.qvm, &v.qvp,
                             &v.qc, &v.qcm, &v.qcp, &v.qr, &v.qrm, &v.qrp, &v.u, &v.w, &v.RKM, &v.RKH,
                             &v.init_th_forcing, &v.evaporation, &v.accretion, &v.autoconversion, &v.condensation};
        for (double **p : fields) *p = buf(f2);
        double **hists[] = {&v.dth_advect, &v.dth_buoyancy, &v.dzeta_advect, &v.dqv_advect, &v.dqc_advect,
                            &v.dqr_advect, &v.dqr_VT};
        for (double **p : hists) *p = buf(2 * f2);
        double **rows[] = {&v.precip, &v.xi, &v.uxi};
        for (double **p : rows) *p = buf(size_t(nx));
        double **profs[] = {&v.rhou, &v.rhow, &v.pib, &v.pb, &v.thb, &v.thbm, &v.thb_init, &v.thb_zeta, &v.thvb,
                            &v.thvbm, &v.qvb, &v.qvb0, &v.qvsb, &v.z, &v.z_zeta, &v.lambda2, &v.Q1LS, &v.Q2LS};
        for (double **p : profs) *p = buf(size_t(nz));
    }
};

}  // namespace

extern "C" {

int vvm_ens_create(const vvm_ens_member *m, int n, int device, int threads, vvm_ens_t **out) {
    if (!m || n <= 0 || !out || threads < 0) return VVM_ERR_ARG;
    *out = nullptr;
    auto *e = new vvm_ens();
    e->device = device;
    e->threads = threads > 0 ? std::min(threads, n) : std::min(n, 8);
    for (int i = 0; i < n; i++) {
        HostState hs(m[i].cfg.nx, m[i].cfg.nz);
        int rc = vvm_init_case(&m[i].cfg, &m[i].params, &hs.v);
        vvm_dev_t *h = nullptr;
        if (rc == VVM_OK) rc = vvm_dev_create(&m[i].cfg, device, &h);
        if (rc == VVM_OK) vvmd::dev_set_phase_overlap(h, false);
        if (rc == VVM_OK) rc = vvm_dev_upload(h, &hs.v);
        if (rc < 0) {
            const std::string msg = "member " + std::to_string(i) + ": " + (h ? vvm_dev_last_error(h) : "create");
            if (h) vvm_dev_destroy(h);
            vvm_ens_destroy(e);
            (void)msg;
            return rc;
        }
        e->dev.push_back(h);
        e->points.push_back(long(m[i].cfg.nx - 2) * (m[i].cfg.nz - 2));
    }
    // largest first onto the least-loaded worker
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return e->points[a] > e->points[b]; });
    e->work.assign(e->threads, {});
    std::vector<long> load(e->threads, 0);
    for (int i : order) {
        const int w = int(std::min_element(load.begin(), load.end()) - load.begin());
        e->work[w].push_back(i);
        load[w] += e->points[i];
    }
    *out = e;
    return VVM_OK;
}

void vvm_ens_destroy(vvm_ens_t *e) {
    if (!e) return;
    for (vvm_dev_t *h : e->dev) vvm_dev_destroy(h);
    delete e;
}

const char *vvm_ens_last_error(const vvm_ens_t *e) { return e ? e->err.c_str() : "null handle"; }

int vvm_ens_size(const vvm_ens_t *e) { return e ? int(e->dev.size()) : 0; }

vvm_dev_t *vvm_ens_handle(vvm_ens_t *e, int i) {
    return (e && i >= 0 && i < int(e->dev.size())) ? e->dev[i] : nullptr;
}

int vvm_ens_set_guess_net(vvm_ens_t *e, vvm_net_t *net) {
    if (!e) return VVM_ERR_ARG;
    for (size_t i = 0; i < e->dev.size(); i++) {
        const int rc = vvm_dev_set_guess_net(e->dev[i], net);
        if (rc < 0) return efail(e, rc, "member " + std::to_string(i) + ": " + vvm_dev_last_error(e->dev[i]));
    }
    return VVM_OK;
}

int vvm_ens_tune(vvm_ens_t *e, int steps) {
    if (!e || steps <= 0) return VVM_ERR_ARG;
    // vvm_dev_tune caches per shape, so the first member of a shape tunes
    // and the others adopt the result
    for (size_t i = 0; i < e->dev.size(); i++) {
        const int rc = vvm_dev_tune(e->dev[i], steps, nullptr);
        if (rc < 0) return efail(e, rc, "member " + std::to_string(i) + ": " + vvm_dev_last_error(e->dev[i]));
    }
    return VVM_OK;
}

int vvm_ens_step(vvm_ens_t *e, int nsteps, vvm_ens_stats *stats) {
    if (!e || nsteps < 0) return VVM_ERR_ARG;
    std::atomic<long long> iters{0};
    std::atomic<int> notconv{0}, first_err{0};
    std::mutex emu;
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&](int w) {
        // the worker's members advance in lock step, so that every member's
        // stream has work queued while another member waits on its solver
        for (int s = 0; s < nsteps && first_err.load() == 0; s++)
            for (int i : e->work[w]) {
                vvm_step_stats st{};
                const int rc = vvm_dev_step(e->dev[i], 1, &st);
                if (rc < 0) {
                    std::lock_guard<std::mutex> lk(emu);
                    if (first_err.load() == 0) {
                        first_err = rc;
                        e->err = "member " + std::to_string(i) + ": " + vvm_dev_last_error(e->dev[i]);
                    }
                    return;
                }
                iters += st.solve_iters_total;
                notconv += st.not_converged;
            }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < e->threads; w++) pool.emplace_back(worker, w);
    worker(0);
    for (auto &t : pool) t.join();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        double pts = 0.;
        for (long p : e->points) pts += double(p) * nsteps;
        stats->seconds = secs;
        stats->grid_point_steps = pts;
        stats->solve_iters_total = iters.load();
        stats->not_converged = notconv.load();
        stats->members = int(e->dev.size());
    }
    return first_err.load();
}

}  // extern "C"
