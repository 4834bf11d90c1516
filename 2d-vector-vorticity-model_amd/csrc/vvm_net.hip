// vvm_net.hip -- host side of the learned initial-guess network: buffers,
// forward / backward / Adam orchestration and the C ABI (vvm_net_*).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "net.h"
#include "vvm_net.h"

namespace vvmd {

namespace {
template <typename T>
T *dmalloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    return static_cast<T *>(p);
}
}  // namespace

int GuessNet::init(int device, hipStream_t stream) {
    st = stream;
    (void)device;
    params = dmalloc<float>(NET_PARAMS);
    grad = dmalloc<float>(NET_PARAMS);
    m = dmalloc<float>(NET_PARAMS);
    v = dmalloc<float>(NET_PARAMS);
    wf = dmalloc<__bf16>(size_t(NET_L) * NET_PACK);
    wb = dmalloc<__bf16>(size_t(NET_L) * NET_PACK);
    partial = dmalloc<float>(size_t(nblk_wgrad) * NET_WGRAD_PARTIAL);
    if (!params || !grad || !m || !v || !wf || !wb || !partial) return VVM_ERR_NOMEM;
    if (hipHostMalloc(reinterpret_cast<void **>(&hloss), sizeof(double) * 4096, hipHostMallocDefault) != hipSuccess)
        return VVM_ERR_NOMEM;
    hipMemsetAsync(params, 0, sizeof(float) * NET_PARAMS, st);
    hipMemsetAsync(grad, 0, sizeof(float) * NET_PARAMS, st);
    reset_adam();
    return VVM_OK;
}

void GuessNet::reset_adam() {
    hipMemsetAsync(m, 0, sizeof(float) * NET_PARAMS, st);
    hipMemsetAsync(v, 0, sizeof(float) * NET_PARAMS, st);
    adam_t = 0;
}

void GuessNet::destroy() {
    if (st) hipStreamSynchronize(st);
    void *ps[] = {params, grad, m, v, wf, wb, partial, lpart, act[0], act[1], act[2], act[3], act[4], act[5],
                  zpre[0], zpre[1], zpre[2], zpre[3], zpre[4], zpre[5], dz[0], dz[1]};
    for (void *p : ps)
        if (p) hipFree(p);
    if (hloss) hipHostFree(hloss);
    *this = GuessNet();
}

int GuessNet::reserve(long npx, bool training) {
    if (npx <= cap && (!training || cap_train)) return VVM_OK;
    const long n = std::max(npx, cap);
    for (int l = 0; l < NET_L; l++) {
        if (act[l]) hipFree(act[l]);
        if (zpre[l]) hipFree(zpre[l]);
        act[l] = zpre[l] = nullptr;
    }
    for (auto &d : dz) {
        if (d) hipFree(d);
        d = nullptr;
    }
    if (lpart) hipFree(lpart);
    const size_t e = size_t(n) * NET_C;
    // inference needs two activation buffers; training keeps every layer
    const int nact = training ? NET_L : 2;
    for (int l = 0; l < nact; l++)
        if (!(act[l] = dmalloc<__bf16>(e))) return VVM_ERR_NOMEM;
    if (training) {
        for (int l = 1; l < NET_L; l++)
            if (!(zpre[l] = dmalloc<__bf16>(e))) return VVM_ERR_NOMEM;
        for (auto &d : dz)
            if (!(d = dmalloc<__bf16>(e))) return VVM_ERR_NOMEM;
        zpre[0] = dmalloc<__bf16>(e);  // scratch Z of inference-shaped calls
    } else {
        zpre[0] = dmalloc<__bf16>(e);
        if (!zpre[0]) return VVM_ERR_NOMEM;
    }
    lpart = dmalloc<double>((n + 255) / 256 + 1);
    if (!lpart) return VVM_ERR_NOMEM;
    cap = n;
    cap_train = training;
    return VVM_OK;
}

void GuessNet::repack() { net_pack(params, wf, wb, st); }

int NetScratch::reserve(long npx) {
    if (npx <= cap) return VVM_OK;
    release();
    const size_t e = size_t(npx) * NET_C;
    a0 = dmalloc<__bf16>(e);
    a1 = dmalloc<__bf16>(e);
    z = dmalloc<__bf16>(e);
    if (!a0 || !a1 || !z) {
        release();
        return VVM_ERR_NOMEM;
    }
    cap = npx;
    return VVM_OK;
}

void NetScratch::release() {
    for (__bf16 *p : {a0, a1, z})
        if (p) hipFree(p);
    a0 = a1 = z = nullptr;
    cap = 0;
}

// Hidden layer l (and l + 1 with it when layer pairs are on, both are
// hidden layers and `pair_ok`) without keeping pre-activations; returns the
// number of layers done.
int GuessNet::hidden_step(const NetGeo &g, int l, const __bf16 *in, __bf16 *out, hipStream_t s,
                          bool pair_ok) const {
    const __bf16 *wl = wf + size_t(l) * NET_PACK;
    const float *bl = params + size_t(l) * NET_LAYER_PARAMS + NET_C * NET_C * NET_T;
    if (pair_ok && net_layer_pairs() && l + 1 < NET_L - 1) {
        const __bf16 *wn = wf + size_t(l + 1) * NET_PACK;
        const float *bn = params + size_t(l + 1) * NET_LAYER_PARAMS + NET_C * NET_C * NET_T;
        net_conv_pair(g, in, wl, bl, wn, bn, out, s);
        return 2;
    }
    net_conv_hidden(g, in, wl, bl, out, nullptr, s);
    return 1;
}

int GuessNet::infer_guess(const NetGeo &g, const NetFeatIn &fin, const NetApply &ap, NetScratch &sc,
                          hipStream_t s) const {
    const long npx = long(g.nb) * g.h * g.w;
    int rc = sc.reserve(npx);
    if (rc) return rc;
    __bf16 *buf[2] = {sc.a0, sc.a1};
    int cur = 1;  // buffer holding the latest activations
    for (int l = 0; l < NET_L;) {
        const __bf16 *wl = wf + size_t(l) * NET_PACK;
        const float *bl = params + size_t(l) * NET_LAYER_PARAMS + NET_C * NET_C * NET_T;
        if (l == 0) {
            net_conv_first_features(g, fin, wl, bl, buf[1], s);
            l++;
        } else if (l == NET_L - 1) {
            net_conv_last_apply(g, buf[cur], wl, bl, ap, s);
            l++;
        } else {
            l += hidden_step(g, l, buf[cur], buf[cur ^ 1], s);
            cur ^= 1;
        }
    }
    return hipGetLastError() == hipSuccess ? VVM_OK : VVM_ERR_HIP;
}

int GuessNet::forward(const NetGeo &g, const float *feat, float *out, bool keep, hipStream_t s) {
    const long npx = long(g.nb) * g.h * g.w;
    int rc = reserve(npx, keep);
    if (rc) return rc;
    const hipStream_t st = s ? s : this->st;
    net_pack_input(g, feat, act[0], st);
    if (!keep) {  // inference: activations alternate between act[0] and act[1]
        int cur = 0;
        for (int l = 0; l < NET_L - 1;) {
            l += hidden_step(g, l, act[cur], act[cur ^ 1], st, l > 0);
            cur ^= 1;
        }
        net_conv_last(g, act[cur], wf + size_t(NET_L - 1) * NET_PACK,
                      params + size_t(NET_L - 1) * NET_LAYER_PARAMS + NET_C * NET_C * NET_T, out, st);
        return hipGetLastError() == hipSuccess ? VVM_OK : VVM_ERR_HIP;
    }
    for (int l = 0; l < NET_L; l++) {
        const __bf16 *wl = wf + size_t(l) * NET_PACK;
        const float *bl = params + size_t(l) * NET_LAYER_PARAMS + NET_C * NET_C * NET_T;
        if (l == NET_L - 1)
            net_conv_last(g, act[l], wl, bl, out, st);
        else
            net_conv_hidden(g, act[l], wl, bl, act[l + 1], zpre[l + 1], st);
    }
    return hipGetLastError() == hipSuccess ? VVM_OK : VVM_ERR_HIP;
}

int GuessNet::backward(const NetGeo &g, const float *out, const float *target, const float *mask, double msum,
                       double *loss) {
    const long npx = long(g.nb) * g.h * g.w;
    if (!cap_train || npx > cap) return VVM_ERR_STATE;
    const int nlb = int((npx + 255) / 256);
    net_loss(npx, out, target, mask, float(1.0 / msum), dz[0], lpart, st);
    int cur = 0;
    for (int l = NET_L - 1; l >= 0; l--) {
        // layer l+1 of the network: input act[l], pre-activation gradient dz[cur]
        net_wgrad(g, act[l], dz[cur], partial, nblk_wgrad, grad, l, st);
        if (l == 0) break;
        net_conv_bwd(g, dz[cur], wb + size_t(l) * NET_PACK, zpre[l], dz[cur ^ 1], st);
        cur ^= 1;
    }
    if (loss) {
        if (nlb > 4096) return VVM_ERR_SIZE;
        hipMemcpyAsync(hloss, lpart, sizeof(double) * nlb, hipMemcpyDeviceToHost, st);
        if (hipStreamSynchronize(st) != hipSuccess) return VVM_ERR_HIP;
        double s = 0.;
        for (int i = 0; i < nlb; i++) s += hloss[i];
        *loss = s / msum;
    }
    return hipGetLastError() == hipSuccess ? VVM_OK : VVM_ERR_HIP;
}

void GuessNet::adam_step(float lr, float b1, float b2, float eps) {
    adam_t++;
    const float c1 = 1.f - std::pow(b1, float(adam_t)), c2 = 1.f - std::pow(b2, float(adam_t));
    net_adam(params, m, v, grad, NET_PARAMS, lr, b1, b2, eps, c1, c2, st);
    repack();
}

}  // namespace vvmd

// ------------------------------------------------------------------ C ABI

using vvmd::GuessNet;
using vvmd::NetGeo;

struct vvm_net {
    int device = 0;
    hipStream_t st = nullptr;
    GuessNet net;
    std::string err;
    float *dfeat = nullptr, *dout = nullptr, *dtgt = nullptr, *dmask = nullptr;
    size_t dcap = 0;
    struct Pool {
        float *buf;  // [count][4][M][N]
        int N, M, count;
    };
    std::vector<Pool> pool;
};

namespace vvmd {
GuessNet *guess_net_of(vvm_net_t *n) { return n ? &n->net : nullptr; }
int guess_net_device(vvm_net_t *n) { return n ? n->device : -1; }
}  // namespace vvmd

namespace {
int nfail(vvm_net *n, int rc, const std::string &msg) {
    if (n) n->err = msg;
    return rc;
}

int stage(vvm_net *n, long npx) {
    if (size_t(npx) <= n->dcap) return VVM_OK;
    hipFree(n->dfeat), hipFree(n->dout), hipFree(n->dtgt), hipFree(n->dmask);
    n->dfeat = n->dout = n->dtgt = n->dmask = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&n->dfeat), sizeof(float) * 3 * npx) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&n->dout), sizeof(float) * npx) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&n->dtgt), sizeof(float) * npx) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&n->dmask), sizeof(float) * npx) != hipSuccess)
        return nfail(n, VVM_ERR_NOMEM, "staging allocation failed");
    n->dcap = size_t(npx);
    return VVM_OK;
}

bool geo_ok(int nb, int h, int w) { return nb > 0 && h > 0 && w > 0 && long(nb) * h * w < (1L << 31); }
}  // namespace

extern "C" {

int vvm_net_create(int device, vvm_net_t **out) {
    if (!out) return VVM_ERR_ARG;
    *out = nullptr;
    auto *n = new vvm_net();
    n->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return VVM_ERR_HIP;
    }
    int rc = n->net.init(device, n->st);
    if (rc) {
        n->net.destroy();
        hipStreamDestroy(n->st);
        delete n;
        return rc;
    }
    *out = n;
    return VVM_OK;
}

void vvm_net_destroy(vvm_net_t *n) {
    if (!n) return;
    hipSetDevice(n->device);
    n->net.destroy();
    for (auto &p : n->pool) hipFree(p.buf);
    hipFree(n->dfeat), hipFree(n->dout), hipFree(n->dtgt), hipFree(n->dmask);
    if (n->st) hipStreamDestroy(n->st);
    delete n;
}

const char *vvm_net_last_error(const vvm_net_t *n) { return n ? n->err.c_str() : "null handle"; }

int vvm_net_param_count(void) { return vvmd::NET_PARAMS; }

int vvm_net_set_params(vvm_net_t *n, const float *host) {
    if (!n || !host) return VVM_ERR_ARG;
    hipSetDevice(n->device);
    if (hipMemcpyAsync(n->net.params, host, sizeof(float) * vvmd::NET_PARAMS, hipMemcpyHostToDevice, n->st) !=
        hipSuccess)
        return nfail(n, VVM_ERR_HIP, "parameter upload failed");
    n->net.repack();
    n->net.reset_adam();
    return hipStreamSynchronize(n->st) == hipSuccess ? VVM_OK : nfail(n, VVM_ERR_HIP, "sync failed");
}

int vvm_net_get_params(vvm_net_t *n, float *host) {
    if (!n || !host) return VVM_ERR_ARG;
    hipSetDevice(n->device);
    hipMemcpyAsync(host, n->net.params, sizeof(float) * vvmd::NET_PARAMS, hipMemcpyDeviceToHost, n->st);
    return hipStreamSynchronize(n->st) == hipSuccess ? VVM_OK : nfail(n, VVM_ERR_HIP, "sync failed");
}

int vvm_net_get_grad(vvm_net_t *n, float *host) {
    if (!n || !host) return VVM_ERR_ARG;
    hipSetDevice(n->device);
    hipMemcpyAsync(host, n->net.grad, sizeof(float) * vvmd::NET_PARAMS, hipMemcpyDeviceToHost, n->st);
    return hipStreamSynchronize(n->st) == hipSuccess ? VVM_OK : nfail(n, VVM_ERR_HIP, "sync failed");
}

// He initialisation with frozen entries (layer 1 inputs >= 3, layer 6
// outputs >= 1) at zero; zero biases.
int vvm_net_init_params(vvm_net_t *n, uint64_t seed) {
    if (!n) return VVM_ERR_ARG;
    std::vector<float> p(vvmd::NET_PARAMS, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < vvmd::NET_L; l++) {
        const int cin = (l == 0) ? vvmd::NET_CIN : vvmd::NET_C;
        const int cout = (l == vvmd::NET_L - 1) ? 1 : vvmd::NET_C;
        // the last layer starts small so that the first guesses stay near
        // the extrapolation
        const float sd = std::sqrt(2.f / float(cin * vvmd::NET_T)) * (l == vvmd::NET_L - 1 ? 0.1f : 1.f);
        std::normal_distribution<float> nd(0.f, sd);
        float *W = p.data() + size_t(l) * vvmd::NET_LAYER_PARAMS;
        for (int co = 0; co < cout; co++)
            for (int ci = 0; ci < cin; ci++)
                for (int t = 0; t < vvmd::NET_T; t++) W[(co * vvmd::NET_C + ci) * vvmd::NET_T + t] = nd(rng);
    }
    return vvm_net_set_params(n, p.data());
}

int vvm_net_forward_host(vvm_net_t *n, int nb, int h, int w, int periodic, const float *input, float *out) {
    if (!n || !input || !out || !geo_ok(nb, h, w)) return VVM_ERR_ARG;
    hipSetDevice(n->device);
    const long npx = long(nb) * h * w;
    int rc = stage(n, npx);
    if (rc) return rc;
    hipMemcpyAsync(n->dfeat, input, sizeof(float) * 3 * npx, hipMemcpyHostToDevice, n->st);
    rc = n->net.forward(NetGeo{nb, h, w, periodic}, n->dfeat, n->dout, false);
    if (rc) return nfail(n, rc, "forward failed");
    hipMemcpyAsync(out, n->dout, sizeof(float) * npx, hipMemcpyDeviceToHost, n->st);
    return hipStreamSynchronize(n->st) == hipSuccess ? VVM_OK : nfail(n, VVM_ERR_HIP, "sync failed");
}

int vvm_net_grad_host(vvm_net_t *n, int nb, int h, int w, int periodic, const float *input, const float *target,
                      const float *mask, float *grad, double *loss) {
    if (!n || !input || !target || !mask || !geo_ok(nb, h, w)) return VVM_ERR_ARG;
    hipSetDevice(n->device);
    const long npx = long(nb) * h * w;
    int rc = stage(n, npx);
    if (rc) return rc;
    double msum = 0.;
    for (long i = 0; i < npx; i++) msum += mask[i];
    if (msum <= 0.) msum = 1.;
    hipMemcpyAsync(n->dfeat, input, sizeof(float) * 3 * npx, hipMemcpyHostToDevice, n->st);
    hipMemcpyAsync(n->dtgt, target, sizeof(float) * npx, hipMemcpyHostToDevice, n->st);
    hipMemcpyAsync(n->dmask, mask, sizeof(float) * npx, hipMemcpyHostToDevice, n->st);
    const NetGeo g{nb, h, w, periodic};
    rc = n->net.forward(g, n->dfeat, n->dout, true);
    if (rc) return nfail(n, rc, "forward failed");
    double l = 0.;
    rc = n->net.backward(g, n->dout, n->dtgt, n->dmask, msum, &l);
    if (rc) return nfail(n, rc, "backward failed");
    if (loss) *loss = l;
    if (grad) return vvm_net_get_grad(n, grad);
    return VVM_OK;
}

int vvm_net_adam(vvm_net_t *n, float lr) {
    if (!n) return VVM_ERR_ARG;
    hipSetDevice(n->device);
    n->net.adam_step(lr);
    return hipStreamSynchronize(n->st) == hipSuccess ? VVM_OK : nfail(n, VVM_ERR_HIP, "sync failed");
}

int vvm_net_set_conv_rows(int rows) {
    if (rows != 2 && rows != 4 && rows != 8) return VVM_ERR_ARG;
    vvmd::net_set_conv_rows(rows);
    return VVM_OK;
}

int vvm_net_set_layer_pairs(int on) {
    if (on != 0 && on != 1) return VVM_ERR_ARG;
    vvmd::net_set_layer_pairs(on != 0);
    return VVM_OK;
}

int vvm_net_add_samples(vvm_net_t *n, vvm_dev_t *h) {
    if (!n || !h) return VVM_ERR_ARG;
    const float *buf = nullptr;
    int N = 0, M = 0, count = 0, device = -1;
    int rc = vvmd::dev_samples(h, &buf, &N, &M, &count, &device);
    if (rc) return nfail(n, rc, "the timestep handle has no solver state");
    if (device != n->device) return nfail(n, VVM_ERR_ARG, "samples live on another device");
    if (count == 0) return VVM_OK;
    hipSetDevice(n->device);
    const size_t bytes = sizeof(float) * 4 * size_t(M) * N * count;
    float *p = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&p), bytes) != hipSuccess) return nfail(n, VVM_ERR_NOMEM, "pool full");
    hipMemcpyAsync(p, buf, bytes, hipMemcpyDeviceToDevice, n->st);
    if (hipStreamSynchronize(n->st) != hipSuccess) return nfail(n, VVM_ERR_HIP, "sample copy failed");
    n->pool.push_back({p, N, M, count});
    return VVM_OK;
}

int vvm_net_pool_size(vvm_net_t *n) {
    int s = 0;
    if (n)
        for (auto &p : n->pool) s += p.count;
    return s;
}

int vvm_net_train(vvm_net_t *n, const vvm_net_train_opts *o, vvm_net_train_stats *s) {
    using vvmd::CropDesc;
    if (!n || !o || o->steps < 0 || o->batch <= 0 || o->crop_h <= 0 || o->crop_w <= 0 || o->margin < 0 ||
        2 * o->margin >= o->crop_w)
        return VVM_ERR_ARG;
    int total = 0;
    for (auto &p : n->pool)
        if (p.M >= o->crop_h && p.N >= o->crop_w) total += p.count;
    if (total == 0) return nfail(n, VVM_ERR_STATE, "no recorded sample is large enough for the crop");
    hipSetDevice(n->device);
    const int B = o->batch, ch = o->crop_h, cw = o->crop_w;
    const long npx = long(B) * ch * cw;
    int rc = stage(n, npx);
    if (rc) return rc;
    CropDesc *dd = nullptr, *hd = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&dd), sizeof(CropDesc) * B * 2) != hipSuccess)
        return nfail(n, VVM_ERR_NOMEM, "descriptor allocation failed");
    if (hipHostMalloc(reinterpret_cast<void **>(&hd), sizeof(CropDesc) * B * 2, hipHostMallocDefault) != hipSuccess) {
        hipFree(dd);
        return nfail(n, VVM_ERR_NOMEM, "descriptor allocation failed");
    }
    hipEvent_t used[2];
    hipEventCreateWithFlags(&used[0], hipEventDisableTiming);
    hipEventCreateWithFlags(&used[1], hipEventDisableTiming);
    hipEventRecord(used[0], n->st);
    hipEventRecord(used[1], n->st);
    std::mt19937_64 rng(o->seed);
    const NetGeo g{B, ch, cw, 0};
    double lsum = 0., first = -1., last = 0.;
    int lcount = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < o->steps && rc == VVM_OK; it++) {
        const int slot = it & 1;
        hipEventSynchronize(used[slot]);  // the gather that read this slot is done
        CropDesc *h = hd + slot * B;
        double msum = 0.;
        for (int b = 0; b < B; b++) {
            int pick = int(rng() % uint64_t(total));
            const vvm_net::Pool *P = nullptr;
            for (auto &p : n->pool) {
                if (p.M < ch || p.N < cw) continue;
                if (pick < p.count) {
                    P = &p;
                    break;
                }
                pick -= p.count;
            }
            const int y0 = int(rng() % uint64_t(P->M - ch + 1)), x0 = int(rng() % uint64_t(P->N));
            h[b] = {P->buf + size_t(pick) * 4 * P->M * P->N, P->N, P->M, y0, x0};
            const int rows = ch - (y0 > 0 ? o->margin : 0) - (y0 + ch < P->M ? o->margin : 0);
            msum += double(std::max(rows, 0)) * (cw - 2 * o->margin);
        }
        hipMemcpyAsync(dd + slot * B, h, sizeof(CropDesc) * B, hipMemcpyHostToDevice, n->st);
        vvmd::net_gather(B, ch, cw, o->margin, dd + slot * B, n->dfeat, n->dtgt, n->dmask, n->st);
        hipEventRecord(used[slot], n->st);
        rc = n->net.forward(g, n->dfeat, n->dout, true);
        if (rc) break;
        const bool log = o->log_every > 0 && (it % o->log_every == 0 || it == o->steps - 1);
        double l = 0.;
        rc = n->net.backward(g, n->dout, n->dtgt, n->dmask, std::max(msum, 1.), log ? &l : nullptr);
        if (rc) break;
        const double pi = 3.14159265358979323846;
        const float lr = o->lr_final + 0.5f * (o->lr - o->lr_final) *
                                           float(1. + std::cos(pi * double(it) / std::max(1, o->steps - 1)));
        n->net.adam_step(lr);
        if (log) {
            if (first < 0.) first = l;
            last = l;
            lsum += l;
            lcount++;
        }
    }
    hipStreamSynchronize(n->st);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    hipEventDestroy(used[0]);
    hipEventDestroy(used[1]);
    hipFree(dd);
    hipHostFree(hd);
    if (rc) return nfail(n, rc, "training step failed");
    if (s) {
        s->loss_first = first;
        s->loss_last = last;
        s->seconds = secs;
        s->steps = o->steps;
    }
    (void)lsum;
    return hipGetLastError() == hipSuccess ? VVM_OK : nfail(n, VVM_ERR_HIP, "training kernels failed");
}

int vvm_net_eval(vvm_net_t *n, vvm_dev_t *dev, double *err) {
    if (!n || !dev || !err) return VVM_ERR_ARG;
    const float *buf = nullptr;
    int N = 0, M = 0, count = 0, device = -1;
    int rc = vvmd::dev_samples(dev, &buf, &N, &M, &count, &device);
    if (rc) return nfail(n, rc, "the timestep handle has no solver state");
    if (device != n->device) return nfail(n, VVM_ERR_ARG, "samples live on another device");
    if (count == 0) return nfail(n, VVM_ERR_STATE, "no recorded samples");
    hipSetDevice(n->device);
    const long npx = long(M) * N;
    rc = stage(n, npx);
    if (rc) return rc;
    std::vector<float> out(npx), tgt(npx);
    double e2 = 0., t2 = 0.;
    for (int k = 0; k < count; k++) {
        const float *smp = buf + size_t(k) * 4 * npx;
        rc = n->net.forward(NetGeo{1, M, N, 1}, smp, n->dout, false);
        if (rc) return nfail(n, rc, "forward failed");
        hipMemcpyAsync(out.data(), n->dout, sizeof(float) * npx, hipMemcpyDeviceToHost, n->st);
        hipMemcpyAsync(tgt.data(), smp + 3 * npx, sizeof(float) * npx, hipMemcpyDeviceToHost, n->st);
        if (hipStreamSynchronize(n->st) != hipSuccess) return nfail(n, VVM_ERR_HIP, "sync failed");
        for (long i = 0; i < npx; i++) {
            const double d = double(out[i]) - tgt[i];
            e2 += d * d;
            t2 += double(tgt[i]) * tgt[i];
        }
    }
    *err = t2 > 0. ? e2 / t2 : 0.;
    return VVM_OK;
}

int vvm_net_save(vvm_net_t *n, const char *path) {
    if (!n || !path) return VVM_ERR_ARG;
    std::vector<float> p(vvmd::NET_PARAMS);
    int rc = vvm_net_get_params(n, p.data());
    if (rc) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return nfail(n, VVM_ERR_ARG, std::string("cannot write ") + path);
    const int32_t cnt = vvmd::NET_PARAMS;
    bool ok = std::fwrite("VVMNET01", 1, 8, f) == 8 && std::fwrite(&cnt, 4, 1, f) == 1 &&
              std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? VVM_OK : nfail(n, VVM_ERR_ARG, std::string("write error on ") + path);
}

int vvm_net_load(vvm_net_t *n, const char *path) {
    if (!n || !path) return VVM_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return nfail(n, VVM_ERR_ARG, std::string("cannot read ") + path);
    char magic[8];
    int32_t cnt = 0;
    std::vector<float> p(vvmd::NET_PARAMS);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "VVMNET01", 8) == 0 &&
              std::fread(&cnt, 4, 1, f) == 1 && cnt == vvmd::NET_PARAMS &&
              std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
    std::fclose(f);
    if (!ok) return nfail(n, VVM_ERR_ARG, std::string("not a network file: ") + path);
    return vvm_net_set_params(n, p.data());
}

}  // extern "C"
