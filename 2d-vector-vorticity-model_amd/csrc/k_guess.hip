// k_guess.hip -- the learned initial guess of the cal_w solve and the
// recording of training samples.
//
// Features on the unknown grid (rows 1..M, N periodic columns), each scaled
// by its RMS over the grid:
//   0: r = f - L xe   (residual of the extrapolated guess xe: 3 (x1 - x2) + x3
//                      once three solutions exist, 2 x1 - x2 before)
//   1: d = x1 - x2    (last solution increment)
//   2: xe
// The network predicts (x - xe) / s_d with s_d = RMS(d); the guess is
// xe + s_d * prediction.  Everything before the network is FP64; the
// features enter the network as bf16.
#pragma clang fp contract(off)

#include <cmath>

#include "poisson.h"

namespace vvmd {

namespace {

constexpr int GX = 64, GY = 4;
inline dim3 ggrid(int N, int M) { return dim3((N + GX - 1) / GX, (M + GY - 1) / GY); }

__device__ inline double xe_at(const double *x1, const double *x2, const double *x3, int ord, size_t a) {
    return extrap_at(x1, x2, x3, ord, a);
}

// Partials: the three features' sums of squares, and ||wD r||^2 of the
// extrapolated guess (the solver's own norm, wD = rhow: the reference
// system's residual), against which the solve judges the network's guess.
__global__ void __launch_bounds__(GX *GY) k_guess_features(int N, int M, const double *__restrict__ f,
                                                           const double *__restrict__ x1,
                                                           const double *__restrict__ x2,
                                                           const double *__restrict__ x3, int ord,
                                                           const double *__restrict__ cx,
                                                           const double *__restrict__ cu,
                                                           const double *__restrict__ wD, float *__restrict__ feat,
                                                           RedSlot rs) {
    const int i = blockIdx.x * GX + threadIdx.x, j = 1 + blockIdx.y * GY + threadIdx.y;
    double part[4] = {0., 0., 0., 0.};
    if (i < N && j <= M) {
        const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
        const size_t r = size_t(j) * N;
        const double c = xe_at(x1, x2, x3, ord, r + i);
        const double Lx = cx[j] * (2. * c - xe_at(x1, x2, x3, ord, r + im) - xe_at(x1, x2, x3, ord, r + ip)) +
                          cu[j] * (c - xe_at(x1, x2, x3, ord, r + N + i)) + cu[j - 1] * (c - xe_at(x1, x2, x3, ord, r - N + i));
        const double res = f[r + i] - Lx, d = x1[r + i] - x2[r + i];
        const size_t plane = size_t(M) * N, q = size_t(j - 1) * N + i;
        feat[q] = float(res);
        feat[plane + q] = float(d);
        feat[2 * plane + q] = float(c);
        part[0] = res * res;
        part[1] = d * d;
        part[2] = c * c;
        const double wr = wD[j] * res;
        part[3] = wr * wr;
    }
    block_partials<4>(part, rs);
}

// sums of squares -> scales (0 for an all-zero feature)
__device__ inline double inv_rms(double s2, double n) { return s2 > 0. ? 1. / sqrt(s2 / n) : 0.; }

__global__ void k_guess_normalize(size_t plane, double npts, const double *__restrict__ gsc, float *__restrict__ feat,
                                  float *__restrict__ dst) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= 3 * plane) return;
    const int c = int(t / plane);
    dst[t] = float(double(feat[t]) * inv_rms(gsc[c], npts));
}

__global__ void __launch_bounds__(GX *GY) k_guess_target(int N, int M, const double *__restrict__ x,
                                                         const double *__restrict__ x1,
                                                         const double *__restrict__ x2,
                                                         const double *__restrict__ x3, int ord,
                                                         const double *__restrict__ gsc, float *__restrict__ tgt) {
    const int i = blockIdx.x * GX + threadIdx.x, j = 1 + blockIdx.y * GY + threadIdx.y;
    if (i >= N || j > M) return;
    const size_t a = size_t(j) * N + i;
    const double isd = inv_rms(gsc[1], double(N) * M);
    tgt[size_t(j - 1) * N + i] = float((x[a] - xe_at(x1, x2, x3, ord, a)) * isd);
}

}  // namespace

// features of the current (f, x_{t-1}, x_{t-2}), normalised, into dst [3][M][N]
void Poisson::guess_features(float *dst, hipStream_t st) {
    const size_t plane = size_t(M) * N;
    if (!gfeat) {
        hipMalloc(reinterpret_cast<void **>(&gfeat), sizeof(float) * 3 * plane);
        hipMalloc(reinterpret_cast<void **>(&gsc), sizeof(double) * 4);
    }
    if (!dst) dst = gfeat;  // in place: each element reads and writes its own slot
    const dim3 g = ggrid(N, M);
    // rows 2-5 of the reduction slots: row 1 may hold the pending ||b||^2
    // partials of the right-hand side
    const RedSlot rf{red.partial + 2 * size_t(red.max_blocks), red.max_blocks};
    hipLaunchKernelGGL(k_guess_features, g, dim3(GX, GY), 0, st, N, M, f, xprev, xprev2, xprev3, xe_order(), pcg64.mg.lev[0].cx,
                       pcg64.mg.lev[0].cu, pcg64.wD, gfeat, rf);
    finalize<4>(rf, nblocks(g), gsc, RED_STORE, 0, st);
    hipLaunchKernelGGL(k_guess_normalize, dim3(unsigned((3 * plane + 255) / 256)), dim3(256), 0, st, plane,
                       double(plane), gsc, gfeat, dst);
}

bool Poisson::learned_guess(hipStream_t st) {
    if (!net || nsolves < 2) return false;
    if (net_skip > 0) {
        net_skip--;
        guess_skipped++;
        return false;
    }
    const size_t plane = size_t(M) * N;
    if (!gfeat) {
        hipMalloc(reinterpret_cast<void **>(&gfeat), sizeof(float) * 3 * plane);
        hipMalloc(reinterpret_cast<void **>(&gsc), sizeof(double) * 4);
    }
    // raw features and their sums of squares; the network's first layer
    // scales them, its last writes the guess into x (rows 1..M; the
    // Dirichlet rows of x stay zero)
    const dim3 g = ggrid(N, M);
    // rows 2-5 of the reduction slots: row 1 may hold the pending ||b||^2
    // partials of the right-hand side
    const RedSlot rf{red.partial + 2 * size_t(red.max_blocks), red.max_blocks};
    hipLaunchKernelGGL(k_guess_features, g, dim3(GX, GY), 0, st, N, M, f, xprev, xprev2, xprev3, xe_order(), pcg64.mg.lev[0].cx,
                       pcg64.mg.lev[0].cu, pcg64.wD, gfeat, rf);
    // the extrapolated guess's residual norm goes to the host as well: the
    // solve keeps the network's guess only if it starts lower
    finalize<4>(rf, nblocks(g), gsc, RED_STORE, 0, st, -1, hscal_dev + H_GUESS);
    const double npts = double(plane);
    const NetGeo geo{1, M, N, 1};
    const NetFeatIn fin{gfeat, gsc, npts};
    const NetApply ap{xprev, xprev2, xprev3, gsc, pcg64.x, N, npts, xe_order()};
    guess_from_net = net->infer_guess(geo, fin, ap, gscratch, st) == VVM_OK;
    return guess_from_net;
}

int Poisson::collect_reserve(int cap) {
    if (samples) hipFree(samples);
    samples = nullptr;
    sample_cap = sample_count = 0;
    sample_open = false;
    if (cap <= 0) return VVM_OK;
    if (hipMalloc(reinterpret_cast<void **>(&samples), sizeof(float) * 4 * size_t(M) * N * cap) != hipSuccess) {
        samples = nullptr;
        return VVM_ERR_NOMEM;
    }
    sample_cap = cap;
    return VVM_OK;
}

void Poisson::collect_features(hipStream_t st) {
    sample_open = false;
    if (!samples || sample_count >= sample_cap || nsolves < 2) return;
    sample_ord = xe_order();
    guess_features(samples + size_t(sample_count) * 4 * M * N, st);
    sample_open = true;
}

void Poisson::collect_target(hipStream_t st) {
    if (!sample_open) return;
    float *tgt = samples + (size_t(sample_count) * 4 + 3) * M * N;
    hipLaunchKernelGGL(k_guess_target, ggrid(N, M), dim3(GX, GY), 0, st, N, M, pcg64.x, xprev, xprev2, xprev3, sample_ord, gsc, tgt);
    sample_count++;
    sample_open = false;
}

}  // namespace vvmd
