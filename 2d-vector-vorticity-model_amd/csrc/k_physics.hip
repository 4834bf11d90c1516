// k_physics.hip -- Kessler microphysics, first-order turbulence closure and
// K-diffusion, nudging and horizontal means of one VVM timestep (phases C-E
// of vvm::Iteration::TimeMarching, src/Iteration.cpp:300-434).
//
// FP64, no FMA contraction (reference rounding, see k_dynamics.hip).
// std::max(a,b) / std::min(a,b) keep their NaN behaviour: (a<b)?b:a and
// (b<a)?b:a.
#pragma clang fp contract(off)

#include "vvm_internal.h"

#include <cstdio>
#include <mutex>

namespace vvmd {

namespace {

constexpr int BX = 64, BY = 4;

__device__ inline double maxs(double a, double b) { return (a < b) ? b : a; }
__device__ inline double mins(double a, double b) { return (b < a) ? b : a; }

inline dim3 grid2d(const Grid &G) { return dim3((G.nx + BX - 1) / BX, (G.nz + BY - 1) / BY); }

// MicroPhysics::autoconversion, accretion, evaporation, condensation
// (MicroPhysics.cpp:29-51, 54-77, 80-114, 7-26).  Each is pointwise, so
// running the four in call order per point equals the reference's four
// sweeps.  Then the first pass of NegativeValueProcess on qvp, qcp, qrp
// (src/Iteration.cpp:309-311); k_finalize computes the three ratios.
// PROC selects the processes (KP_*); the timestep runs them all with the
// negative-value pass (KP_ALL).  One process alone (the reference's
// MicroPhysics statics) writes only the fields and the rate it changes.
enum { KP_AUTO = 1, KP_ACCR = 2, KP_EVAP = 4, KP_COND = 8, KP_NEG = 16, KP_ALL = 31 };

template <int PROC>
__global__ void __launch_bounds__(BX *BY) k_kessler(Grid G, StepPtrs S, RedSlot rs) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    const bool inside = (i >= 1 && i <= G.nx - 2 && k >= 1 && k <= G.nz - 2);
    double part[6] = {0., 0., 0., 0., 0., 0.};
    if (inside) {
        const size_t p = ix(G, i, k);
        const double dt = G.dt, pib = S.P.pib[k], rhou = S.P.rhou[k], pb = S.P.pb[k], pc = S.P.pcoef[k];
        double qc = S.fut[F_QC][p], qr = S.fut[F_QR][p], qv = S.fut[F_QV][p], th = S.fut[F_TH][p];
        if constexpr ((PROC & KP_AUTO) != 0) {  // autoconversion
            const double qcplus = maxs(0., qc);
            double ar = 0.001 * (qcplus - 0.001);
            ar = maxs(0., ar);
            const double arcrdt = mins(ar * dt, qcplus);
            qc = qc - arcrdt;
            qr = qr + arcrdt;
            S.autoc[p] = arcrdt;
        }
        if constexpr ((PROC & KP_ACCR) != 0) {  // accretion
            const double qcplus = maxs(0., qc), qrplus = maxs(0., qr);
            const double cr = 2.2 * qcplus * (pow(qrplus, 0.875));
            const double arcrdt = mins(cr * dt, qcplus);
            qc = qc - arcrdt;
            qr = qr + arcrdt;
            S.accr[p] = arcrdt;
        }
        if constexpr ((PROC & KP_EVAP) != 0) {  // rain evaporation
            const double qrplus = maxs(0., qr), qvplus = maxs(0., qv);
            const double qvs = pc * exp(17.27 * (pib * th - 273.) / (pib * th - 36.));
            const double deficit = maxs((1. - qvplus / qvs), 0.);
            const double coef = 1.6 + 124.9 * pow((1E-3 * rhou * qrplus), 0.2046);
            const double er = coef * deficit * (pow(1E-3 * rhou * qrplus, 0.525)) /
                              ((5.4E5 + 2.55E6 / (1E-2 * pb * qvs)) * 1E-3 * rhou);
            const double erdt = mins(qrplus, maxs(0., er * dt));
            qr = qr - erdt;
            qv = qv + erdt;
            th = th - G.Lv * erdt / (G.Cp * pib);
            S.evap[p] = erdt;
        }
        if constexpr ((PROC & KP_COND) != 0) {  // saturation adjustment
            const double qvs = pc * exp(17.27 * (pib * th - 273.) / (pib * th - 36.));
            const double tm = th * pib - 36.;
            const double phi = qvs * (17.27 * 237. * G.Lv) / (G.Cp * (tm * tm));
            double C = (qv - qvs) / (1 + phi);
            if (fabs(C) > qc && C < 0) C = -qc;
            qv = qv - C;
            qc = qc + C;
            th = th + G.Lv / (G.Cp * pib) * C;
            S.cond[p] = C;
        }
        double q3[3] = {qv, qc, qr};
        if constexpr ((PROC & KP_NEG) != 0) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                if (q3[q] >= 0.) part[2 * q] += q3[q];
                else {
                    part[2 * q + 1] += q3[q];
                    q3[q] = 0.;
                }
            }
        }
        constexpr bool wv = (PROC & (KP_EVAP | KP_COND | KP_NEG)) != 0;
        constexpr bool wc = (PROC & (KP_AUTO | KP_ACCR | KP_COND | KP_NEG)) != 0;
        constexpr bool wr = (PROC & (KP_AUTO | KP_ACCR | KP_EVAP | KP_NEG)) != 0;
        constexpr bool wt = (PROC & (KP_EVAP | KP_COND)) != 0;
        if (wv) S.fut[F_QV][p] = q3[0];
        if (wc) S.fut[F_QC][p] = q3[1];
        if (wr) S.fut[F_QR][p] = q3[2];
        if (wt) S.fut[F_TH][p] = th;
    }
    if constexpr ((PROC & KP_NEG) != 0) block_partials<6>(part, rs);
}

// MicroPhysics::NegativeValueProcess (MicroPhysics.cpp:116-138) on one
// field: the first pass zeroes the negatives and sums positives and
// negatives (k_finalize forms the ratio), the second rescales.
__global__ void __launch_bounds__(BX *BY) k_negval_sums(Grid G, double *a, RedSlot rs) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    double part[2] = {0., 0.};
    if (i >= 1 && i <= G.nx - 2 && k >= 1 && k <= G.nz - 2) {
        const size_t p = ix(G, i, k);
        const double v = a[p];
        if (v >= 0.) part[0] += v;
        else {
            part[1] += v;
            a[p] = 0.;
        }
    }
    block_partials<2>(part, rs);
}
__global__ void __launch_bounds__(BX *BY) k_negval_scale(Grid G, double *a, const double *ratio) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const double r = *ratio;
    const size_t p = ix(G, i, k);
    if (r >= 0. && a[p] > 0) a[p] = a[p] * r;
}

// Turbulence::RKM_RKH coefficient part (Turbulent.cpp:13-45), with the
// center halos of RKM and RKH.
__global__ void __launch_bounds__(BX *BY) k_turb_coeffs(Grid G, StepPtrs S) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const double *u = S.u, *w = S.w, *thp = S.fut[F_TH];
    const double Rz =
        0.5 * ((w[ix(G, i + 1, k + 1)] + w[ix(G, i + 1, k)]) - (w[ix(G, i - 1, k + 1)] + w[ix(G, i - 1, k)])) * G.r2dx +
        0.5 * ((u[ix(G, i + 1, k + 1)] + u[ix(G, i, k + 1)]) - (u[ix(G, i + 1, k - 1)] + u[ix(G, i, k - 1)])) * G.r2dz;
    const double dv = (u[ix(G, i + 1, k)] - u[ix(G, i, k)]) * G.rdx + (w[ix(G, i, k + 1)] - w[ix(G, i, k)]) * G.rdz;
    const double Ro = dv * dv;
    const double Ri = (G.GRAVITY / thp[ix(G, i, k)] * (thp[ix(G, i, k + 1)] - thp[ix(G, i, k - 1)]) * G.r2dz) /
                      (Rz * Rz + 2. * Ro);
    const double l2 = S.P.lambda2[k];
    double km, kh;
    if (Ri < 0) {
        km = l2 * sqrt(Rz * Rz + 2. * Ro) * sqrt(1. - 16. * Ri);
        kh = l2 * sqrt(Rz * Rz + 2. * Ro) * 1.4 * sqrt(1. - 40. * Ri);
    } else if (0 < Ri && Ri < 0.25) {
        km = l2 * sqrt(Rz * Rz + 2. * Ro) * pow(1. - 4. * Ri, 4.);
        kh = l2 * sqrt(Rz * Rz + 2. * Ro) * 1.4 * (1. - 1.2 * Ri) * pow(1. - 4. * Ri, 4.);
    } else {
        km = 0.;
        kh = 0.;
    }
    const double kmax = 0.8 * G.dx * G.dz / G.dt;
    km = maxs(km, 10.);
    kh = maxs(kh, 10.);
    km = mins(km, kmax);
    kh = mins(kh, kmax);
    store_center(S.RKM, G, i, k, km);
    store_center(S.RKH, G, i, k, kh);
}

// NumericalProcess::Nudge_qv (NumericalProcess.cpp:86-95), all points.
__global__ void __launch_bounds__(BX *BY) k_nudge_qv(Grid G, StepPtrs S, double tau) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i >= G.nx || k >= G.nz) return;
    const size_t p = ix(G, i, k);
    S.fut[F_QV][p] = S.fut[F_QV][p] + G.dt * (S.P.qvb0[k] - S.now[F_QV][p]) / tau;
}

// Iteration::updateMean (src/Iteration.cpp:42-86) as called by
// nextTimeStep: one block per level.  thb_zeta is not read by the step and
// is formed on download.
__global__ void __launch_bounds__(256) k_update_mean(Grid G, StepPtrs S, int shift_ubar) {
    const int k = blockIdx.x + 1;
    const double *thp = S.fut[F_TH], *qvp = S.fut[F_QV];
    double acc[2] = {0., 0.};
    for (int i = 1 + threadIdx.x; i < G.nx - 1; i += blockDim.x) {
        acc[0] += thp[ix(G, i, k)];
        acc[1] += qvp[ix(G, i, k)];
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        double *thb = S.P.thb, *qvb = S.P.qvb, *thvb = S.P.thvb;
        const double tb = acc[0] / (double)(G.nx - 2.), qb = acc[1] / (double)(G.nx - 2.);
        const double tv = tb * (1 + 0.608 * qb);
        S.P.thbm[k] = thb[k];
        S.P.thvbm[k] = thvb[k];
        thb[k] = tb;
        qvb[k] = qb;
        thvb[k] = tv;
        if (k == 1) {
            S.P.thbm[0] = thb[0];
            S.P.thvbm[0] = thvb[0];
            thb[0] = tb;
            qvb[0] = qb;
            thvb[0] = tv;
            if (shift_ubar) {
                // nextTimeStep's copy loop (src/Iteration.cpp:166-167)
                S.s[S_UBARTOPM] = S.s[S_UBARTOPP];
                S.s[S_UBARTOP] = S.s[S_UBARTOPP];
            }
        }
        if (k == G.nz - 2) {
            S.P.thbm[k + 1] = thb[k + 1];
            S.P.thvbm[k + 1] = thvb[k + 1];
            thb[k + 1] = tb;
            qvb[k + 1] = qb;
            thvb[k + 1] = tv;
        }
    }
}

}  // namespace

void launch_kessler(const Grid &G, const StepPtrs &S, const RedSlot &rs, hipStream_t st) {
    hipLaunchKernelGGL(k_kessler<KP_ALL>, grid2d(G), dim3(BX, BY), 0, st, G, S, rs);
    finalize<6>(rs, nblocks(grid2d(G)), S.s, RED_NEGVAL, S_NEG_RATIO, st);
}
void launch_kessler_process(const Grid &G, const StepPtrs &S, int process, hipStream_t st) {
    const RedSlot none{nullptr, 0};
    switch (process) {
        case 0: hipLaunchKernelGGL(k_kessler<KP_AUTO>, grid2d(G), dim3(BX, BY), 0, st, G, S, none); break;
        case 1: hipLaunchKernelGGL(k_kessler<KP_ACCR>, grid2d(G), dim3(BX, BY), 0, st, G, S, none); break;
        case 2: hipLaunchKernelGGL(k_kessler<KP_EVAP>, grid2d(G), dim3(BX, BY), 0, st, G, S, none); break;
        default: hipLaunchKernelGGL(k_kessler<KP_COND>, grid2d(G), dim3(BX, BY), 0, st, G, S, none); break;
    }
}
void launch_negative_value(const Grid &G, const StepPtrs &S, double *a, const RedSlot &rs, hipStream_t st) {
    hipLaunchKernelGGL(k_negval_sums, grid2d(G), dim3(BX, BY), 0, st, G, a, rs);
    finalize<2>(rs, nblocks(grid2d(G)), S.s, RED_NEGVAL, S_NEG_RATIO, st);
    hipLaunchKernelGGL(k_negval_scale, grid2d(G), dim3(BX, BY), 0, st, G, a, S.s + S_NEG_RATIO);
}
void launch_turbulence_coeffs(const Grid &G, const StepPtrs &S, hipStream_t st) {
    hipLaunchKernelGGL(k_turb_coeffs, grid2d(G), dim3(BX, BY), 0, st, G, S);
}

// Nudge_theta and Nudge_zeta alone (NumericalProcess.cpp:60-84), pointwise;
// the timestep fuses them into the diffusion sweep instead.
__global__ void __launch_bounds__(BX *BY) k_nudge(Grid G, StepPtrs S, int nudge) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const size_t p = ix(G, i, k);
    const double rt = S.P.rate_th[k], rz = S.P.rate_z[k];
    if ((nudge & 1) && rt != 0.) S.fut[F_TH][p] -= rt * (S.fut[F_TH][p] - S.P.thb_init[k]);
    if ((nudge & 2) && rz != 0.) S.fut[F_ZETA][p] -= rz * (S.fut[F_ZETA][p]);
}

void launch_nudge(const Grid &G, const StepPtrs &S, int nudge, hipStream_t st) {
    hipLaunchKernelGGL(k_nudge, grid2d(G), dim3(BX, BY), 0, st, G, S, nudge);
}
void launch_nudge_qv(const Grid &G, const StepPtrs &S, double tau, hipStream_t st) {
    hipLaunchKernelGGL(k_nudge_qv, grid2d(G), dim3(BX, BY), 0, st, G, S, tau);
}
void launch_update_mean(const Grid &G, const StepPtrs &S, bool shift_ubar, hipStream_t st) {
    hipLaunchKernelGGL(k_update_mean, dim3(G.nz - 2), dim3(256), 0, st, G, S, int(shift_ubar));
}

}  // namespace vvmd

// ------------------------------------------------ no-allocation entry point
//
// vvm_negative_value_noalloc runs the kernels of the handle's
// NegativeValueProcess operator (k_negval_sums, k_finalize, k_negval_scale)
// on a caller's device field.  Their scratch -- the two partial sums of each
// workgroup, then the correction ratio -- is the file-scope array below
// instead of a handle's buffers, so the call allocates nothing.  Every
// workgroup overwrites its own two partials and k_finalize reads only this
// call's workgroups and writes the ratio before k_negval_scale reads it, so
// nothing needs clearing between calls.  Calls that overlap on one GPU share
// the array and race.

namespace vvmd {
constexpr int kNoAllocMaxBlocks = 8192;  // 64x4-point workgroups: 2050x992 points at most
}  // namespace vvmd

static __device__ __align__(256) double scratch[2 * vvmd::kNoAllocMaxBlocks + 32];

namespace {
// Device address of `scratch` on each device, taken once with that device
// current (each GPU holds its own copy of the array at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return VVM_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "vvm_negative_value_noalloc: hipGetSymbolAddress(scratch) on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return VVM_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return VVM_OK;
}
}  // namespace

extern "C" int vvm_negative_value_noalloc(double *d_var, int nx, int nz, int pitch, void *stream) {
    using namespace vvmd;
    if (!d_var || nx < 3 || nz < 3 || pitch < nx) {
        fprintf(stderr, "vvm_negative_value_noalloc: bad arguments (var %p, nx %d, nz %d, pitch %d)\n",
                static_cast<void *>(d_var), nx, nz, pitch);
        return VVM_ERR_ARG;
    }
    Grid G{};
    G.nx = nx;
    G.nz = nz;
    G.pitch = pitch;
    const dim3 g = grid2d(G);
    const int nb = nblocks(g);
    if (nb > kNoAllocMaxBlocks) {
        fprintf(stderr,
                "vvm_negative_value_noalloc: a %d x %d field needs %d workgroups of %dx%d points; the "
                "no-allocation scratch holds %d (use vvm_dev_op_field for larger grids)\n",
                nx, nz, nb, BX, BY, kNoAllocMaxBlocks);
        return VVM_ERR_SIZE;
    }
    double *s = nullptr;
    if (const int rc = scratch_for_current_device(&s)) return rc;
    const RedSlot rs{s, kNoAllocMaxBlocks};
    double *ratio = s + 2 * kNoAllocMaxBlocks;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_negval_sums, g, dim3(BX, BY), 0, st, G, d_var, rs);
    finalize<2>(rs, nb, ratio, RED_NEGVAL, 0, st);
    hipLaunchKernelGGL(k_negval_scale, g, dim3(BX, BY), 0, st, G, d_var, ratio);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "vvm_negative_value_noalloc: launch failed: %s\n", hipGetErrorString(e));
        return VVM_ERR_HIP;
    }
    return VVM_OK;
}
