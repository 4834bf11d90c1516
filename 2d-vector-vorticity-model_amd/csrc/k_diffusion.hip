// k_diffusion.hip -- Turbulence::Mparam / Hparam (src/Turbulent.cpp:68-114)
// on the device, with Nudge_theta / Nudge_zeta fused into the store.
//
// The reference updates each field in place, k outer and i inner, so a new
// value depends on the *new* west and lower neighbours:
//   new(i,k) = a(i,k) new(i-1,k) + c(i,k) + d(i,k) new(i,k-1),
// where a, d hold the diffusion coefficients and c the old centre, east and
// upper values; ghost cells are never updated during the sweep.
//
// Two kernels.  k_diff_coeffs forms a, c, d for every point of the five
// fields in parallel over the whole GPU (a and d of Hparam depend only on
// RKH and are shared by th, qv, qc, qr).  k_diff_sweep then does only the
// sequential part: row k is a first-order linear recurrence in i once row
// k-1 is final.  One workgroup per field; wavefront w owns a strip of 64*E
// consecutive columns and walks the rows in order.  In each row a wave
// composes its lanes' affine maps y -> a y + e (e = c + d y_below) with a
// 64-lane scan, takes the carry (the new value at the last column of strip
// w-1) from LDS, applies it and publishes its own last value.  The waves
// form a pipeline along the rows.
//
// The carry slots start as a NaN sentinel that arithmetic never produces; a
// wave spins (relaxed LDS loads, no fence) until its slot holds a value.  A
// wave waits only for the wave on its left and wave 0 never waits, so every
// wait ends; a spin limit turns a wait that does not end into an error flag
// instead of a hang.  Values agree with the sequential sweep to rounding.
#pragma clang fp contract(off)

#include <algorithm>
#include <mutex>
#include <vector>

#include "vvm_internal.h"

namespace vvmd {

namespace {

constexpr int kMaxWaves = 16;
constexpr int kSpinLimit = 1 << 24;
constexpr unsigned long long kEmpty = 0x7FF4DEADBEEF0001ull;  // a NaN payload arithmetic never makes
constexpr int BX = 64, BY = 4;
constexpr int kMinStripWaves = 2;   // narrowest strip: 2 compute waves (128 columns)
constexpr int kDefStripWaves = 4;   // default strip: 256 columns
__host__ __device__ inline int strip_waves(const StepPtrs &S) { return S.dstrip_cw > 0 ? S.dstrip_cw : kDefStripWaves; }
__host__ __device__ inline int strips_of(const Grid &G, int cw) { return (G.nx - 2 + cw * 64 - 1) / (cw * 64); }
constexpr int kMaxStrips = 64;
constexpr int kSlotSpinLimit = 1 << 18;

struct Affine {
    double a, b;  // y = a x + b
};

// g after f: y = g.a (f.a x + f.b) + g.b
__device__ inline Affine after(const Affine &g, const Affine &f) { return {g.a * f.a, g.a * f.b + g.b}; }

// A double moved between lanes by a DPP control (two 32-bit moves); lanes
// whose source is outside the pattern, or outside row_mask, keep `old`.
template <int CTRL, int ROWS>
__device__ inline double dppd(double old, double v) {
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(int(o & 0xffffffffll), int(x & 0xffffffffll), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(int(o >> 32), int(x >> 32), CTRL, ROWS, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Inclusive scan of affine maps over the 64 lanes, inc_l = f_l o ... o f_0,
// on DPP moves (a few cycles each, no LDS round trip): row_shr 1, 2, 4, 8
// inside each 16-lane row, then row_bcast 15 (rows 1, 3) and row_bcast 31
// (rows 2, 3) across rows.  Lanes without a source combine with the
// identity map.
template <int CTRL, int ROWS>
__device__ inline void scan_step(Affine &f) {
    const Affine p{dppd<CTRL, ROWS>(1., f.a), dppd<CTRL, ROWS>(0., f.b)};
    f = after(f, p);
}
__device__ inline Affine wave_scan(Affine f) {
    scan_step<0x111, 0xF>(f);  // row_shr:1
    scan_step<0x112, 0xF>(f);  // row_shr:2
    scan_step<0x114, 0xF>(f);  // row_shr:4
    scan_step<0x118, 0xF>(f);  // row_shr:8
    scan_step<0x142, 0xA>(f);  // row_bcast:15 into rows 1 and 3
    scan_step<0x143, 0xC>(f);  // row_bcast:31 into rows 2 and 3
    return f;
}

// a, c, d of every interior point (Turbulent.cpp:81-86 for Mparam on zeta,
// :105-110 for Hparam on th, qv, qc, qr).  Row k = 1 folds the (old) ghost
// row below into c.
__global__ void __launch_bounds__(BX *BY) k_diff_coeffs(Grid G, StepPtrs S) {
    if (S.dslots) {  // empty carry slots for the sweep that follows
        const int nslot = NPROG * strips_of(G, strip_waves(S)) * G.nz;
        const int nthr = gridDim.x * gridDim.y * BX * BY;
        for (int t = (blockIdx.y * gridDim.x + blockIdx.x) * (BX * BY) + threadIdx.y * BX + threadIdx.x; t < nslot;
             t += nthr)
            __hip_atomic_store(S.dslots + t, kEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const size_t p = ix(G, i, k);
    const double A = G.rdx2 * G.dt;
    const double *rhou = S.P.rhou, *rhow = S.P.rhow;
    {  // Mparam
        const double *K = S.RKM, *f = S.fut[F_ZETA];
        const double Ke = 0.5 * (K[ix(G, i, k)] + K[ix(G, i, k - 1)]);
        const double Kw = 0.5 * (K[ix(G, i - 1, k)] + K[ix(G, i - 1, k - 1)]);
        const double Ct = rhou[k] * 0.5 * (K[ix(G, i, k)] + K[ix(G, i - 1, k)]);
        const double Cb = rhou[k - 1] * 0.5 * (K[ix(G, i, k - 1)] + K[ix(G, i - 1, k - 1)]);
        const double Bk = G.rdz2 * G.dt / rhow[k];
        const double a = A * Kw, d = Bk * Cb;
        const double old = f[p], oldE = f[ix(G, i + 1, k)], oldT = f[ix(G, i, k + 1)];
        double c = old + A * (Ke * (oldE - old) - Kw * old) + Bk * (Ct * (oldT - old) - Cb * old);
        if (k == 1) c = c + d * f[ix(G, i, 0)];
        S.dcoef[DC_AM][p] = a;
        S.dcoef[DC_DM][p] = (k == 1) ? 0. : d;
        S.dcoef[DC_C0 + F_ZETA][p] = c;
    }
    {  // Hparam
        const double *K = S.RKH;
        const double Ke = 0.5 * (K[ix(G, i + 1, k)] + K[ix(G, i, k)]);
        const double Kw = 0.5 * (K[ix(G, i, k)] + K[ix(G, i - 1, k)]);
        const double Ct = rhow[k + 1] * 0.5 * (K[ix(G, i, k + 1)] + K[ix(G, i, k)]);
        const double Cb = rhow[k] * 0.5 * (K[ix(G, i, k)] + K[ix(G, i, k - 1)]);
        const double Bk = G.rdz2 * G.dt / rhou[k];
        const double a = A * Kw, d = Bk * Cb;
        S.dcoef[DC_AH][p] = a;
        S.dcoef[DC_DH][p] = (k == 1) ? 0. : d;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int fld = F_TH + q;
            const double *f = S.fut[fld];
            const double old = f[p], oldE = f[ix(G, i + 1, k)], oldT = f[ix(G, i, k + 1)];
            double c = old + A * (Ke * (oldE - old) - Kw * old) + Bk * (Ct * (oldT - old) - Cb * old);
            if (k == 1) c = c + d * f[ix(G, i, 0)];
            S.dcoef[DC_C0 + fld][p] = c;
        }
    }
}

// Rows of coefficients kept in flight per lane: the row-to-row chain is far
// shorter than a load from memory, so loads are issued PF rows ahead into a
// circular register buffer (the row loop is unrolled by PF so the buffer is
// statically indexed).
constexpr int PF = 4;

template <int E>
__device__ inline void load_row(const Grid &G, const double *__restrict__ A_, const double *__restrict__ C_,
                                const double *__restrict__ D_, int i0, int n, int k, int nz, double (&a)[E],
                                double (&c)[E], double (&d)[E]) {
    // Unconditional loads from a clamped (valid) position: a load under a
    // branch merges with a default value, and that merge makes the compiler
    // wait for the load at once, which would serialise the prefetch.  Values
    // of positions past the last row or column are never used.
    const int kk = min(k, nz - 2);
#pragma unroll
    for (int j = 0; j < E; j++) {
        const size_t p = ix(G, min(i0 + j, n), kk);
        a[j] = A_[p];
        c[j] = C_[p];
        d[j] = D_[p];
    }
}

template <int E>
__device__ void sweep(const Grid &G, double *__restrict__ f, const double *__restrict__ A_,
                      const double *__restrict__ C_, const double *__restrict__ D_, const double *rate, int is_theta,
                      const double *thb_init, unsigned long long *carry, double *err) {
    const int n = G.nx - 2, nz = G.nz;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int strip0 = 1 + w * kWave * E;               // first column of the strip
    const int strip1 = min(strip0 + kWave * E - 1, n);  // last column of the strip
    const int i0 = strip0 + lane * E;

    double ba[PF][E], bc[PF][E], bd[PF][E], ynew[E];
#pragma unroll
    for (int j = 0; j < E; j++) ynew[j] = 0.;
#pragma unroll
    for (int q = 0; q < PF; q++) load_row<E>(G, A_, C_, D_, i0, n, 1 + q, nz, ba[q], bc[q], bd[q]);
    for (int k0 = 1; k0 <= nz - 2; k0 += PF) {
#pragma unroll
    for (int q = 0; q < PF; q++) {
        const int k = k0 + q;
        if (k > nz - 2) break;
        double a[E], e[E];
        Affine m{1., 0.};
#pragma unroll
        for (int j = 0; j < E; j++) {
            a[j] = ba[q][j];
            e[j] = bc[q][j] + bd[q][j] * ynew[j];
            if (i0 + j <= n) m = after(Affine{a[j], e[j]}, m);
        }
        // refill this slot with row k + PF
        load_row<E>(G, A_, C_, D_, i0, n, k + PF, nz, ba[q], bc[q], bd[q]);
        // inclusive scan over the wave, then the exclusive prefix of this lane
        Affine inc = m;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const double pa = __shfl_up(inc.a, off, kWave), pb = __shfl_up(inc.b, off, kWave);
            if (lane >= off) inc = after(inc, Affine{pa, pb});
        }
        const double ea = __shfl_up(inc.a, 1, kWave), eb = __shfl_up(inc.b, 1, kWave);
        // carry: the new value left of the strip
        double y0;
        if (w == 0) {
            y0 = f[ix(G, 0, k)];  // ghost column, old as in the reference
        } else {
            unsigned long long bits = kEmpty;
            if (lane == 0) {
                int spins = 0;
                unsigned long long *slot = carry + (w - 1) * nz + k;
                while ((bits = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) == kEmpty) {
                    if (++spins > kSpinLimit) {
                        *err = 1.;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            y0 = __shfl(__longlong_as_double(static_cast<long long>(bits)), 0, kWave);
        }
        double y = (lane == 0) ? y0 : ea * y0 + eb;
        const double r = rate ? rate[k] : 0.;
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int i = i0 + j;
            if (i > n) continue;
            y = a[j] * y + e[j];
            ynew[j] = y;
            double out = y;
            if (r != 0.) out = is_theta ? y - r * (y - thb_init[k]) : y - r * y;
            f[ix(G, i, k)] = out;
            if (i == strip1)
                __hip_atomic_store(carry + w * nz + k, static_cast<unsigned long long>(__double_as_longlong(y)),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    }
}

// blockIdx.x: 0 zeta (Mparam), 1..4 th, qv, qc, qr (Hparam).
// nudge bit 0 fuses Nudge_theta, bit 1 Nudge_zeta.
template <int E>
__global__ void __launch_bounds__(kMaxWaves *kWave) k_diff_sweep(Grid G, StepPtrs S, int nudge, int fbase) {
    extern __shared__ unsigned long long carry[];  // [waves][nz]
    const int waves = blockDim.x / kWave;
    for (int t = threadIdx.x; t < waves * G.nz; t += blockDim.x) carry[t] = kEmpty;
    __syncthreads();
    const int fld = fbase + blockIdx.x;  // F_ZETA .. F_QR
    const bool mpar = (fld == F_ZETA);
    const double *rate = mpar ? ((nudge & 2) ? S.P.rate_z : nullptr) : ((fld == F_TH && (nudge & 1)) ? S.P.rate_th : nullptr);
    sweep<E>(G, S.fut[fld], S.dcoef[mpar ? DC_AM : DC_AH], S.dcoef[DC_C0 + fld], S.dcoef[mpar ? DC_DM : DC_DH], rate,
             fld == F_TH, S.P.thb_init, carry, S.s + S_DIFF_ERR);
}

// ---------------------------------------------------------------------------
// Multi-workgroup sweep.  The columns of a field are cut into strips of
// W = cw * 64 * E columns; strip s of field f is workgroup (s, f), and the
// strips of a field run on different CUs as a pipeline along the rows.
// Inside a workgroup, cw compute wavefronts run the same wave pipeline as
// the one-workgroup sweep (LDS carries between neighbouring waves).  One
// more wavefront, the courier, moves carries between workgroups: it fetches
// the carries of the strip on the left (up to 64 rows per round trip) into
// LDS for compute wave 0, and publishes the last compute wave's carries for
// the strip on the right.  Only the courier touches the inter-workgroup
// slots, so its long-latency loads never hold up the compute waves' own
// memory counters.
//
// The slots live in uncached device memory: a relaxed agent-scope store is
// visible to a load on any CU without cache maintenance.  Each carry is its
// own ready flag (the kEmpty NaN payload marks an empty slot; k_diff_coeffs
// empties the slots before every sweep).  Workgroups are numbered
// strip-fastest, so every producer is dispatched before its consumer, and
// all workgroups of a launch (strips x 5 fields) are resident at once.
// Every wait has a spin limit; a courier that gives up fills the remaining
// carries with NaN so that its compute waves finish, and sets the error flag.
constexpr int kMaxCompute = 8;

__device__ inline unsigned long long slot_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long lds_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_store(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Lanes 0..63 look at rows k0 + lane; returns how many rows from k0 on are
// ready (values not kEmpty), counting rows past kend as ready.
__device__ inline int ready_run(unsigned long long v, int k0, int kend) {
    const int lane = threadIdx.x % kWave;
    const unsigned long long empty = __ballot(k0 + lane <= kend && v == kEmpty);
    const int span = min(kWave, kend - k0 + 1);
    return empty ? min(__builtin_ctzll(empty), span) : span;
}

__device__ void courier(const unsigned long long *in, unsigned long long *out, unsigned long long *cin,
                        const unsigned long long *clast, int nz, double *err) {
    const int lane = threadIdx.x % kWave;
    const int kend = nz - 2;
    int kin = in ? 1 : kend + 1, kout = out ? 1 : kend + 1;
    int idle = 0;
    while (kin <= kend || kout <= kend) {
        bool moved = false;
        if (kin <= kend) {
            const int k = kin + lane;
            const unsigned long long v = (k <= kend) ? slot_load(in + k) : 0ull;
            const int r = ready_run(v, kin, kend);
            if (lane < r) lds_store(cin + k, v);
            kin += r;
            moved = moved || r > 0;
        }
        if (kout <= kend) {
            const int k = kout + lane;
            const unsigned long long v = (k <= kend) ? lds_load(clast + k) : 0ull;
            const int r = ready_run(v, kout, kend);
            if (lane < r)
                __hip_atomic_store(out + k, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            kout += r;
            moved = moved || r > 0;
        }
        if (moved) {
            idle = 0;
        } else if (++idle > kSlotSpinLimit) {
            if (lane == 0) *err = 1.;
            // unblock compute wave 0: the rows not fetched become NaN
            const unsigned long long qnan = 0x7FF8000000000000ull;
            for (int k = kin + lane; k <= kend; k += kWave) lds_store(cin + k, qnan);
            break;
        } else {
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

template <int E, int CW>
__global__ void __launch_bounds__((CW + 1) * kWave) k_diff_strips(Grid G, StepPtrs S, int nudge, int nstrip,
                                                                   int fbase) {
    // LDS: carry[cw][nz], cin[nz] (u64), then ghost[nz], rate[nz], thbi[nz]
    // (doubles): the per-row values a compute wave reads come from LDS, so
    // its only global loads are the prefetched coefficient rows (a global
    // load consumed at once would make the wave wait for every earlier,
    // prefetched load as well)
    extern __shared__ unsigned long long lds[];
    constexpr int cw = CW;
    const int nz = G.nz, n = G.nx - 2;
    unsigned long long *carry = lds, *cin = lds + size_t(cw) * nz;
    double *ghost = reinterpret_cast<double *>(lds + size_t(cw + 1) * nz), *rate_l = ghost + nz, *thbi = rate_l + nz;
    const int s = blockIdx.x, fld = fbase + blockIdx.y;
    {
        const bool mz = (fld == F_ZETA);
        const double *rt = mz ? ((nudge & 2) ? S.P.rate_z : nullptr) : ((fld == F_TH && (nudge & 1)) ? S.P.rate_th : nullptr);
        const double *fv = S.fut[fld];
        for (int t = threadIdx.x; t < (cw + 1) * nz; t += blockDim.x) lds[t] = kEmpty;
        for (int k = threadIdx.x; k < nz; k += blockDim.x) {
            ghost[k] = fv[ix(G, 0, k)];
            rate_l[k] = rt ? rt[k] : 0.;
            thbi[k] = S.P.thb_init[k];
        }
    }
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const unsigned long long *in = (s > 0) ? S.dslots + (size_t(fld) * nstrip + (s - 1)) * nz : nullptr;
    unsigned long long *out = (s + 1 < nstrip) ? S.dslots + (size_t(fld) * nstrip + s) * nz : nullptr;
    if (w == cw) {
        courier(in, out, cin, carry + size_t(cw - 1) * nz, nz, S.s + S_DIFF_ERR);
        return;
    }
    const bool mpar = (fld == F_ZETA);
    const double *__restrict__ A_ = S.dcoef[mpar ? DC_AM : DC_AH];
    const double *__restrict__ C_ = S.dcoef[DC_C0 + fld];
    const double *__restrict__ D_ = S.dcoef[mpar ? DC_DM : DC_DH];
    double *f = S.fut[fld];
    const bool is_theta = (fld == F_TH);
    const int W = cw * kWave * E;
    const int strip0 = 1 + s * W + w * kWave * E;       // first column of this wave
    const int strip1 = min(strip0 + kWave * E - 1, n);  // last column of this wave
    const int i0 = strip0 + lane * E;
    // waves past the last column of a partial strip have no work; every
    // wave after them is past it too, so nobody waits for their carries
    if (strip0 > n) return;
    // where this wave's carry comes from: the wave on its left, the
    // courier (wave 0 of strips s > 0), or the old ghost column (strip 0)
    const unsigned long long *src = (w > 0) ? carry + size_t(w - 1) * nz : (in ? cin : nullptr);
    unsigned long long *dst = carry + size_t(w) * nz;
    const bool publish = (w < cw - 1) || (out != nullptr);

    // rows in flight per lane: a wave's row takes well under a memory
    // latency, so the loads run PFS rows ahead (8; 4 for the widest strip,
    // whose nine wavefronts leave fewer registers per lane)
    constexpr int PFS = (CW >= 8) ? 4 : 8;
    double ba[PFS][E], bc[PFS][E], bd[PFS][E], ynew[E];
#pragma unroll
    for (int j = 0; j < E; j++) ynew[j] = 0.;
#pragma unroll
    for (int q = 0; q < PFS; q++) load_row<E>(G, A_, C_, D_, i0, n, 1 + q, nz, ba[q], bc[q], bd[q]);
    bool gave_up = false;
    // The scan of a row's affine maps y -> a y + e splits in two chains: the
    // products of the a's (P), which do not depend on the row below, and
    // the offsets (B), which do.  P of row k+1 is formed while row k's B
    // chain runs (independent instructions the compiler interleaves with
    // the dependent ones), so the per-row critical path is the B chain
    // alone: one DPP move pair, a multiply and an add per step.  Same
    // operations, same rounding as the combined affine scan.
    double Pc[7];  // P at the start of scan steps 1..6 and the inclusive P, current row
    auto pchain = [&](const int q, double (&P)[7]) {
        double p = 1.;
#pragma unroll
        for (int j = 0; j < E; j++)
            if (i0 + j <= n) p = ba[q][j] * p;
        P[0] = p;
        p = p * dppd<0x111, 0xF>(1., p);
        P[1] = p;
        p = p * dppd<0x112, 0xF>(1., p);
        P[2] = p;
        p = p * dppd<0x114, 0xF>(1., p);
        P[3] = p;
        p = p * dppd<0x118, 0xF>(1., p);
        P[4] = p;
        p = p * dppd<0x142, 0xA>(1., p);
        P[5] = p;
        p = p * dppd<0x143, 0xC>(1., p);
        P[6] = p;
    };
    pchain(0, Pc);
    // one row from prefetch slot q; with `refill` the slot is then loaded
    // again for row k + PFS
    auto row = [&](const int k, const int q, const bool refill) {
        // The carry and the row's LDS values are read first, by every lane
        // (one broadcast word), so their latency passes under the scan; the
        // wait below only spins when the wave on the left has not yet
        // published this row.
        unsigned long long bits = src ? lds_load(src + k) : 0ull;
        const double r = rate_l[k], tb = thbi[k], gk = ghost[k];
        double a[E], e[E];
        double b = 0.;
#pragma unroll
        for (int j = 0; j < E; j++) {
            a[j] = ba[q][j];
            e[j] = bc[q][j] + bd[q][j] * ynew[j];
            if (i0 + j <= n) b = a[j] * b + e[j];
        }
        // the B chain of this row, step by step beside the P chain of the
        // next row (written interleaved so that the two fill each other's
        // latency)
        const int qn = (q + 1) % PFS;
        double pn = 1.;
#pragma unroll
        for (int j = 0; j < E; j++)
            if (i0 + j <= n) pn = ba[qn][j] * pn;
        double Pn[7];
        Pn[0] = pn;
        const double d1 = dppd<0x111, 0xF>(0., b), p1 = dppd<0x111, 0xF>(1., pn);  // row_shr:1
        b = Pc[0] * d1 + b;
        pn = pn * p1;
        Pn[1] = pn;
        const double d2 = dppd<0x112, 0xF>(0., b), p2 = dppd<0x112, 0xF>(1., pn);  // row_shr:2
        b = Pc[1] * d2 + b;
        pn = pn * p2;
        Pn[2] = pn;
        const double d3 = dppd<0x114, 0xF>(0., b), p3 = dppd<0x114, 0xF>(1., pn);  // row_shr:4
        b = Pc[2] * d3 + b;
        pn = pn * p3;
        Pn[3] = pn;
        const double d4 = dppd<0x118, 0xF>(0., b), p4 = dppd<0x118, 0xF>(1., pn);  // row_shr:8
        b = Pc[3] * d4 + b;
        pn = pn * p4;
        Pn[4] = pn;
        const double d5 = dppd<0x142, 0xA>(0., b), p5 = dppd<0x142, 0xA>(1., pn);  // row_bcast:15, rows 1, 3
        b = Pc[4] * d5 + b;
        pn = pn * p5;
        Pn[5] = pn;
        const double d6 = dppd<0x143, 0xC>(0., b), p6 = dppd<0x143, 0xC>(1., pn);  // row_bcast:31, rows 2, 3
        b = Pc[5] * d6 + b;
        pn = pn * p6;
        Pn[6] = pn;
        // exclusive prefix: the inclusive one of the lane below (identity at lane 0)
        const double ea = dppd<0x138, 0xF>(1., Pc[6]), eb = dppd<0x138, 0xF>(0., b);
        double y0;
        if (!src) {
            y0 = gk;  // ghost column, old as in the reference
        } else {
            // every lane holds the same word, so the loop is uniform
            int spins = 0;
            while (bits == kEmpty && !gave_up) {
                if (++spins > kSpinLimit) {
                    S.s[S_DIFF_ERR] = 1.;
                    gave_up = true;
                }
                bits = lds_load(src + k);
            }
            y0 = __longlong_as_double(static_cast<long long>(bits));
        }
        double y = (lane == 0) ? y0 : ea * y0 + eb;
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int i = i0 + j;
            if (i > n) continue;
            y = a[j] * y + e[j];
            ynew[j] = y;
        }
        // publish the carry before the nudging and the store, which the
        // wave on the right does not wait for
        if (publish && i0 + E - 1 >= strip1 && i0 <= strip1)
            lds_store(dst + k, static_cast<unsigned long long>(__double_as_longlong(ynew[strip1 - i0])));
#pragma unroll
        for (int j = 0; j < E; j++) {
            const int i = i0 + j;
            if (i > n) continue;
            const double yj = ynew[j];
            double o = yj;
            if (r != 0.) o = is_theta ? yj - r * (yj - tb) : yj - r * yj;
            f[ix(G, i, k)] = o;
        }
#pragma unroll
        for (int t = 0; t < 7; t++) Pc[t] = Pn[t];
        // refill after the last use of the slot, so that the loaded values
        // stay in the slot's own registers (no copy that would wait for them)
        if (refill) load_row<E>(G, A_, C_, D_, i0, n, k + PFS, nz, ba[q], bc[q], bd[q]);
    };
    // Whole groups of PFS rows first, with no exit inside a group: an exit
    // between the rows of an unrolled group makes the compiler copy each
    // refilled slot at once, which waits for its load and undoes the
    // prefetch.  The last, partial group refills nothing.
    const int kend = nz - 2;
    int k0 = 1;
    for (; k0 + PFS - 1 <= kend; k0 += PFS) {
#pragma unroll
        for (int q = 0; q < PFS; q++) row(k0 + q, q, true);
    }
#pragma unroll
    for (int q = 0; q < PFS; q++)
        if (k0 + q <= kend) row(k0 + q, q, false);
}


template <int E>
void launch_e(const Grid &G, const StepPtrs &S, int nudge, int fbase, int nfld, hipStream_t st) {
    const int n = G.nx - 2;
    const int waves = (n + kWave * E - 1) / (kWave * E);
    const size_t lds = size_t(waves) * G.nz * sizeof(unsigned long long);
    hipLaunchKernelGGL(k_diff_sweep<E>, dim3(nfld), dim3(waves * kWave), lds, st, G, S, nudge, fbase);
}

}  // namespace

bool diffusion_strip_waves_ok(int cw) { return cw >= kMinStripWaves && cw <= kMaxCompute; }

// nx - 2 <= 16 waves * 64 lanes * 8 columns, and the carry rows fit in LDS.
bool diffusion_supported(const Grid &G) {
    return G.nx - 2 <= kMaxWaves * kWave * 8 && size_t(kMaxWaves) * G.nz * sizeof(double) <= 60 * 1024;
}

// Carry-slot regions of the multi-workgroup sweep, one per device handle.
// Uncached memory is never returned to the runtime: a region of a
// destroyed handle goes back to this pool for the next handle.
namespace {
struct SlotRegion {
    int device;
    size_t bytes;
    unsigned long long *p;
    bool busy;
};
std::mutex g_slot_mu;
std::vector<SlotRegion> g_slots;
}  // namespace

unsigned long long *diffusion_slots(const Grid &G, int device) {
    // room for the narrowest strips the tuner may pick
    const int ns = strips_of(G, kMinStripWaves);
    if (ns < 2 || ns > kMaxStrips) return nullptr;
    const size_t need = size_t(NPROG) * ns * G.nz * sizeof(unsigned long long);
    std::lock_guard<std::mutex> lk(g_slot_mu);
    for (SlotRegion &r : g_slots)
        if (!r.busy && r.device == device && r.bytes >= need) {
            r.busy = true;
            return r.p;
        }
    const size_t bytes = std::max(need, size_t(64) << 10);
    void *p = nullptr;
    if (hipExtMallocWithFlags(&p, bytes, hipDeviceMallocUncached) != hipSuccess || !p) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_slots.push_back({device, bytes, static_cast<unsigned long long *>(p), true});
    return g_slots.back().p;
}

void diffusion_slots_release(unsigned long long *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_slot_mu);
    for (SlotRegion &r : g_slots)
        if (r.p == p) r.busy = false;
}

void launch_diffusion(const Grid &G, const StepPtrs &S, int nudge, hipStream_t st, int only_field) {
    hipLaunchKernelGGL(k_diff_coeffs, dim3((G.nx + BX - 1) / BX, (G.nz + BY - 1) / BY), dim3(BX, BY), 0, st, G, S);
    const int n = G.nx - 2;
    const int fbase = only_field >= 0 ? only_field : 0, nfld = only_field >= 0 ? 1 : NPROG;
    if (S.dslots) {
        const int cw = strip_waves(S);  // compute waves of one column per lane
        const int ns = strips_of(G, cw);
        const size_t lds = size_t(cw + 4) * G.nz * sizeof(unsigned long long);
        const dim3 grid(ns, nfld), block((cw + 1) * kWave);
        if (cw == 2) hipLaunchKernelGGL((k_diff_strips<1, 2>), grid, block, lds, st, G, S, nudge, ns, fbase);
        else if (cw == 8) hipLaunchKernelGGL((k_diff_strips<1, 8>), grid, block, lds, st, G, S, nudge, ns, fbase);
        else hipLaunchKernelGGL((k_diff_strips<1, 4>), grid, block, lds, st, G, S, nudge, ns, fbase);
        return;
    }
    if (n <= kMaxWaves * kWave) launch_e<1>(G, S, nudge, fbase, nfld, st);
    else if (n <= 2 * kMaxWaves * kWave) launch_e<2>(G, S, nudge, fbase, nfld, st);
    else if (n <= 4 * kMaxWaves * kWave) launch_e<4>(G, S, nudge, fbase, nfld, st);
    else launch_e<8>(G, S, nudge, fbase, nfld, st);
}

}  // namespace vvmd
