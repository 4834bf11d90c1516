// k_poisson.hip -- the elliptic part of a VVM timestep on the device:
// PoissonSolver::cal_w (src/PoissonSolver.cpp:439-477) as multigrid-
// preconditioned CG in FP64, or FP32 MG-PCG inside FP64 iterative
// refinement, and PoissonSolver::cal_u (src/PoissonSolver.cpp:580-653).
//
// Multigrid: geometric, vertex coarsening (N -> N/2 periodic columns,
// M -> (M-1)/2 Dirichlet rows), operators rediscretised with harmonic-mean
// vertical conductances, bilinear prolongation P and restriction P^T,
// damped-Jacobi V(2,2) smoothing and a single-workgroup coarsest solve.
// Every piece is symmetric, so the V-cycle is a valid CG preconditioner.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "poisson.h"

namespace vvmd {

namespace {

constexpr int LX = 64, LY = 4;

inline dim3 lgrid(int N, int M) { return dim3((N + LX - 1) / LX, (M + LY - 1) / LY); }

template <typename T>
__device__ inline T applyL(const T *__restrict__ x, int N, int i, int j, T cxj, T cuj, T cujm) {
    const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
    const size_t r = size_t(j) * N;
    const T c = x[r + i];
    return cxj * (T(2) * c - x[r + im] - x[r + ip]) + cuj * (c - x[r + N + i]) + cujm * (c - x[r - N + i]);
}

// ------------------------------------------------------------------ CG

// The two CG vector kernels run on at most kCGBlocks workgroups, each
// walking LX x LY tiles, so that the update kernel can add the p.q partials
// of every pq workgroup itself (no k_finalize between them).
constexpr int kCGBlocks = 1024;
inline int cg_blocks(int N, int M) { return std::min(kCGBlocks, int(nblocks(lgrid(N, M)))); }

// p = z + beta p (beta = 0 on the first iteration), q = L p, partials of p.q
// into rs_pq.  After the first iteration every workgroup adds the n_rz
// partials of r.z the preconditioner left in rs_rz, beta = r.z / sc[old];
// workgroup 0 stores r.z in sc[fresh] for the update.
template <typename T>
__global__ void __launch_bounds__(LX *LY) k_pcg_pq(int N, int M, const T *z, const T *pin, T *pout, T *q, const T *cx,
                                                   const T *cu, double *sc, RedSlot rs_pq, int first, RedSlot rs_rz,
                                                   int n_rz, int old, int fresh) {
    const int tiles_x = (N + LX - 1) / LX, ntiles = tiles_x * ((M + LY - 1) / LY);
    T beta = T(0);
    if (!first) {
        const double rz = block_total(rs_rz, n_rz);
        beta = T(rz / sc[old]);
        if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) sc[fresh] = rz;
    }
    double part[1] = {0.};
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int i = (t % tiles_x) * LX + threadIdx.x, j = 1 + (t / tiles_x) * LY + threadIdx.y;
        if (i < N && j <= M) {
            const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
            const size_t r = size_t(j) * N;
            auto pn = [&](size_t a) -> T { return first ? z[a] : z[a] + beta * pin[a]; };
            const T c = pn(r + i);
            const T Lp = cx[j] * (T(2) * c - pn(r + im) - pn(r + ip)) + cu[j] * (c - pn(r + N + i)) +
                         cu[j - 1] * (c - pn(r - N + i));
            pout[r + i] = c;
            q[r + i] = Lp;
            part[0] += double(c) * double(Lp);
        }
    }
    block_partials<1>(part, rs_pq);
}

// alpha = sc[rz] / p.q (p.q from the npq partials of k_pcg_pq), x += alpha p,
// r -= alpha q, and ||w r||^2 with the row weights w (rhow: the reference
// system's residual) as partials in rs.
template <typename T>
__global__ void __launch_bounds__(LX *LY) k_pcg_update(int N, int M, T *x, T *r, const T *p, const T *q,
                                                       const T *wD, double *sc, int rz, RedSlot rs_pq, int npq,
                                                       RedSlot rs) {
    const int tiles_x = (N + LX - 1) / LX, ntiles = tiles_x * ((M + LY - 1) / LY);
    const T alpha = T(sc[rz] / block_total(rs_pq, npq));
    double part[1] = {0.};
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int i = (t % tiles_x) * LX + threadIdx.x, j = 1 + (t / tiles_x) * LY + threadIdx.y;
        if (i < N && j <= M) {
            const size_t c = size_t(j) * N + i;
            x[c] = x[c] + alpha * p[c];
            const T rn = r[c] - alpha * q[c];
            r[c] = rn;
            const double wr = double(wD[j]) * double(rn);
            part[0] += wr * wr;
        }
    }
    block_partials<1>(part, rs);
}

// The solve's first residual with the guess or the refinement correction
// folded in.  XF_EXTRAP: x = the extrapolation of order ord (2 x1 - x2, or
// 3 (x1 - x2) + x3), XF_ZERO:
// x = 0, XF_CORR: x = xin + e32 (the FP32 correction; written to another
// buffer, since neighbouring workgroups read xin), XF_NONE: x as it is.
// Every value of x the stencil needs is formed on the fly with the same
// arithmetic the separate kernels used, so r and x are bit for bit what they
// gave.  With r32 the FP32 right-hand side of the next correction (r as
// float, x32 = 0) is written too.
enum { XF_NONE = 0, XF_EXTRAP, XF_ZERO, XF_CORR };
template <int XF>
__global__ void __launch_bounds__(LX *LY) k_residual_x(int N, int M, const double *f, const double *xin,
                                                       double *xout, const double *x1, const double *x2,
                                                       const double *x3, int ord, const float *e32, double *r, const double *cx,
                                                       const double *cu, const double *wD, RedSlot rs, float *r32,
                                                       float *x32) {
    const int i = blockIdx.x * LX + threadIdx.x, j = 1 + blockIdx.y * LY + threadIdx.y;
    auto xv = [&](size_t a) -> double {
        if (XF == XF_EXTRAP) return extrap_at(x1, x2, x3, ord, a);
        if (XF == XF_ZERO) return 0.;
        if (XF == XF_CORR) return xin[a] + double(e32[a]);
        return xin[a];
    };
    double part[1] = {0.};
    if (i < N && j <= M) {
        const size_t c = size_t(j) * N + i;
        const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
        const size_t rw = size_t(j) * N;
        const double xc = xv(rw + i);
        const double Lx = cx[j] * (2. * xc - xv(rw + im) - xv(rw + ip)) + cu[j] * (xc - xv(rw + N + i)) +
                          cu[j - 1] * (xc - xv(rw - N + i));
        const double rv = f[c] - Lx;
        r[c] = rv;
        if (XF != XF_NONE) {
            xout[c] = xc;
            // the boundary rows: XF_EXTRAP and XF_ZERO form them as the
            // separate kernels did, XF_CORR carries them over
            if (j == 1) xout[i] = (XF == XF_CORR) ? xin[i] : xv(i);
            if (j == M) xout[rw + N + i] = (XF == XF_CORR) ? xin[rw + N + i] : xv(rw + N + i);
        }
        if (r32) {
            r32[c] = float(rv);
            x32[c] = 0.f;
        }
        const double wr = wD[j] * rv;
        part[0] = wr * wr;
    }
    block_partials<1>(part, rs);
}

// ------------------------------------------------------------- cal_w glue

// b = rhow^2 (zetap[i+1] - zetap[i]) dx (src/PoissonSolver.cpp:444-450) for
// k = 2..nz-2 and f = -b / rhow; ||b||^2 into sc[P_BNORM2].
__global__ void __launch_bounds__(LX *LY) k_calw_rhs(Grid G, const double *zetap, const double *rhow, double *f,
                                                     double *sc, RedSlot rs) {
    const int N = G.nx - 2, M = G.nz - 3;
    const int ii = blockIdx.x * LX + threadIdx.x, j = 1 + blockIdx.y * LY + threadIdx.y;
    double part[1] = {0.};
    if (ii < N && j <= M) {
        const int i = ii + 1, k = j + 1;
        const double b = rhow[k] * rhow[k] * (zetap[ix(G, i + 1, k)] - zetap[ix(G, i, k)]) * G.dx;
        f[size_t(j) * N + ii] = -b / rhow[k];
        part[0] = b * b;
    }
    block_partials<1>(part, rs);
}

__global__ void __launch_bounds__(LX *LY) k_rhs_from_b(int N, int M, const double *bref, const double *wD,
                                                       double *f, double *sc, RedSlot rs) {
    const int ii = blockIdx.x * LX + threadIdx.x, j = 1 + blockIdx.y * LY + threadIdx.y;
    double part[1] = {0.};
    if (ii < N && j <= M) {
        const double b = bref[size_t(j - 1) * N + ii];
        f[size_t(j) * N + ii] = -b / wD[j];
        part[0] = b * b;
    }
    block_partials<1>(part, rs);
}

__global__ void k_x_ref(int N, int M, double *x, double *xref, int to_ref) {
    const int ii = blockIdx.x * LX + threadIdx.x, j = 1 + blockIdx.y * LY + threadIdx.y;
    if (ii >= N || j > M) return;
    if (to_ref) xref[size_t(j - 1) * N + ii] = x[size_t(j) * N + ii];
    else x[size_t(j) * N + ii] = xref[size_t(j - 1) * N + ii];
}

// w[i][k] = x / rhow[k] and BoundaryProcess2D_westdown(w) (:469-476).
__global__ void __launch_bounds__(LX *LY) k_calw_back(Grid G, const double *x, double *w, const double *rhow) {
    const int i = blockIdx.x * LX + threadIdx.x, k = blockIdx.y * LY + threadIdx.y;
    if (i >= G.nx || k >= G.nz) return;
    if (k == 0 || k == 1 || k == G.nz - 1) {
        w[ix(G, i, k)] = 0.;
        return;
    }
    if (i < 1 || i > G.nx - 2) return;
    const int N = G.nx - 2;
    store_westdown(w, G, i, k, x[size_t(k - 1) * N + (i - 1)] / rhow[k]);
}

__global__ void k_extrapolate(size_t n, double *x, const double *x1, const double *x2, const double *x3, int ord) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    x[t] = extrap_at(x1, x2, x3, ord, t);
}

// ------------------------------------------------------------------ cal_u

// Sums over 1024 threads of per-thread element runs: exclusive prefix of the
// thread totals (LDS scratch for 16 waves).
__device__ inline double block_excl_sum(double v, double *total) {
    __shared__ double ws[16];
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave, nw = blockDim.x / kWave;
    double inc = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const double o = __shfl_up(inc, off, kWave);
        if (lane >= off) inc += o;
    }
    const double below = __shfl_up(inc, 1, kWave);
    if (lane == kWave - 1) ws[wid] = inc;
    __syncthreads();
    double wpre = 0., all = 0.;
    for (int w2 = 0; w2 < nw; w2++) {
        if (w2 < wid) wpre += ws[w2];
        all += ws[w2];
    }
    __syncthreads();
    *total = all;
    return (lane == 0) ? wpre : wpre + below;
}

// Top row: h (mean removed), the periodic 1-D problem G y = h solved in
// closed form through its first differences D_r = y_r - y_{r-1}:
// D_{r+1} - D_r = h_r and sum_r D_r = 0, so D_r = S_r - mean(S) with S the
// exclusive prefix sum of h.  uxi = D rdx (mean removed as in the
// reference), u[i][nz-2] = uxi + ubarTopp, and xi = the mean-free y.
__global__ void __launch_bounds__(1024) k_calu_top(Grid G, StepPtrs S) {
    extern __shared__ double sh[];  // N doubles
    const int N = G.nx - 2, nz = G.nz;
    const int E = (N + blockDim.x - 1) / blockDim.x;
    const int r0 = threadIdx.x * E;
    const double *w = S.w, *rhow = S.P.rhow, *rhou = S.P.rhou;
    // h and its mean (src/PoissonSolver.cpp:585-594)
    double loc = 0.;
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        const double h = -(0. - rhow[nz - 2] * w[ix(G, r + 1, nz - 2)]) / rhou[nz - 2] * G.dx;
        sh[r] = h;
        loc += h;
    }
    double tot;
    block_excl_sum(loc, &tot);
    const double hm = tot / (N);
    loc = 0.;
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        sh[r] = sh[r] - hm;
        loc += sh[r];
    }
    // exclusive prefix S_r of h, written over h
    double pre = block_excl_sum(loc, &tot);
    double sloc = 0.;
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        const double h = sh[r];
        sh[r] = pre;
        sloc += pre;
        pre += h;
    }
    block_excl_sum(sloc, &tot);
    const double D0 = -tot / N;
    // uxi = D rdx, then its mean removed (:621-634)
    double usum = 0.;
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        const double D = D0 + sh[r];
        sh[r] = D;
        usum += D * G.rdx;
    }
    block_excl_sum(usum, &tot);
    const double um = tot / (N);
    const double ubp = S.s[S_UBARTOPP];
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        const double uxi = sh[r] * G.rdx - um;
        S.uxi[r + 1] = uxi;
        if (r + 1 == N) S.uxi[0] = uxi;
        if (r == 0) S.uxi[N + 1] = uxi;
        store_center(S.u, G, r + 1, nz - 2, uxi + ubp);
    }
    // xi: y_r = sum_{j<=r} D_j minus its mean (min-norm solution)
    double dl = 0.;
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        dl += (r == 0) ? 0. : sh[r];
    }
    double ypre = block_excl_sum(dl, &tot);
    double ysum = 0.;
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        ypre += (r == 0) ? 0. : sh[r];
        ysum += ypre;
    }
    block_excl_sum(ysum, &tot);
    const double ym = tot / N;
    ypre = block_excl_sum(dl, &tot);
    for (int e = 0; e < E; e++) {
        const int r = r0 + e;
        if (r >= N) break;
        ypre += (r == 0) ? 0. : sh[r];
        const double xi = ypre - ym;
        S.xi[r + 1] = xi;
        if (r + 1 == N) S.xi[0] = xi;
        if (r == 0) S.xi[N + 1] = xi;
    }
}

// Downward integration u[i][k] = u[i][nz-2] - dz sum_{k'=k}^{nz-3}
// ((w[i][k'+1]-w[i-1][k'+1]) rdx - rhow[k'+1] zetap[i][k'+1]) (:643-650):
// UW columns x UC vertical chunks per block (narrow blocks, so that the
// 2-D grid has enough workgroups for every CU), chunk sums scanned in LDS.
constexpr int UW = 16, UC = 64;
__global__ void __launch_bounds__(UW *UC) k_calu_cols(Grid G, StepPtrs S) {
    __shared__ double cs[UC][UW];
    const int i = 1 + blockIdx.x * UW + threadIdx.x, c = threadIdx.y;
    const int K = G.nz - 3;  // levels nz-3 .. 1
    const int C = (K + UC - 1) / UC;
    const bool col = (i <= G.nx - 2);
    const double *w = S.w, *zp = S.fut[F_ZETA], *rhow = S.P.rhow;
    double loc = 0.;
    if (col) {
        for (int kk = c * C; kk < min((c + 1) * C, K); kk++) {
            const int k = G.nz - 3 - kk;
            loc += ((w[ix(G, i, k + 1)] - w[ix(G, i - 1, k + 1)]) * G.rdx - rhow[k + 1] * zp[ix(G, i, k + 1)]) * -G.dz;
        }
    }
    cs[c][threadIdx.x] = loc;
    __syncthreads();
    if (!col) return;
    double area = 0.;
    for (int c2 = 0; c2 < c; c2++) area += cs[c2][threadIdx.x];
    const double utop = S.u[ix(G, i, G.nz - 2)];
    for (int kk = c * C; kk < min((c + 1) * C, K); kk++) {
        const int k = G.nz - 3 - kk;
        area += ((w[ix(G, i, k + 1)] - w[ix(G, i - 1, k + 1)]) * G.rdx - rhow[k + 1] * zp[ix(G, i, k + 1)]) * -G.dz;
        store_center(S.u, G, i, k, area + utop);
    }
}

// LDS budget of the one-workgroup coarse kernel for an FP64-sized budget:
// the FP32 hierarchy takes the levels the FP64 one takes.
template <typename T>
size_t coarse_budget(size_t fp64_bytes) {
    return fp64_bytes * sizeof(T) / sizeof(double);
}

template <typename T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    hipMemset(p, 0, n * sizeof(T));
    return static_cast<T *>(p);
}

template <typename T>
bool build_hier(MGHier<T> &H, int N, int M, const std::vector<double> &cx0, const std::vector<double> &cu0) {
    std::vector<double> cx = cx0, cu = cu0;
    int l = 0;
    while (true) {
        MGLevel<T> &L = H.lev[l];
        L.N = N;
        L.M = M;
        std::vector<T> tcx(M + 2, T(0)), tcu(M + 1), tdi(M + 2, T(0));
        for (int j = 0; j <= M + 1; j++) tcx[j] = T(cx[j]);
        for (int j = 0; j <= M; j++) tcu[j] = T(cu[j]);
        for (int j = 1; j <= M; j++) tdi[j] = T(1. / (2. * cx[j] + cu[j] + cu[j - 1]));
        L.cx = dalloc<T>(M + 2);
        L.cu = dalloc<T>(M + 1);
        L.dinv = dalloc<T>(M + 2);
        if (!L.cx || !L.cu || !L.dinv) return false;
        hipMemcpy(L.cx, tcx.data(), sizeof(T) * (M + 2), hipMemcpyHostToDevice);
        hipMemcpy(L.cu, tcu.data(), sizeof(T) * (M + 1), hipMemcpyHostToDevice);
        hipMemcpy(L.dinv, tdi.data(), sizeof(T) * (M + 2), hipMemcpyHostToDevice);
        if (l > 0) {
            const size_t n = size_t(M + 2) * N;
            L.x = dalloc<T>(n);
            L.b = dalloc<T>(n);
            L.t = dalloc<T>(n);
            if (!L.x || !L.b || !L.t) return false;
        }
        l++;
        const bool can = (M >= 3) && (M % 2 == 1) && (N >= 4) && (N % 2 == 0) && (l < kMaxLevels);
        if (!can) break;
        const int Mc = (M - 1) / 2, Nc = N / 2;
        std::vector<double> ccx(Mc + 2, 0.), ccu(Mc + 1);
        for (int jc = 1; jc <= Mc; jc++) ccx[jc] = cx[2 * jc];
        for (int jc = 0; jc <= Mc; jc++) ccu[jc] = 2. / (1. / cu[2 * jc] + 1. / cu[2 * jc + 1]);
        cx.swap(ccx);
        cu.swap(ccu);
        N = Nc;
        M = Mc;
    }
    H.nlev = l;
    // exact solve of a small coarsest level: dense inverse of its operator
    // (SPD: diagonal 2 cx + cu_j + cu_{j-1}, -cx to the periodic x
    // neighbours, -cu to the z neighbours, Dirichlet rows dropped), by
    // Gauss-Jordan elimination in FP64
    {
        const int n = N * M;
        if (n > 0 && n <= 512) {
            std::vector<double> A(size_t(n) * n, 0.), I(size_t(n) * n, 0.);
            for (int j = 1; j <= M; j++)
                for (int i = 0; i < N; i++) {
                    const int t = (j - 1) * N + i;
                    A[size_t(t) * n + t] += 2. * cx[j] + cu[j] + cu[j - 1];
                    A[size_t(t) * n + (j - 1) * N + (i + N - 1) % N] -= cx[j];
                    A[size_t(t) * n + (j - 1) * N + (i + 1) % N] -= cx[j];
                    if (j + 1 <= M) A[size_t(t) * n + j * N + i] -= cu[j];
                    if (j - 1 >= 1) A[size_t(t) * n + (j - 2) * N + i] -= cu[j - 1];
                }
            for (int t = 0; t < n; t++) I[size_t(t) * n + t] = 1.;
            bool ok = true;
            for (int c = 0; c < n && ok; c++) {
                int piv = c;
                for (int r = c + 1; r < n; r++)
                    if (std::fabs(A[size_t(r) * n + c]) > std::fabs(A[size_t(piv) * n + c])) piv = r;
                if (A[size_t(piv) * n + c] == 0.) {
                    ok = false;
                    break;
                }
                if (piv != c)
                    for (int k = 0; k < n; k++) {
                        std::swap(A[size_t(c) * n + k], A[size_t(piv) * n + k]);
                        std::swap(I[size_t(c) * n + k], I[size_t(piv) * n + k]);
                    }
                const double inv = 1. / A[size_t(c) * n + c];
                for (int k = 0; k < n; k++) {
                    A[size_t(c) * n + k] *= inv;
                    I[size_t(c) * n + k] *= inv;
                }
                for (int r = 0; r < n; r++) {
                    if (r == c) continue;
                    const double fct = A[size_t(r) * n + c];
                    if (fct == 0.) continue;
                    for (int k = 0; k < n; k++) {
                        A[size_t(r) * n + k] -= fct * A[size_t(c) * n + k];
                        I[size_t(r) * n + k] -= fct * I[size_t(c) * n + k];
                    }
                }
            }
            if (ok) {
                std::vector<T> Ti(I.begin(), I.end());
                H.cinv = dalloc<T>(size_t(n) * n);
                if (!H.cinv) return false;
                hipMemcpy(H.cinv, Ti.data(), sizeof(T) * Ti.size(), hipMemcpyHostToDevice);
                H.cinv_n = n;
            }
        }
    }
    // the one-workgroup kernel takes every level that fits in 64 KiB of LDS
    // in FP64 (the same levels in FP32: one more level there, 4x the
    // points, made the one-CU kernel 50% slower than the tiled launch it
    // saves); the finest levels run as tiled down/up kernels
    return mg_plan(H, coarse_budget<T>(size_t(64) * 1024)) == 0;
}

template <typename T>
void free_hier(MGHier<T> &H) {
    for (int l = 0; l < H.nlev; l++) {
        MGLevel<T> &L = H.lev[l];
        hipFree(L.cx);
        hipFree(L.cu);
        hipFree(L.dinv);
        if (l > 0) {
            hipFree(L.x);
            hipFree(L.b);
            hipFree(L.t);
        }
    }
    if (H.cws) hipFree(H.cws);
    H.cws = nullptr;
    if (H.cinv) hipFree(H.cinv);
    H.cinv = nullptr;
    H.cinv_n = 0;
    H.nlev = 0;
}

template <typename T>
bool pcg_fused(const PCG<T> &P, const RedSlot &rs, const double *hpart_dev) {
    return hpart_dev && mg_cg_fusable(P.mg, rs.max_blocks);
}

// The first V-cycle of pcg_loop (on P.r), which a caller may queue before it
// has read the initial residual, so that the GPU does not wait for the host;
// returns what pcg_loop's n_rz_pre takes.
template <typename T>
int pcg_first_vcycle(PCG<T> &P, const RedSlot &rs, const double *hpart_dev, hipStream_t st) {
    const bool fused = pcg_fused(P, rs, hpart_dev);
    // fused: r.z and w.z partials; otherwise r.z into sc[P_RZ]
    const int n = mg_vcycle(P.mg, P.r, P.zA, P.zB, rs, P.sc, true, st, false, fused ? P.w : nullptr);
    return fused ? n : 0;  // unfused: the first pq kernel takes beta = 0 and needs no partials
}

// Preconditioned CG iterations on P (x, r and sc[P_RNORM2] set up by the
// caller; rn_start = that ||w r||^2) until ||w r||^2 <= target or max_iter.
// The convergence test lags one V-cycle behind the GPU so the stream never
// drains: the next V-cycle is queued before the host reads the residual.
// When the residual history predicts that this iteration meets the target
// (geometric extrapolation of the last ratio), the V-cycle is queued only
// after the test, since a speculative V-cycle after the last iteration is
// wasted work; a wrong prediction costs one drained queue instead.  Returns
// the iteration count; *rn2 gets the last residual norm^2 seen.
//
// On a hierarchy with a tiled level 0 (every grid but the smallest) the
// iteration is the single-reduction form: the V-cycle's level-0 up kernel
// leaves r.z and w.z (w = L z), and the level-0 down kernel of the next
// V-cycle forms alpha and beta from them and applies the whole update, so an
// iteration is one V-cycle and nothing else.  Otherwise (a hierarchy that is
// only the coarse kernel) it is standard PCG with its own vector kernels.
template <typename T>
int pcg_loop(PCG<T> &P, double target, int max_iter, bool stall_stop, const RedSlot &rs, double *hs,
             double *hs_dev, const double *hpart, double *hpart_dev, hipEvent_t ev, hipStream_t st,
             double rn_start, double *rn2, bool *conv, int n_rz_pre = -1) {
    const bool fused = pcg_fused(P, rs, hpart_dev);
    const int nb = cg_blocks(P.N, P.M);
    const dim3 b(LX, LY);
    // p.q partials in the last row of the reduction slots, which the update
    // kernel reads while it writes its own partials to row 0
    const RedSlot rs_pq{rs.partial + size_t(7) * rs.max_blocks, rs.max_blocks};
    int n_rz = n_rz_pre >= 0 ? n_rz_pre : pcg_first_vcycle(P, rs, hpart_dev, st);
    int it = 0;
    double best = 1e300;
    int stall = 0;
    double prev = rn_start, prev2 = -1.;
    *conv = false;
    while (it < max_iter) {
        it++;
        const int rz_old = (it & 1) ? P_RZ2 : P_RZ, rz_now = (it & 1) ? P_RZ : P_RZ2;  // it = 1: P_RZ
        int nfused = 0;
        if (fused) {
            const int a_old = (it & 1) ? P_ALPHA2 : P_ALPHA, a_now = (it & 1) ? P_ALPHA : P_ALPHA2;
            const CGFuse<T> fz{P.x,  P.r,   P.rB,   P.pA,  P.q,         P.qB, P.w, P.wD, P.sc, rz_old,
                               rz_now, a_old, a_now, int(it == 1), rs, n_rz, hpart_dev};
            nfused = mg_down0_cg(P.mg, P.zA, fz, rs.max_blocks, st);
            std::swap(P.r, P.rB);
            std::swap(P.q, P.qB);
        } else {
            hipLaunchKernelGGL(k_pcg_pq<T>, dim3(nb), b, 0, st, P.N, P.M, P.zA, P.pA, P.pB, P.q, P.mg.lev[0].cx,
                               P.mg.lev[0].cu, P.sc, rs_pq, int(it == 1), rs, n_rz, rz_old, rz_now);
            std::swap(P.pA, P.pB);
            hipLaunchKernelGGL(k_pcg_update<T>, dim3(nb), b, 0, st, P.N, P.M, P.x, P.r, P.pA, P.q, P.wD, P.sc,
                               rz_now, rs_pq, nb, rs);
            finalize<1>(rs, nb, P.sc, RED_STORE, P_RNORM2, st, -1, hs_dev);
        }
        hipEventRecord(ev, st);
        bool ahead = it < max_iter;
        if (ahead && !stall_stop) {
            // an MG-PCG step cuts ||r||^2 by ~1e-2; the last observed ratio is a
            // better guess once there is one
            const double q = (prev2 > 0.) ? std::min(prev / prev2, 1.) : 1e-2;
            if (prev * q <= target) ahead = false;
        }
        T *wz = fused ? P.w : nullptr;
        if (ahead) n_rz = mg_vcycle(P.mg, P.r, P.zA, P.zB, rs, P.sc, false, st, fused, wz);
        hipEventSynchronize(ev);
        double rn = 0.;
        if (fused) {
            const volatile double *hp = hpart;
            for (int k = 0; k < nfused; k++) rn += hp[k];  // fixed order: deterministic
        } else {
            rn = *static_cast<volatile double *>(hs);
        }
        *rn2 = rn;
        if (!std::isfinite(rn)) break;  // breakdown (NaN or Inf): stop, not converged
        if (!(rn > target)) {
            *conv = true;
            break;
        }
        if (!ahead && it < max_iter) n_rz = mg_vcycle(P.mg, P.r, P.zA, P.zB, rs, P.sc, false, st, fused, wz);
        prev2 = prev;
        prev = rn;
        if (stall_stop) {
            if (rn < 0.25 * best) {
                best = rn;
                stall = 0;
            } else if (++stall >= 3) {
                *conv = true;
                break;
            }
        }
    }
    return it;
}

}  // namespace

int Poisson::init(const Grid &G, const double *rhou, const double *rhow, const RedSlot &rs, bool with_fp32) {
    N = G.nx - 2;
    M = G.nz - 3;
    red = rs;
    std::vector<double> cx(M + 2, 0.), cu(M + 1, 0.);
    for (int j = 1; j <= M; j++) cx[j] = 1. / rhow[j + 1];
    for (int j = 0; j <= M; j++) cu[j] = 1. / rhou[j + 1];
    const size_t n = size_t(M + 2) * N;
    std::vector<double> wd(M + 2, 0.);
    for (int j = 1; j <= M; j++) wd[j] = rhow[j + 1];
    {
        PCG<double> &P = pcg64;
        P.N = N;
        P.M = M;
        if (!build_hier(P.mg, N, M, cx, cu)) return VVM_ERR_NOMEM;
        P.x = dalloc<double>(n);
        P.r = dalloc<double>(n);
        P.rB = dalloc<double>(n);
        P.zA = dalloc<double>(n);
        P.zB = dalloc<double>(n);
        P.pA = dalloc<double>(n);
        P.pB = dalloc<double>(n);
        P.q = dalloc<double>(n);
        P.qB = dalloc<double>(n);
        P.w = dalloc<double>(n);
        P.wD = dalloc<double>(M + 2);
        P.sc = dalloc<double>(NPSCAL);
        if (!P.x || !P.r || !P.rB || !P.zA || !P.zB || !P.pA || !P.pB || !P.q || !P.qB || !P.w || !P.wD || !P.sc)
            return VVM_ERR_NOMEM;
        hipMemcpy(P.wD, wd.data(), sizeof(double) * (M + 2), hipMemcpyHostToDevice);
    }
    have32 = with_fp32;
    if (with_fp32) {
        PCG<float> &P = pcg32;
        P.N = N;
        P.M = M;
        if (!build_hier(P.mg, N, M, cx, cu)) return VVM_ERR_NOMEM;
        P.x = dalloc<float>(n);
        P.r = dalloc<float>(n);
        P.rB = dalloc<float>(n);
        P.zA = dalloc<float>(n);
        P.zB = dalloc<float>(n);
        P.pA = dalloc<float>(n);
        P.pB = dalloc<float>(n);
        P.q = dalloc<float>(n);
        P.qB = dalloc<float>(n);
        P.w = dalloc<float>(n);
        P.wD = dalloc<float>(M + 2);
        P.sc = dalloc<double>(NPSCAL);
        if (!P.x || !P.r || !P.rB || !P.zA || !P.zB || !P.pA || !P.pB || !P.q || !P.qB || !P.w || !P.wD || !P.sc)
            return VVM_ERR_NOMEM;
        std::vector<float> wf(wd.begin(), wd.end());
        hipMemcpy(P.wD, wf.data(), sizeof(float) * (M + 2), hipMemcpyHostToDevice);
        r64 = dalloc<double>(n);
        x32B = dalloc<float>(n);
        if (!r64 || !x32B) return VVM_ERR_NOMEM;
    }
    f = dalloc<double>(n);
    xprev = dalloc<double>(n);
    xprev2 = dalloc<double>(n);
    xprev3 = dalloc<double>(n);
    x64B = dalloc<double>(n);
    if (!f || !xprev || !xprev2 || !xprev3 || !x64B) return VVM_ERR_NOMEM;
    // pinned, coherent and mapped: k_finalize writes the CG residual norm
    // here directly, and the host reads it after an event
    if (hipHostMalloc(reinterpret_cast<void **>(&hscal), 16 * sizeof(double),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return VVM_ERR_NOMEM;
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&hscal_dev), hscal, 0) != hipSuccess) return VVM_ERR_HIP;
    // per-workgroup residual norms of the fused CG update, read by the host
    if (hipHostMalloc(reinterpret_cast<void **>(&hpart), sizeof(double) * size_t(rs.max_blocks),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        return VVM_ERR_NOMEM;
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&hpart_dev), hpart, 0) != hipSuccess) return VVM_ERR_HIP;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return VVM_ERR_HIP;
    if (hipEventCreateWithFlags(&ev_res, hipEventDisableTiming) != hipSuccess) return VVM_ERR_HIP;
    return VVM_OK;
}

template <typename T>
static void free_pcg(PCG<T> &P) {
    free_hier(P.mg);
    for (void *p : {(void *)P.x, (void *)P.r, (void *)P.rB, (void *)P.zA, (void *)P.zB, (void *)P.pA,
                    (void *)P.pB, (void *)P.q, (void *)P.qB, (void *)P.w, (void *)P.wD, (void *)P.sc})
        if (p) hipFree(p);
    P = PCG<T>();
}

void Poisson::destroy() {
    free_pcg(pcg64);
    if (have32) free_pcg(pcg32);
    for (void *p : {(void *)f, (void *)xprev, (void *)xprev2, (void *)xprev3, (void *)x64B, (void *)x32B, (void *)bcg, (void *)bsc,
                    (void *)r64, (void *)gfeat,
                    (void *)gsc, (void *)samples})
        if (p) hipFree(p);
    f = xprev = xprev2 = xprev3 = r64 = x64B = bcg = bsc = nullptr;
    gfeat = samples = x32B = nullptr;
    bn_pending = 0;
    x_pending = 0;
    corr_pending = false;
    gsc = nullptr;
    gscratch.release();
    sample_cap = sample_count = 0;
    net = nullptr;
    if (hscal) hipHostFree(hscal);
    hscal = nullptr;
    hscal_dev = nullptr;
    if (hpart) hipHostFree(hpart);
    hpart = hpart_dev = nullptr;
    if (ev) hipEventDestroy(ev);
    if (ev_res) hipEventDestroy(ev_res);
    ev_res = nullptr;
    ev = nullptr;
}

int Poisson::replan(size_t budget) {
    if (budget > size_t(64) * 1024) return VVM_ERR_ARG;
    if (mg_plan(pcg64.mg, coarse_budget<double>(budget)) != 0) return VVM_ERR_NOMEM;
    if (have32 && mg_plan(pcg32.mg, coarse_budget<float>(budget)) != 0) return VVM_ERR_NOMEM;
    lds_budget = budget;
    return VVM_OK;
}

// ||b||^2 partials go to row 1 of red; the first residual's reduction adds
// them up with its own (flush_pending, if something needs red first)
void Poisson::rhs_from_zetap(const Grid &G, const double *zetap, const double *rhow_dev, hipStream_t st) {
    const RedSlot row1{red.partial + red.max_blocks, red.max_blocks};
    hipLaunchKernelGGL(k_calw_rhs, lgrid(N, M), dim3(LX, LY), 0, st, G, zetap, rhow_dev, f, pcg64.sc, row1);
    bn_pending = nblocks(lgrid(N, M));
}

void Poisson::rhs_from_b(const double *b, hipStream_t st) {
    // the FP64 row weights wD[j] are rhow[j+1]
    const RedSlot row1{red.partial + red.max_blocks, red.max_blocks};
    hipLaunchKernelGGL(k_rhs_from_b, lgrid(N, M), dim3(LX, LY), 0, st, N, M, b, pcg64.wD, f, pcg64.sc, row1);
    bn_pending = nblocks(lgrid(N, M));
}

void Poisson::flush_pending(hipStream_t st) {
    if (bn_pending) {
        const RedSlot row1{red.partial + red.max_blocks, red.max_blocks};
        finalize<1>(row1, bn_pending, pcg64.sc, RED_STORE, P_BNORM2, st, -1, hscal_dev + P_BNORM2);
        bn_pending = 0;
    }
    const size_t n = size_t(M + 2) * N;
    if (x_pending == XF_ZERO)
        hipMemsetAsync(pcg64.x, 0, n * sizeof(double), st);
    else if (x_pending == XF_EXTRAP)
        hipLaunchKernelGGL(k_extrapolate, dim3((n + 255) / 256), dim3(256), 0, st, n, pcg64.x, xprev, xprev2, xprev3,
                           x_pending_ord);
    x_pending = XF_NONE;
}

void Poisson::first_residual(double *r, bool fp32_rhs, hipStream_t st) {
    const dim3 g = lgrid(N, M), b(LX, LY);
    PCG<double> &D = pcg64;
    const float *e32 = have32 ? pcg32.x : nullptr;
    const int xf = corr_pending ? int(XF_CORR) : x_pending;
    double *xo = (xf == XF_CORR) ? x64B : D.x;
    // the FP32 copy of r, and the FP32 x zeroed for the next correction; the
    // correction being applied is read from pcg32.x, so the zeros then go to
    // the other FP32 buffer
    float *r32 = fp32_rhs ? pcg32.r : nullptr;
    float *x32 = fp32_rhs ? ((xf == XF_CORR) ? x32B : pcg32.x) : nullptr;
#define VVM_RESIDUAL_X(XF)                                                                                        \
    hipLaunchKernelGGL((k_residual_x<XF>), g, b, 0, st, N, M, f, D.x, xo, xprev, xprev2, xprev3, x_pending_ord, e32, r, D.mg.lev[0].cx, \
                       D.mg.lev[0].cu, D.wD, red, r32, x32)
    switch (xf) {
        case XF_EXTRAP: VVM_RESIDUAL_X(XF_EXTRAP); break;
        case XF_ZERO: VVM_RESIDUAL_X(XF_ZERO); break;
        case XF_CORR: VVM_RESIDUAL_X(XF_CORR); break;
        default: VVM_RESIDUAL_X(XF_NONE); break;
    }
#undef VVM_RESIDUAL_X
    if (xf == XF_CORR) std::swap(D.x, x64B);
    if (fp32_rhs && xf == XF_CORR) std::swap(pcg32.x, x32B);
    x_pending = XF_NONE;
    corr_pending = false;
    // the norms land in mapped host memory (hscal) directly: no copy in the stream
    if (bn_pending) {
        // rows 0 (r) and 1 (b), same grid: sc[P_RNORM2], sc[P_BNORM2]
        finalize<2>(red, nblocks(g), D.sc, RED_STORE, P_RNORM2, st, -1, hscal_dev + P_RNORM2);
        bn_pending = 0;
    } else {
        finalize<1>(red, nblocks(g), D.sc, RED_STORE, P_RNORM2, st, -1, hscal_dev + P_RNORM2);
    }
}

void Poisson::x_to_ref(double *xref, hipStream_t st) {
    hipLaunchKernelGGL(k_x_ref, lgrid(N, M), dim3(LX, LY), 0, st, N, M, pcg64.x, xref, 1);
}

void Poisson::x_from_ref(const double *xref, hipStream_t st) {
    hipLaunchKernelGGL(k_x_ref, lgrid(N, M), dim3(LX, LY), 0, st, N, M, pcg64.x, const_cast<double *>(xref), 0);
}

void Poisson::back_to_w(const Grid &G, double *w, const double *rhow_dev, hipStream_t st) {
    hipLaunchKernelGGL(k_calw_back, dim3((G.nx + LX - 1) / LX, (G.nz + LY - 1) / LY), dim3(LX, LY), 0, st, G, pcg64.x,
                       w, rhow_dev);
}

void Poisson::set_guess(int mode, hipStream_t st) {
    guess_from_net = false;
    // formed by the solve's first residual kernel (first_residual)
    x_pending = XF_NONE;
    if (mode == VVM_GUESS_ZERO) {
        x_pending = XF_ZERO;
    } else if ((mode == VVM_GUESS_EXTRAP || mode == VVM_GUESS_EXTRAP2) && nsolves >= 2) {
        x_pending = XF_EXTRAP;
        x_pending_ord = (mode == VVM_GUESS_EXTRAP2) ? xe_order() : 1;
    }
    // VVM_GUESS_PREV, and the extrapolations before two solutions exist,
    // start from the last solution (x holds the oldest one after
    // remember_solution); the learned guess is written by its own kernels.
    if (x_pending == XF_NONE && mode != VVM_GUESS_ZERO && nsolves >= 1)
        hipMemcpyAsync(pcg64.x, xprev, size_t(M + 2) * N * sizeof(double), hipMemcpyDeviceToDevice, st);
}

void Poisson::remember_solution(hipStream_t st) {
    flush_pending(st);  // nothing is pending after a solve; a guess set without one is formed now
    // x_{t-3} <- x_{t-2} <- x_{t-1} <- x by rotating the buffers, no copy:
    // the oldest becomes the next solve's x, which set_guess fills (or the
    // first residual kernel, or the network, rewrites in full)
    double *oldest = xprev3;
    xprev3 = xprev2;
    xprev2 = xprev;
    xprev = pcg64.x;
    pcg64.x = oldest;
    nsolves++;
}

template <typename T>
void Poisson::start_solve(double *r, bool fp32_rhs, PCG<T> *pre_vcycle, int *n_rz_pre, double *bn2, double *rn2,
                          hipStream_t st) {
    const bool judge = guess_from_net;
    guess_from_net = false;
    for (int attempt = 0;; attempt++) {
        first_residual(r, fp32_rhs, st);
        hipEventRecord(ev_res, st);
        if (pre_vcycle) *n_rz_pre = pcg_first_vcycle(*pre_vcycle, red, hpart_dev, st);
        hipEventSynchronize(ev_res);
        *bn2 = hscal[P_BNORM2];
        *rn2 = hscal[P_RNORM2];
        if (!judge || attempt > 0) return;
        const double rn2_xe = hscal[H_GUESS + 3];
        if (std::isfinite(*rn2) && *rn2 <= rn2_xe) {
            net_backoff = 0;
            return;
        }
        net_backoff = std::min(std::max(1, 4 * net_backoff), kMaxNetSkip);
        net_skip = net_backoff;
        // the network's guess starts worse than the extrapolation it corrects
        // (or is not finite): start from the extrapolation instead; the first
        // V-cycle queued on the rejected residual is simply overwritten
        guess_fallbacks++;
        x_pending = XF_EXTRAP;
        x_pending_ord = xe_order();
    }
}

SolveResult Poisson::solve_fp64(double tol, int max_iter, hipStream_t st) {
    SolveResult R;
    PCG<double> &P = pcg64;
    // the first V-cycle goes in before the host reads the residual (a guess
    // that already meets the stop rule wastes it, which is rare)
    int n_rz_pre = 0;
    double bn2 = 0., rn2 = 0.;
    start_solve(P.r, false, &P, &n_rz_pre, &bn2, &rn2, st);
    if (!std::isfinite(bn2) || !std::isfinite(rn2)) {  // a non-finite right-hand side or guess
        R.rel0 = R.rel = std::sqrt(rn2 / bn2);
        return R;
    }
    if (bn2 == 0.) {
        hipMemsetAsync(P.x, 0, size_t(M + 2) * N * sizeof(double), st);
        R.converged = true;
        return R;
    }
    R.rel0 = std::sqrt(rn2 / bn2);
    (void)P;
    const bool parity = (tol <= 0.);
    const double target = parity ? 1e-30 * bn2 : tol * tol * bn2;
    if (!(rn2 > target)) {
        R.rel = R.rel0;
        R.converged = true;
        return R;
    }
    double last = rn2;
    bool conv = false;
    R.iters = pcg_loop(P, target, max_iter, parity, red, hscal + 8, hscal_dev + 8, hpart, hpart_dev, ev, st, rn2, &last,
                       &conv, n_rz_pre);
    R.rel = std::sqrt(last / bn2);
    R.converged = conv;
    return R;
}

// smallest relative residual reduction asked of one FP32 inner solve
constexpr double kInnerFloor = 1e-6;

SolveResult Poisson::solve_mixed(double tol, int max_outer, int max_inner, double inner_tol, hipStream_t st) {
    SolveResult R;
    if (!have32) return solve_fp64(tol, max_inner * max_outer, st);
    PCG<double> &D = pcg64;
    PCG<float> &F = pcg32;
    double bn2 = -1.;
    for (int outer = 0; outer <= max_outer; outer++) {
        // r64, and its FP32 copy for the correction solve (unused if converged);
        // on the first sweep the inner solve's first V-cycle goes in before
        // the host reads the residual (the guess almost never meets the stop
        // rule); on later sweeps the residual is usually the final check
        int n_rz_pre = -1;
        double rn2 = 0.;
        start_solve(r64, true, outer == 0 ? &F : static_cast<PCG<float> *>(nullptr), &n_rz_pre, &bn2, &rn2, st);
        if (!std::isfinite(bn2) || !std::isfinite(rn2)) {  // a non-finite right-hand side, guess or correction
            R.rel = std::sqrt(rn2 / bn2);
            if (outer == 0) R.rel0 = R.rel;
            return R;
        }
        if (bn2 == 0.) {
            hipMemsetAsync(D.x, 0, size_t(M + 2) * N * sizeof(double), st);
            R.converged = true;
            return R;
        }
        const double rel = std::sqrt(rn2 / bn2);
        if (outer == 0) R.rel0 = rel;
        R.rel = rel;
        if (rel <= tol) {
            R.converged = true;
            return R;
        }
        if (outer == max_outer) break;
        R.refine++;
        // The inner solve reduces its residual just enough to meet tol (with
        // a margin for the FP64 residual after the correction), by no more
        // than FP32 resolves and by at least inner_tol: one refinement when
        // the guess is good, instead of a fixed 1e-4 per refinement that
        // overshoots the target.  Its starting norm is the FP64 residual's
        // (the FP32 copy differs in the 8th digit), so no read-back is needed.
        const double need = 0.3 * tol / rel;
        const double itol = std::min(inner_tol, std::max(need, kInnerFloor));
        const double r0 = rn2;
        double last = r0;
        bool conv = false;
        R.iters += pcg_loop(F, itol * itol * r0, max_inner, false, red, hscal + 8, hscal_dev + 8, hpart, hpart_dev, ev, st, r0, &last,
                            &conv, n_rz_pre);
        corr_pending = true;  // x += e, in the next sweep's residual kernel
    }
    (void)D;
    return R;
}

// ---------------------------------------------------------------- cal_u

bool cal_u_supported(const Grid &G) { return G.nx - 2 <= 8192; }

void launch_cal_u(const Grid &G, const StepPtrs &S, hipStream_t st) {
    const int N = G.nx - 2;
    hipLaunchKernelGGL(k_calu_top, dim3(1), dim3(1024), size_t(N) * sizeof(double), st, G, S);
    hipLaunchKernelGGL(k_calu_cols, dim3((N + UW - 1) / UW), dim3(UW, UC), 0, st, G, S);
}

}  // namespace vvmd
