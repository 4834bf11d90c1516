// k_mg.hip -- the multigrid V(2,2) preconditioner of the elliptic solve.
//
// Levels hold (M+2) x N arrays, x-contiguous, rows 0 and M+1 the Dirichlet
// zeros, columns periodic.  One V-cycle is:
//   down(l), l = 0..Lc-1: two damped-Jacobi sweeps from zero on b_l (the
//     first is pointwise), residual, restriction P^T to b_{l+1}; the
//     smoothed x_l is stored for the way up.  One kernel per level: an LDS
//     tile of the fine points a block of coarse points needs, with a 3-point
//     halo (restriction 1 + residual 1 + second sweep 1).
//   coarse(Lc..L-1): every remaining level in one workgroup, all arrays in
//     LDS (the same steps, and damped-Jacobi iterations on the coarsest).
//   up(l), l = Lc-1..0: x_l += P x_{l+1} (bilinear), two Jacobi sweeps; one
//     kernel per level with a 2-point halo.  On level 0 the kernel also
//     writes the partial sums of r.z for the CG.
// Jacobi sweeps, P and P^T are symmetric in combination, so the cycle is a
// symmetric positive definite preconditioner.
#include "mg.h"

#include <algorithm>

namespace vvmd {

namespace {

constexpr int CX = 32, CY = 4;  // coarse points per block of the down kernel
constexpr int TX = 64, TY = 4;  // fine points per block of the up kernel
// the level-0 kernels of the single-reduction CG carry more halo (and more
// vectors), so they take taller tiles
constexpr int CYG = 8, TYG = 8;
constexpr int DW = 2 * CX + 5;

__host__ __device__ inline int wrapi(int i, int N) {
    i %= N;
    return i < 0 ? i + N : i;
}

template <typename T>
__device__ inline T lap(const T *X, int W, int t, T cxj, T cuj, T cujm) {
    // operator L on an LDS tile (row stride W) at linear index t
    const T c = X[t];
    return cxj * (T(2) * c - X[t - 1] - X[t + 1]) + cuj * (c - X[t + W]) + cujm * (c - X[t - W]);
}

// CG: the right-hand side is the CG residual after its update, formed here
// (cg, see CGFuse); otherwise b.
template <typename T, bool CG>
__global__ void __launch_bounds__(256) k_mg_down(int N, int M, const T *__restrict__ b, T *__restrict__ x2out,
                                                 const T *__restrict__ cx, const T *__restrict__ cu,
                                                 const T *__restrict__ dinv, T omega, int Nc, int Mc,
                                                 T *__restrict__ bc, CGFuse<T> cg) {
    constexpr int CYk = CG ? CYG : CY, DH = 2 * CYk + 5;
    __shared__ T Bs[DH * DW], X1[DH * DW], X2[DH * DW];
    __shared__ T Rx[DH], Ru[DH], Rd[DH];  // cx, cu, dinv of the tile's rows
    T alpha = T(0), beta = T(0);
    if (CG) {
        // gamma = r.z, delta = w.z of the last V-cycle; beta = gamma /
        // gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
        const double gam = block_total(cg.rs_gd, cg.ngd);
        const double del = block_total(RedSlot{cg.rs_gd.partial + cg.rs_gd.max_blocks, cg.rs_gd.max_blocks}, cg.ngd);
        double a, bt = 0.;
        if (cg.first) {
            a = gam / del;
        } else {
            bt = gam / cg.sc[cg.g_old];
            a = gam / (del - bt * gam / cg.sc[cg.a_old]);
        }
        alpha = T(a);
        beta = T(bt);
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
            cg.sc[cg.g_new] = gam;
            cg.sc[cg.a_new] = a;
        }
    }
    const int ic0 = blockIdx.x * CX, jc0 = 1 + blockIdx.y * CYk;
    const int ic1 = min(ic0 + CX - 1, Nc - 1), jc1 = min(jc0 + CYk - 1, Mc);
    const int fi0 = 2 * ic0 - 3, fj0 = 2 * jc0 - 3;
    const int W = 2 * (ic1 - ic0) + 7, H = 2 * (jc1 - jc0) + 7;
    const int n = W * H;
    if (threadIdx.x < H) {  // loaded with the tile, so later steps read LDS only
        const int j = fj0 + threadIdx.x;
        const bool in = (j >= 1 && j <= M);
        Rx[threadIdx.x] = in ? cx[j] : T(0);
        Ru[threadIdx.x] = (j >= 0 && j <= M) ? cu[j] : T(0);
        Rd[threadIdx.x] = in ? dinv[j] : T(0);
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int tj = t / W, ti = t % W;
        const int j = fj0 + tj, i = wrapi(fi0 + ti, N);
        const bool in = (j >= 1 && j <= M);
        T bv = T(0);
        if (in) {
            const size_t a = size_t(j) * N + i;
            if (CG) {
                // q = w + beta q_old (= L p), r = r_old - alpha q; q is
                // stored by the workgroup that owns the point
                const T qn = cg.first ? cg.w[a] : cg.w[a] + beta * cg.qin[a];
                bv = cg.rin[a] - alpha * qn;
                const int oi = i - 2 * ic0, oj = j - (2 * jc0 - 1);
                const int jtop = (jc1 == Mc) ? 2 * jc1 + 1 : 2 * jc1;
                if (fi0 + ti == i && oi >= 0 && oi < 2 * (ic1 - ic0) + 2 && oj >= 0 && j <= jtop) cg.qout[a] = qn;
            } else {
                bv = b[a];
            }
        }
        Bs[tj * DW + ti] = bv;
        X1[tj * DW + ti] = in ? omega * dinv[j] * bv : T(0);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int tj = t / W, ti = t % W;
        const int j = fj0 + tj;
        T v = T(0);
        if (ti >= 1 && ti <= W - 2 && tj >= 1 && tj <= H - 2 && j >= 1 && j <= M) {
            const int s = tj * DW + ti;
            v = X1[s] + omega * Rd[tj] * (Bs[s] - lap(X1, DW, s, Rx[tj], Ru[tj], Ru[tj - 1]));
        }
        X2[tj * DW + ti] = v;
    }
    __syncthreads();
    // smoothed fine values this block owns: columns 2ic0..2ic1+1, rows
    // 2jc0-1..2jc1 (and row M with the top coarse row)
    const int jtop = (jc1 == Mc) ? 2 * jc1 + 1 : 2 * jc1;
    const int ow = 2 * (ic1 - ic0) + 2, oh = jtop - (2 * jc0 - 1) + 1;
    double part[1] = {0.};
    for (int t = threadIdx.x; t < ow * oh; t += blockDim.x) {
        const int oj = t / ow, oi = t % ow;
        const int i = 2 * ic0 + oi, j = 2 * jc0 - 1 + oj;
        if (j >= 1 && j <= M) {
            const size_t a = size_t(j) * N + i;
            const int s = (j - fj0) * DW + (i - fi0);
            if (CG) {  // the CG update on the points this workgroup owns
                // p = z + beta p_old; z (in x2out's buffer) is read before
                // the smoothed value below replaces it
                const T pn = cg.first ? x2out[a] : x2out[a] + beta * cg.p[a];
                cg.p[a] = pn;
                cg.x[a] = cg.x[a] + alpha * pn;
                const T rn = Bs[s];
                cg.rout[a] = rn;
                const double wr = double(cg.wD[j]) * double(rn);
                part[0] += wr * wr;
            }
            x2out[a] = X2[s];
        }
    }
    if (CG) {
        block_sum<1>(part);  // two barriers: also orders the LDS reads below
        if (threadIdx.x == 0) {
            const int bid = blockIdx.x + gridDim.x * blockIdx.y;
            __hip_atomic_store(cg.hpart + bid, part[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    // residual on the restriction footprint, into X1
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int tj = t / W, ti = t % W;
        const int j = fj0 + tj;
        if (ti >= 2 && ti <= W - 3 && tj >= 2 && tj <= H - 3) {
            const int s = tj * DW + ti;
            X1[s] = (j >= 1 && j <= M) ? Bs[s] - lap(X2, DW, s, Rx[tj], Ru[tj], Ru[tj - 1]) : T(0);
        }
    }
    __syncthreads();
    const int cw = ic1 - ic0 + 1, ch = jc1 - jc0 + 1;
    for (int t = threadIdx.x; t < cw * ch; t += blockDim.x) {
        const int ic = ic0 + t % cw, jc = jc0 + t / cw;
        const int s = (2 * jc - fj0) * DW + (2 * ic - fi0);
        const T r = X1[s] + T(0.5) * (X1[s - 1] + X1[s + 1] + X1[s - DW] + X1[s + DW]) +
                    T(0.25) * (X1[s - DW - 1] + X1[s - DW + 1] + X1[s + DW - 1] + X1[s + DW + 1]);
        bc[size_t(jc) * Nc + ic] = r;
    }
}

// Bilinear prolongation of the coarse correction at fine (i, j).
template <typename T>
__device__ inline T prolong(const T *__restrict__ xc, int Nc, int i, int j) {
    const int ic = i >> 1, jc = j >> 1;
    const int ic1 = (ic + 1 == Nc) ? 0 : ic + 1;
    const T *r0 = xc + size_t(jc) * Nc, *r1 = r0 + Nc;
    if (!(i & 1) && !(j & 1)) return r0[ic];
    if ((i & 1) && !(j & 1)) return T(0.5) * (r0[ic] + r0[ic1]);
    if (!(i & 1)) return T(0.5) * (r0[ic] + r1[ic]);
    return T(0.25) * (r0[ic] + r0[ic1] + r1[ic] + r1[ic1]);
}

// CGW (level 0 of the single-reduction CG): the tile carries a 3-point
// halo, the second sweep runs on the owned points and a 1-point ring around
// them (kept in LDS), and the kernel also writes w = L z and the partials of
// r.z and w.z (rows 0 and 1 of rs).  The ring values are the ones the
// neighbouring workgroups compute for their own points, by the same code.
template <typename T, bool CGW>
__global__ void __launch_bounds__(256) k_mg_up(int N, int M, const T *__restrict__ b, const T *__restrict__ x2,
                                               T *__restrict__ xout, const T *__restrict__ cx,
                                               const T *__restrict__ cu, const T *__restrict__ dinv, T omega, int Nc,
                                               const T *__restrict__ xc, RedSlot rs, int dot, T *__restrict__ wout) {
    constexpr int TYk = CGW ? TYG : TY, HL = CGW ? 3 : 2, W = TX + 2 * HL, H = TYk + 2 * HL;
    __shared__ T Xs[H * W], X3[H * W], Bs[H * W];
    __shared__ T Rx[H], Ru[H], Rd[H];  // cx, cu, dinv of the tile's rows
    const int i0 = blockIdx.x * TX, j0 = 1 + blockIdx.y * TYk;
    const int fi0 = i0 - HL, fj0 = j0 - HL;
    const int n = W * H;
    if (threadIdx.x < H) {  // loaded with the tile, so later steps read LDS only
        const int j = fj0 + threadIdx.x;
        const bool in = (j >= 1 && j <= M);
        Rx[threadIdx.x] = in ? cx[j] : T(0);
        Ru[threadIdx.x] = (j >= 0 && j <= M) ? cu[j] : T(0);
        Rd[threadIdx.x] = in ? dinv[j] : T(0);
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int tj = t / W, ti = t % W;
        const int j = fj0 + tj, i = wrapi(fi0 + ti, N);
        const bool in = (j >= 1 && j <= M);
        T v = T(0), bv = T(0);
        if (in) {
            const size_t p = size_t(j) * N + i;
            v = x2[p];
            if (xc) v += prolong(xc, Nc, i, j);
            bv = b[p];
        }
        Xs[t] = v;
        Bs[t] = bv;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int tj = t / W, ti = t % W;
        const int j = fj0 + tj;
        T v = T(0);
        if (ti >= 1 && ti <= W - 2 && tj >= 1 && tj <= H - 2 && j >= 1 && j <= M)
            v = Xs[t] + omega * Rd[tj] * (Bs[t] - lap(Xs, W, t, Rx[tj], Ru[tj], Ru[tj - 1]));
        X3[t] = v;
    }
    __syncthreads();
    double part[2] = {0., 0.};
    if (!CGW) {
        for (int t = threadIdx.x; t < TX * TYk; t += blockDim.x) {
            const int i = i0 + t % TX, j = j0 + t / TX;
            if (i >= N || j > M) continue;
            const int s = (j - fj0) * W + (i - fi0), tj = j - fj0;
            const T z = X3[s] + omega * Rd[tj] * (Bs[s] - lap(X3, W, s, Rx[tj], Ru[tj], Ru[tj - 1]));
            xout[size_t(j) * N + i] = z;
            part[0] += double(Bs[s]) * double(z);
        }
        double p1[1] = {part[0]};
        if (dot) block_partials<1>(p1, rs);
        return;
    }
    // second sweep on the owned points and their 1-point ring, into Xs
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int tj = t / W, ti = t % W;
        const int j = fj0 + tj;
        if (ti >= 2 && ti <= W - 3 && tj >= 2 && tj <= H - 3)
            Xs[t] = (j >= 1 && j <= M) ? X3[t] + omega * Rd[tj] * (Bs[t] - lap(X3, W, t, Rx[tj], Ru[tj], Ru[tj - 1]))
                                       : T(0);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < TX * TYk; t += blockDim.x) {
        const int i = i0 + t % TX, j = j0 + t / TX;
        if (i >= N || j > M) continue;
        const int s = (j - fj0) * W + (i - fi0), tj = j - fj0;
        const T z = Xs[s];
        const T w = lap(Xs, W, s, Rx[tj], Ru[tj], Ru[tj - 1]);
        const size_t a = size_t(j) * N + i;
        xout[a] = z;
        wout[a] = w;
        part[0] += double(Bs[s]) * double(z);
        part[1] += double(w) * double(z);
    }
    block_partials<2>(part, rs);
}

// All levels from Lc down, in one workgroup with every array in LDS.
template <typename T, bool GWS>
__global__ void __launch_bounds__(1024) k_mg_coarse(CoarseParams<T> cp, const T *__restrict__ bin,
                                                    T *__restrict__ xout, T *ws) {
    extern __shared__ unsigned char lds_raw[];
    // GWS: the levels live in ws, a global workspace of the same layout,
    // because they exceed LDS (__syncthreads orders global memory within the
    // workgroup as well).  The two cases are separate instantiations so that
    // the LDS one addresses LDS directly (ds_read/ds_write) rather than
    // through generic pointers, whose flat accesses wait on both the memory
    // and the LDS counters.
    T *lds = GWS ? ws : reinterpret_cast<T *>(lds_raw);
    const T om = cp.omega;
    const int L = cp.nlev;
    auto B = [&](int l) { return lds + cp.offB[l]; };
    auto X = [&](int l) { return lds + cp.offX[l]; };
    // the level coefficients, staged here once: every step below then reads
    // only LDS, with no global-memory latency between two barriers
    auto CX = [&](int l) { return lds + cp.offC[l]; };
    auto CU = [&](int l) { return lds + cp.offC[l] + (cp.M[l] + 2); };
    auto DI = [&](int l) { return lds + cp.offC[l] + 2 * (cp.M[l] + 2); };
    T *Tm = lds + cp.offT;
    // b of the first level; zero every x (its Dirichlet rows stay zero)
    {
        const int N = cp.N[0], M = cp.M[0];
        for (int t = threadIdx.x; t < (M + 2) * N; t += blockDim.x) {
            const int j = t / N;
            B(0)[t] = (j >= 1 && j <= M) ? bin[t] : T(0);
        }
        for (int l = 0; l < L; l++)
            for (int t = threadIdx.x; t < (cp.M[l] + 2) * cp.N[l]; t += blockDim.x) X(l)[t] = T(0);
        for (int t = threadIdx.x; t < cp.Tsize; t += blockDim.x) Tm[t] = T(0);
        if (cp.offI >= 0)
            for (int t = threadIdx.x; t < cp.cinv_n * cp.cinv_n; t += blockDim.x) lds[cp.offI + t] = cp.cinv[t];
        for (int l = 0; l < L; l++) {
            const int m2 = cp.M[l] + 2;
            for (int t = threadIdx.x; t < m2; t += blockDim.x) {
                CX(l)[t] = cp.cx[l][t];
                if (t < m2 - 1) CU(l)[t] = cp.cu[l][t];
                DI(l)[t] = cp.dinv[l][t];
            }
        }
    }
    __syncthreads();
    auto sweep = [&](int l, const T *xin, T *xo) {
        const int N = cp.N[l], M = cp.M[l];
        const T *cx = CX(l), *cu = CU(l), *di = DI(l), *b = B(l);
        for (int t = threadIdx.x; t < M * N; t += blockDim.x) {
            const int j = 1 + t / N, i = t % N;
            const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
            const int s = j * N + i;
            const T c = xin[s];
            const T Lx = cx[j] * (T(2) * c - xin[j * N + im] - xin[j * N + ip]) + cu[j] * (c - xin[s + N]) +
                         cu[j - 1] * (c - xin[s - N]);
            xo[s] = c + om * di[j] * (b[s] - Lx);
        }
        // Dirichlet rows of the output (Tm is shared by levels of other sizes)
        for (int t = threadIdx.x; t < N; t += blockDim.x) {
            xo[t] = T(0);
            xo[(M + 1) * N + t] = T(0);
        }
        __syncthreads();
    };
    for (int l = 0; l < L - 1; l++) {
        const int N = cp.N[l], M = cp.M[l];
        for (int t = threadIdx.x; t < M * N; t += blockDim.x) {
            const int j = 1 + t / N, s = j * N + t % N;
            X(l)[s] = om * DI(l)[j] * B(l)[s];
        }
        __syncthreads();
        sweep(l, X(l), Tm);
        // residual into X(l) (x2 stays in Tm), then restriction
        {
            const T *cx = CX(l), *cu = CU(l), *b = B(l);
            for (int t = threadIdx.x; t < M * N; t += blockDim.x) {
                const int j = 1 + t / N, i = t % N;
                const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
                const int s = j * N + i;
                const T c = Tm[s];
                const T Lx = cx[j] * (T(2) * c - Tm[j * N + im] - Tm[j * N + ip]) + cu[j] * (c - Tm[s + N]) +
                             cu[j - 1] * (c - Tm[s - N]);
                X(l)[s] = b[s] - Lx;
            }
            __syncthreads();
            const int Nc = cp.N[l + 1], Mc = cp.M[l + 1];
            for (int t = threadIdx.x; t < Mc * Nc; t += blockDim.x) {
                const int jc = 1 + t / Nc, ic = t % Nc;
                const int j = 2 * jc, i = 2 * ic;
                const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
                const T *R = X(l);
                const T r = R[j * N + i] +
                            T(0.5) * (R[j * N + im] + R[j * N + ip] + R[(j - 1) * N + i] + R[(j + 1) * N + i]) +
                            T(0.25) * (R[(j - 1) * N + im] + R[(j - 1) * N + ip] + R[(j + 1) * N + im] +
                                       R[(j + 1) * N + ip]);
                B(l + 1)[jc * Nc + ic] = r;
            }
            // x2 back into X(l) for the way up
            __syncthreads();
            for (int t = threadIdx.x; t < M * N; t += blockDim.x) {
                const int s = (1 + t / N) * N + t % N;
                X(l)[s] = Tm[s];
            }
            __syncthreads();
        }
    }
    {  // coarsest: exact (dense inverse) when small, else damped Jacobi from zero
        const int l = L - 1;
        const int N = cp.N[l], M = cp.M[l], n = N * M;
        if (cp.cinv && cp.cinv_n == n) {
            const T *ci = (cp.offI >= 0) ? lds + cp.offI : cp.cinv;
            for (int t = threadIdx.x; t < n; t += blockDim.x) {
                const T *row = ci + size_t(t) * n;
                T s = T(0);
                for (int q = 0; q < n; q++) s += row[q] * B(l)[N + q];  // rows 1..M start at offset N
                X(l)[N + t] = s;
            }
            __syncthreads();
        } else {
            for (int it = 0; it < cp.coarse_iters; it += 2) {
                sweep(l, X(l), Tm);
                sweep(l, Tm, X(l));
            }
        }
    }
    for (int l = L - 2; l >= 0; l--) {
        const int N = cp.N[l], M = cp.M[l], Nc = cp.N[l + 1];
        const T *xc = X(l + 1);
        for (int t = threadIdx.x; t < M * N; t += blockDim.x) {
            const int j = 1 + t / N, i = t % N;
            const int ic = i >> 1, jc = j >> 1, ic1 = (ic + 1 == Nc) ? 0 : ic + 1;
            const T *r0 = xc + jc * Nc, *r1 = r0 + Nc;
            T v;
            if (!(i & 1) && !(j & 1)) v = r0[ic];
            else if ((i & 1) && !(j & 1)) v = T(0.5) * (r0[ic] + r0[ic1]);
            else if (!(i & 1)) v = T(0.5) * (r0[ic] + r1[ic]);
            else v = T(0.25) * (r0[ic] + r0[ic1] + r1[ic] + r1[ic1]);
            X(l)[j * N + i] += v;
        }
        __syncthreads();
        sweep(l, X(l), Tm);
        sweep(l, Tm, X(l));
    }
    const int N = cp.N[0], M = cp.M[0];
    for (int t = threadIdx.x; t < M * N; t += blockDim.x) {
        const int s = (1 + t / N) * N + t % N;
        xout[s] = X(0)[s];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_dot(int n, const T *a, const T *b, RedSlot rs) {
    double part[1] = {0.};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        part[0] += double(a[t]) * double(b[t]);
    block_partials<1>(part, rs);
}

}  // namespace

template <typename T>
int mg_plan(MGHier<T> &H, size_t lds_budget) {
    if (H.cws) hipFree(H.cws);  // a previous plan's global workspace
    H.cws = nullptr;
    // smallest start level whose remaining levels fit in LDS
    H.coarse_from = H.nlev;
    for (int lc = H.nlev - 1; lc >= 0; lc--) {
        size_t tot = 0, tmax = 0;
        for (int l = lc; l < H.nlev; l++) {
            const size_t s = size_t(H.lev[l].M + 2) * H.lev[l].N;
            tot += 2 * s + 3 * size_t(H.lev[l].M + 2);
            tmax = std::max(tmax, s);
        }
        if ((tot + tmax) * sizeof(T) > lds_budget || H.nlev - lc > kMaxCoarse) break;
        H.coarse_from = lc;
    }
    // grids whose z size stops the coarsening early: the coarsest level
    // alone, from global memory
    const bool global = (H.coarse_from >= H.nlev);
    if (global) H.coarse_from = H.nlev - 1;
    CoarseParams<T> &cp = H.cp;
    cp = CoarseParams<T>();
    cp.nlev = H.nlev - H.coarse_from;
    int off = 0, tmax = 0;
    for (int q = 0; q < cp.nlev; q++) {
        const MGLevel<T> &L = H.lev[H.coarse_from + q];
        const int s = (L.M + 2) * L.N;
        cp.N[q] = L.N;
        cp.M[q] = L.M;
        cp.cx[q] = L.cx;
        cp.cu[q] = L.cu;
        cp.dinv[q] = L.dinv;
        cp.offB[q] = off;
        off += s;
        cp.offX[q] = off;
        off += s;
        tmax = std::max(tmax, s);
    }
    for (int q = 0; q < cp.nlev; q++) {
        cp.offC[q] = off;
        off += 3 * (cp.M[q] + 2);
    }
    cp.offT = off;
    cp.Tsize = tmax;
    cp.omega = H.omega;
    cp.cinv = H.cinv;
    cp.cinv_n = H.cinv_n;
    cp.coarse_iters = H.coarse_iters;
    // a small dense inverse (up to 32 x 32) is staged in LDS with the rest
    cp.offI = -1;
    if (H.cinv && H.cinv_n <= 32 && !global &&
        size_t(off + tmax + H.cinv_n * H.cinv_n) * sizeof(T) <= 64 * 1024) {
        cp.offI = off + tmax;
        tmax += H.cinv_n * H.cinv_n;
    }
    H.coarse_lds = size_t(off + tmax) * sizeof(T);
    if (global) {
        if (hipMalloc(reinterpret_cast<void **>(&H.cws), H.coarse_lds) != hipSuccess) {
            H.cws = nullptr;
            return -1;
        }
        H.coarse_lds = 0;
    }
    return 0;
}

template <typename T>
bool mg_cg_fusable(const MGHier<T> &H, int max_blocks) {
    if (H.coarse_from < 1 || H.nlev < 2) return false;
    const MGLevel<T> &F = H.lev[0];
    const MGLevel<T> &C = H.lev[1];
    const long nd = long((C.N + CX - 1) / CX) * ((C.M + CYG - 1) / CYG);
    const long nu = long((F.N + TX - 1) / TX) * ((F.M + TYG - 1) / TYG);
    return nd <= max_blocks && nu <= max_blocks;
}

template <typename T>
int mg_down0_cg(MGHier<T> &H, T *zA, const CGFuse<T> &f, int max_blocks, hipStream_t st) {
    if (!mg_cg_fusable(H, max_blocks)) return 0;
    const MGLevel<T> &F = H.lev[0];
    const MGLevel<T> &C = H.lev[1];
    const dim3 g((C.N + CX - 1) / CX, (C.M + CYG - 1) / CYG);
    hipLaunchKernelGGL((k_mg_down<T, true>), g, dim3(256), 0, st, F.N, F.M, f.rin, zA, F.cx, F.cu, F.dinv,
                       H.omega, C.N, C.M, C.b, f);
    return int(g.x * g.y);
}

template <typename T>
int mg_vcycle(MGHier<T> &H, const T *r, T *&zA, T *&zB, const RedSlot &rs, double *sc, bool init, hipStream_t st,
              bool skip_down0, T *w) {
    int nrz = 0;
    const int Lc = H.coarse_from, L = H.nlev;
    // level-l arrays: level 0 uses the CG's r (b) and z buffers
    auto bptr = [&](int l) -> const T * { return l == 0 ? r : H.lev[l].b; };
    for (int l = skip_down0 ? 1 : 0; l < Lc && l < L - 1; l++) {
        const MGLevel<T> &F = H.lev[l];
        const MGLevel<T> &C = H.lev[l + 1];
        T *x2 = (l == 0) ? zA : H.lev[l].x;
        const dim3 g((C.N + CX - 1) / CX, (C.M + CY - 1) / CY);
        hipLaunchKernelGGL((k_mg_down<T, false>), g, dim3(256), 0, st, F.N, F.M, bptr(l), x2, F.cx, F.cu, F.dinv,
                           H.omega, C.N, C.M, C.b, CGFuse<T>{});
    }
    if (Lc < L) {
        T *xo = (Lc == 0) ? zA : H.lev[Lc].x;
        if (H.cws)
            hipLaunchKernelGGL((k_mg_coarse<T, true>), dim3(1), dim3(1024), 0, st, H.cp, bptr(Lc), xo, H.cws);
        else
            hipLaunchKernelGGL((k_mg_coarse<T, false>), dim3(1), dim3(1024), H.coarse_lds, st, H.cp, bptr(Lc), xo,
                               H.cws);
    }
    const bool dot_on_up = (Lc > 0);
    for (int l = std::min(Lc, L) - 1; l >= 0; l--) {
        MGLevel<T> &F = H.lev[l];
        const T *xc = (l + 1 < L) ? H.lev[l + 1].x : nullptr;
        T *x2 = (l == 0) ? zA : F.x;
        T *xo = (l == 0) ? zB : F.t;
        const bool cgw = (l == 0 && w);
        const dim3 g((F.N + TX - 1) / TX, (F.M + (cgw ? TYG : TY) - 1) / (cgw ? TYG : TY));
        const int Nc = (l + 1 < L) ? H.lev[l + 1].N : 1;
        if (cgw)
            hipLaunchKernelGGL((k_mg_up<T, true>), g, dim3(256), 0, st, F.N, F.M, bptr(l), x2, xo, F.cx, F.cu,
                               F.dinv, H.omega, Nc, xc, rs, 1, w);
        else
            hipLaunchKernelGGL((k_mg_up<T, false>), g, dim3(256), 0, st, F.N, F.M, bptr(l), x2, xo, F.cx, F.cu,
                               F.dinv, H.omega, Nc, xc, rs, int(l == 0), static_cast<T *>(nullptr));
        if (l == 0) {
            std::swap(zA, zB);
            nrz = int(g.x * g.y);
            if (init && !w) finalize<1>(rs, nrz, sc, RED_RZ_INIT, 0, st);
        } else {
            std::swap(F.x, F.t);
        }
    }
    if (!dot_on_up) {
        const MGLevel<T> &F = H.lev[0];
        const int n = (F.M + 2) * F.N;
        const int nb = std::min(1024, (n + 255) / 256);
        hipLaunchKernelGGL(k_dot<T>, dim3(nb), dim3(256), 0, st, n, r, zA, rs);
        nrz = nb;
        if (init) finalize<1>(rs, nb, sc, RED_RZ_INIT, 0, st);
    }
    return nrz;
}

template int mg_plan<double>(MGHier<double> &, size_t);
template int mg_plan<float>(MGHier<float> &, size_t);
template int mg_vcycle<double>(MGHier<double> &, const double *, double *&, double *&, const RedSlot &, double *,
                               bool, hipStream_t, bool, double *);
template int mg_vcycle<float>(MGHier<float> &, const float *, float *&, float *&, const RedSlot &, double *, bool,
                              hipStream_t, bool, float *);
template bool mg_cg_fusable<double>(const MGHier<double> &, int);
template bool mg_cg_fusable<float>(const MGHier<float> &, int);
template int mg_down0_cg<double>(MGHier<double> &, double *, const CGFuse<double> &, int, hipStream_t);
template int mg_down0_cg<float>(MGHier<float> &, float *, const CGFuse<float> &, int, hipStream_t);

}  // namespace vvmd
