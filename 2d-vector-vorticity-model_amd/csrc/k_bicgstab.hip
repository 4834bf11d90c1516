// k_bicgstab.hip -- the cal_w solve as the reference's GPU back-end runs
// it: AMGX BiCGSTAB with an ILU0 preconditioner, tolerance 1e-12 on the L2
// norm of b - A x (src/PoissonSolver_GPU_mpi.cpp:41, AMGX's default
// ABSOLUTE criterion), at most 10000 iterations.  A comparison mode beside
// the multigrid solvers, not a fast path.
//
// ILU0 on the GPU is the multicolour kind (AMGX's ILU0 colours the matrix
// too): with the red-black colouring (i + j even / odd) of the 5-point
// operator there are no red-red or black-black couplings, so in red-black
// order A = [D_R C; C^T D_B], and ILU0 keeps L = [I 0; C^T D_R^-1 I],
// U = [D_R C; 0 S] with S the diagonal of D_B - C^T D_R^-1 C.  Applying
// (LU)^-1 is one pass over the black points and one over the red points.
// The periodic columns keep the colouring only for an even column count,
// which this mode therefore requires.
//
// Right-preconditioned BiCGSTAB (van der Vorst 1992, Algorithm 3.1 with
// the preconditioner on the right): per iteration two operator products,
// two preconditioner applications and three grid-wide reductions; the
// scalars are formed on the device by the kernels that use them, and the
// host reads ||b - A x||^2 once per iteration (mapped memory, after an
// event) for the stop rule.
#pragma clang fp contract(off)

#include <cmath>

#include "poisson.h"

namespace vvmd {

namespace {

constexpr int BX = 64, BY = 4;
inline dim3 bgrid(int N, int M) { return dim3((N + BX - 1) / BX, (M + BY - 1) / BY); }

// scalars of the iteration (bsc): rho of even / odd iterations, r^.v,
// t.s, t.t, alpha and omega of the last iteration
enum { B_RHO0 = 0, B_RHO1, B_RV, B_TS, B_TT, B_ALPHA, B_OMEGA, B_RN, NBSC = 8 };

// The operator of the unknown grid (rows 1..M, periodic columns, zero rows
// 0 and M+1): diagonal cx[j] * 2 + cu[j] + cu[j-1], off-diagonals -cx[j]
// (west, east), -cu[j] (north), -cu[j-1] (south).
__device__ inline double diag(const double *cx, const double *cu, int j) { return 2. * cx[j] + cu[j] + cu[j - 1]; }

__device__ inline double applyA(const double *x, int N, int i, int j, double cxj, double cuj, double cujm) {
    const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
    const size_t r = size_t(j) * N;
    const double c = x[r + i];
    return cxj * (2. * c - x[r + im] - x[r + ip]) + cuj * (c - x[r + N + i]) + cujm * (c - x[r - N + i]);
}

// z = M^-1 r, black pass: z_b = (r_b - sum_q a_bq r_q / d_q) / S_b over the
// red neighbours q (rows outside 1..M are not unknowns and add nothing)
__global__ void __launch_bounds__(BX *BY) k_ilu_black(int N, int M, const double *cx, const double *cu,
                                                      const double *r, double *z) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    if (i >= N || j > M || ((i + j) & 1) == 0) return;
    const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
    const size_t row = size_t(j) * N;
    const double dj = diag(cx, cu, j);
    // S_b = d_j - sum_q a_bq a_qb / d_q (the operator is symmetric)
    double S = dj - 2. * cx[j] * cx[j] / dj;
    double acc = -cx[j] * (r[row + im] + r[row + ip]) / dj;
    if (j < M) {
        const double dn = diag(cx, cu, j + 1);
        S -= cu[j] * cu[j] / dn;
        acc += -cu[j] * r[row + N + i] / dn;
    }
    if (j > 1) {
        const double ds = diag(cx, cu, j - 1);
        S -= cu[j - 1] * cu[j - 1] / ds;
        acc += -cu[j - 1] * r[row - N + i] / ds;
    }
    z[row + i] = (r[row + i] - acc) / S;
}

// red pass: z_r = (r_r - sum_q a_rq z_q) / d_r over the black neighbours
__global__ void __launch_bounds__(BX *BY) k_ilu_red(int N, int M, const double *cx, const double *cu,
                                                    const double *r, double *z) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    if (i >= N || j > M || ((i + j) & 1) == 1) return;
    const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
    const size_t row = size_t(j) * N;
    double acc = -cx[j] * (z[row + im] + z[row + ip]);
    if (j < M) acc += -cu[j] * z[row + N + i];
    if (j > 1) acc += -cu[j - 1] * z[row - N + i];
    z[row + i] = (r[row + i] - acc) / diag(cx, cu, j);
}

// p = r + beta (p - omega v), beta = (rho / rho_old) (alpha / omega); p = r
// on the first iteration
__global__ void __launch_bounds__(BX *BY) k_bcg_p(int N, int M, const double *r, double *p, const double *v,
                                                  const double *bsc, int first, int rho_now, int rho_old) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    if (i >= N || j > M) return;
    const size_t c = size_t(j) * N + i;
    if (first) {
        p[c] = r[c];
        return;
    }
    const double omega = bsc[B_OMEGA];
    const double beta = (bsc[rho_now] / bsc[rho_old]) * (bsc[B_ALPHA] / omega);
    p[c] = r[c] + beta * (p[c] - omega * v[c]);
}

// y = A x and the partials of a.y (and y.y with two)
template <int NV>
__global__ void __launch_bounds__(BX *BY) k_bcg_A(int N, int M, const double *cx, const double *cu,
                                                  const double *x, double *y, const double *a, RedSlot rs) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    double part[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) part[q] = 0.;
    if (i < N && j <= M) {
        const size_t c = size_t(j) * N + i;
        const double yv = applyA(x, N, i, j, cx[j], cu[j], cu[j - 1]);
        y[c] = yv;
        part[0] = a[c] * yv;
        if (NV == 2) part[NV - 1] = yv * yv;
    }
    block_partials<NV>(part, rs);
}

// alpha = rho / r^.v; s = r - alpha v (workgroup 0 keeps alpha)
__global__ void __launch_bounds__(BX *BY) k_bcg_s(int N, int M, const double *r, const double *v, double *s,
                                                  double *bsc, int rho_now) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    const double alpha = bsc[rho_now] / bsc[B_RV];
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) bsc[B_ALPHA] = alpha;
    if (i >= N || j > M) return;
    const size_t c = size_t(j) * N + i;
    s[c] = r[c] - alpha * v[c];
}

// omega = t.s / t.t; x += alpha p^ + omega s^; r = s - omega t; partials
// of r^.r (the next rho) and ||wD r||^2 (workgroup 0 keeps omega)
__global__ void __launch_bounds__(BX *BY) k_bcg_update(int N, int M, double *x, double *r, const double *rh,
                                                       const double *ph, const double *sh, const double *s,
                                                       const double *t, const double *wD, double *bsc,
                                                       RedSlot rs) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    const double alpha = bsc[B_ALPHA];
    const double omega = bsc[B_TT] != 0. ? bsc[B_TS] / bsc[B_TT] : 0.;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) bsc[B_OMEGA] = omega;
    double part[2] = {0., 0.};
    if (i < N && j <= M) {
        const size_t c = size_t(j) * N + i;
        x[c] = x[c] + alpha * ph[c] + omega * sh[c];
        const double rn = s[c] - omega * t[c];
        r[c] = rn;
        part[0] = rh[c] * rn;
        const double wr = wD[j] * rn;
        part[1] = wr * wr;
    }
    block_partials<2>(part, rs);
}

// partials of a.b over the unknowns
__global__ void __launch_bounds__(BX *BY) k_bcg_dot(int N, int M, const double *a, const double *b, RedSlot rs) {
    const int i = blockIdx.x * BX + threadIdx.x, j = 1 + blockIdx.y * BY + threadIdx.y;
    double part[1] = {0.};
    if (i < N && j <= M) {
        const size_t c = size_t(j) * N + i;
        part[0] = a[c] * b[c];
    }
    block_partials<1>(part, rs);
}

__global__ void k_bcg_copy(size_t n, const double *a, double *b) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

}  // namespace

bool bicgstab_supported(int N) { return N % 2 == 0; }

SolveResult Poisson::solve_bicgstab(double tol, bool absolute, int max_iter, hipStream_t st) {
    SolveResult R;
    guess_from_net = false;  // BiCGSTAB starts from whatever guess x holds
    const size_t n = size_t(M + 2) * N;
    if (!bcg) {
        // r^, p, v, s, t, p^, s^ (zeroed, so their Dirichlet rows are zero)
        if (hipMalloc(reinterpret_cast<void **>(&bcg), 7 * n * sizeof(double)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&bsc), NBSC * sizeof(double)) != hipSuccess) {
            R.converged = false;
            return R;
        }
        hipMemsetAsync(bcg, 0, 7 * n * sizeof(double), st);
        hipMemsetAsync(bsc, 0, NBSC * sizeof(double), st);
    }
    double *rh = bcg, *p = bcg + n, *v = bcg + 2 * n, *s = bcg + 3 * n, *t = bcg + 4 * n, *ph = bcg + 5 * n,
           *sh = bcg + 6 * n;
    PCG<double> &P = pcg64;
    const double *cx = P.mg.lev[0].cx, *cu = P.mg.lev[0].cu;
    const dim3 g = bgrid(N, M), b(BX, BY);
    const int nb = nblocks(g);
    // r = f - A x with the pending guess, ||wD r||^2 and ||b||^2 into hscal;
    // r^ = r, rho_0 = r^.r
    first_residual(P.r, false, st);
    hipLaunchKernelGGL(k_bcg_copy, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, n, P.r, rh);
    hipLaunchKernelGGL(k_bcg_dot, g, b, 0, st, N, M, P.r, P.r, red);
    finalize<1>(red, nb, bsc, RED_STORE, B_RHO0, st);
    hipStreamSynchronize(st);
    const double bn2 = hscal[P_BNORM2], rn2 = hscal[P_RNORM2];
    R.rel0 = bn2 > 0. ? std::sqrt(rn2 / bn2) : 0.;
    auto done = [&](double r2) { return absolute ? r2 <= tol * tol : r2 <= tol * tol * bn2; };
    R.rel = R.rel0;
    if (bn2 == 0. || done(rn2)) {
        R.converged = true;
        return R;
    }
    double last = rn2;
    int it = 0;
    while (it < max_iter) {
        const int rho_now = (it & 1) ? B_RHO1 : B_RHO0, rho_next = (it & 1) ? B_RHO0 : B_RHO1;
        hipLaunchKernelGGL(k_bcg_p, g, b, 0, st, N, M, P.r, p, v, bsc, int(it == 0), rho_now, rho_next);
        hipLaunchKernelGGL(k_ilu_black, g, b, 0, st, N, M, cx, cu, p, ph);
        hipLaunchKernelGGL(k_ilu_red, g, b, 0, st, N, M, cx, cu, p, ph);
        hipLaunchKernelGGL((k_bcg_A<1>), g, b, 0, st, N, M, cx, cu, ph, v, rh, red);
        finalize<1>(red, nb, bsc, RED_STORE, B_RV, st);
        hipLaunchKernelGGL(k_bcg_s, g, b, 0, st, N, M, P.r, v, s, bsc, rho_now);
        hipLaunchKernelGGL(k_ilu_black, g, b, 0, st, N, M, cx, cu, s, sh);
        hipLaunchKernelGGL(k_ilu_red, g, b, 0, st, N, M, cx, cu, s, sh);
        hipLaunchKernelGGL((k_bcg_A<2>), g, b, 0, st, N, M, cx, cu, sh, t, s, red);
        finalize<2>(red, nb, bsc, RED_STORE, B_TS, st);
        hipLaunchKernelGGL(k_bcg_update, g, b, 0, st, N, M, P.x, P.r, rh, ph, sh, s, t, P.wD, bsc, red);
        // rho of the next iteration, and ||wD r||^2 to the host
        finalize<1>(red, nb, bsc, RED_STORE, rho_next, st);
        const RedSlot row1{red.partial + red.max_blocks, red.max_blocks};
        finalize<1>(row1, nb, bsc, RED_STORE, B_RN, st, -1, hscal_dev + P_RNORM2);
        hipEventRecord(ev, st);
        hipEventSynchronize(ev);
        it++;
        last = *static_cast<volatile double *>(hscal + P_RNORM2);
        if (!(last == last)) break;  // breakdown (NaN): stop, not converged
        if (done(last)) {
            R.converged = true;
            break;
        }
    }
    R.iters = it;
    R.rel = std::sqrt(last / bn2);
    return R;
}

}  // namespace vvmd
