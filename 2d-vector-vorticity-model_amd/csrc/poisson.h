// poisson.h -- the cal_w elliptic solve on the device.
//
// The reference assembles A (src/PoissonSolver.cpp:760-796) for the unknown
// x = rhow[k] w[i][k], i = 1..nx-2 (periodic), k = 2..nz-2, with Dirichlet
// zeros at k = 1 and k = nz-1.  Row (i,k) of A divided by rhow[k] is
//   -L x = cx_k (x_{i-1} - 2 x_i + x_{i+1}) + cu_k (x_{k+1} - x_k) - cu_{k-1} (x_k - x_{k-1})
// with cx_k = 1/rhow[k] and cu_k = 1/rhou[k]: L is symmetric positive
// definite, so the solve is preconditioned CG on L x = f, f = -b / rhow.
// The true residual of the reference system is ||b - A x|| = ||rhow r||.
//
// Unknown arrays are (M+2) rows x N columns, x-contiguous, rows 0 and M+1
// held at zero (the Dirichlet rows); N = nx-2, M = nz-3.
#pragma once

#include "mg.h"
#include "net.h"

namespace vvmd {


// One CG instance in precision T (vectors on level 0 of its hierarchy).
template <typename T>
struct PCG {
    int N = 0, M = 0;
    MGHier<T> mg;
    T *x = nullptr, *r = nullptr, *zA = nullptr, *zB = nullptr, *pA = nullptr, *pB = nullptr, *q = nullptr;
    T *rB = nullptr;       // the other residual buffer of the fused update (swapped with r)
    T *qB = nullptr;       // the other L p buffer of the single-reduction CG (swapped with q)
    T *w = nullptr;        // L z of the last V-cycle (single-reduction CG)
    T *wD = nullptr;       // row weights rhow[k] for ||D r||, [M+2]
    double *sc = nullptr;  // NPSCAL device scalars
};

struct SolveResult {
    int iters = 0;
    int refine = 0;
    double rel = 0., rel0 = 0.;
    bool converged = false;
};

class Poisson {
   public:
    // rhou/rhow are the host profiles (nz long).  Allocates with hipMalloc.
    int init(const Grid &G, const double *rhou, const double *rhow, const RedSlot &rs, bool with_fp32);
    void destroy();

    // cal_w right-hand side: f from zetap, and ||b||^2.
    void rhs_from_zetap(const Grid &G, const double *zetap, const double *rhow_dev, hipStream_t st);
    // b given in the reference's row order (device pointer); fills f and ||b||^2.
    void rhs_from_b(const double *b, hipStream_t st);
    // x (the reference's row order, device) <-> internal solution array.
    void x_to_ref(double *xref, hipStream_t st);
    void x_from_ref(const double *xref, hipStream_t st);
    // w = x / rhow and the westdown halo.
    void back_to_w(const Grid &G, double *w, const double *rhow_dev, hipStream_t st);

    // Guess handling: VVM_GUESS_ZERO clears x, VVM_GUESS_EXTRAP forms
    // 2 x_{t-1} - x_{t-2}, VVM_GUESS_EXTRAP2 3 (x_{t-1} - x_{t-2}) + x_{t-3}
    // (the linear form until three solutions exist); the others keep the
    // current contents of x.
    void set_guess(int mode, hipStream_t st);
    // Record x as the latest solution (for the extrapolated guess).
    void remember_solution(hipStream_t st);
    double *xprev_ptr() const { return xprev; }
    double *xprev2_ptr() const { return xprev2; }
    double *xprev3_ptr() const { return xprev3; }
    int nsolves = 0;
    // order of the extrapolation the learned guess corrects and the
    // recorded samples use: quadratic once three solutions exist
    int xe_order() const { return nsolves >= 3 ? 2 : 1; }
    // Plain FP64 MG-PCG.  tol: relative true residual ||b-Ax||/||b||.
    SolveResult solve_fp64(double tol, int max_iter, hipStream_t st);
    // FP32 MG-PCG inside FP64 iterative refinement.
    SolveResult solve_mixed(double tol, int max_outer, int max_inner, double inner_tol, hipStream_t st);
    // The reference GPU back-end's solver (k_bicgstab.hip): BiCGSTAB with a
    // red-black ILU0 preconditioner; stops at ||b - A x||_2 <= tol
    // (absolute) or <= tol ||b||_2.  Needs an even column count.
    SolveResult solve_bicgstab(double tol, bool absolute, int max_iter, hipStream_t st);

    // Re-plan which multigrid levels the one-workgroup coarse kernel takes
    // (LDS budget in bytes, at most 64 KiB); a tuning knob, the
    // preconditioner's arithmetic is the same either way.
    int replan(size_t lds_budget);
    size_t lds_budget = size_t(64) * 1024;

    double *x64() const { return pcg64.x; }
    double *f64() const { return f; }
    int N = 0, M = 0;

    // Learned initial guess (k_guess.hip).  With a network attached and two
    // previous solutions, x = xe + s_d * net(features), xe the extrapolation
    // of order xe_order(); returns false (nothing written) otherwise.
    GuessNet *net = nullptr;
    bool learned_guess(hipStream_t st);
    // The network's guess was written for the next solve.  That solve keeps
    // it only if its initial residual is finite and below the extrapolated
    // guess's (which the feature pass measured); otherwise it restarts from
    // the extrapolation, so a poor or broken network never costs iterations
    // or accuracy.  guess_fallbacks counts the restarts.
    bool guess_from_net = false;
    int guess_fallbacks = 0;
    // A network whose guesses keep losing stops costing its inference: after
    // a restart it sits out net_skip solves, 1, 4, 16, then 64 while its
    // guesses keep losing; a kept guess resets the count.
    static constexpr int kMaxNetSkip = 64;
    int net_skip = 0, net_backoff = 0;
    int guess_skipped = 0;
    // Sample recording for training: features before the solve, target
    // (x - xe) / s_d after it.  cap 0 frees the buffer.
    int collect_reserve(int cap);
    void collect_features(hipStream_t st);  // before the solve (when recording)
    void collect_target(hipStream_t st);    // after the solve
    float *samples = nullptr;               // [cap][4][M][N]: 3 features, target
    int sample_cap = 0, sample_count = 0;
    bool sample_open = false;               // features written, target pending
    int sample_ord = 1;                     // extrapolation order of the open sample

   private:
    void vcycle64(hipStream_t st, bool first_done, double **zcur);
    SolveResult pcg_run64(double tol, int max_iter, hipStream_t st);
    PCG<double> pcg64;
    PCG<float> pcg32;
    bool have32 = false;
    double *f = nullptr;       // rhs of L x = f, (M+2)*N
    double *xprev = nullptr;   // previous solution (for extrapolation)
    double *xprev2 = nullptr;
    double *xprev3 = nullptr;
    double *r64 = nullptr;     // FP64 residual of the refinement loop
    // pinned host copy of scalars: [P_RNORM2], [P_BNORM2] of the first
    // residual, [8..] the CG loop's, [H_GUESS..H_GUESS+3] the feature pass's
    // sums (H_GUESS + 3: ||wD r||^2 of the extrapolated guess)
    static constexpr int H_GUESS = 12;
    double *hscal = nullptr;
    double *hscal_dev = nullptr;  // the same memory as the device sees it
    double *hpart = nullptr, *hpart_dev = nullptr;  // per-workgroup norms (mapped host memory)
    RedSlot red;
    float *gfeat = nullptr;    // [3][M][N] features of the learned guess
    double *gsc = nullptr;     // sums of squares of the three features
    NetScratch gscratch;       // this solver's network activations
    void guess_features(float *dst, hipStream_t st);
    hipEvent_t ev = nullptr;
    hipEvent_t ev_res = nullptr;  // after the first residual: the host waits on it, not on the whole stream
    // Work folded into the next solve's first residual kernel (one launch
    // instead of two each): the ||b||^2 partials the right-hand-side kernel
    // left in row 1 of red (their block count, 0 for none), the guess
    // (GuessMode below) and, between refinement sweeps, the FP32
    // correction.  flush_pending runs the first two on their own, for
    // callers that use red or x before the solve.
    int bn_pending = 0;
    int x_pending = 0;
    int x_pending_ord = 1;   // extrapolation order of a pending XF_EXTRAP guess
    bool corr_pending = false;
    double *x64B = nullptr;  // the other FP64 solution buffer (the correction writes it)
    double *bcg = nullptr, *bsc = nullptr;  // BiCGSTAB vectors and scalars (allocated on first use)
    float *x32B = nullptr;   // the other FP32 solution buffer (zeroed while the correction reads x)
    void flush_pending(hipStream_t st);
    // r = f - L x (x with the pending guess or correction applied, and
    // written), ||w r||^2 (and ||b||^2 if pending) into sc and hscal; with
    // fp32_rhs also the FP32 copy of r and a zero FP32 x.
    void first_residual(double *r, bool fp32_rhs, hipStream_t st);
    // The first residual of a solve, with its first V-cycle queued before the
    // host reads the norms (pre_vcycle: that V-cycle's partial count into
    // *n_rz_pre), and the learned guess judged against the extrapolation.
    // Returns ||b||^2 and ||w r||^2 as the host read them.
    template <typename T>
    void start_solve(double *r, bool fp32_rhs, PCG<T> *pre_vcycle, int *n_rz_pre, double *bn2, double *rn2,
                     hipStream_t st);
};

// cal_u on the device (src/PoissonSolver.cpp:580-653): top-row potential
// flow and the downward integration of u.
bool cal_u_supported(const Grid &G);  // nx - 2 <= 8192
bool bicgstab_supported(int N);      // red-black colouring: N even
void launch_cal_u(const Grid &G, const StepPtrs &S, hipStream_t st);

}  // namespace vvmd
