// mg.h -- multigrid hierarchy of the elliptic solve (see k_mg.hip).
#pragma once

#include "vvm_internal.h"

namespace vvmd {

template <typename T>
struct MGLevel {
    int N = 0, M = 0;
    T *cx = nullptr, *cu = nullptr, *dinv = nullptr;  // [M+2], [M+1], [M+2]
    T *x = nullptr, *b = nullptr, *t = nullptr;       // (M+2)*N each (levels >= 1)
};

constexpr int kMaxLevels = 16;
constexpr int kMaxCoarse = 10;  // levels handled by the one-workgroup kernel

// Levels coarse_from..nlev-1 run in one workgroup with LDS offsets (in T).
template <typename T>
struct CoarseParams {
    int nlev = 0;
    int N[kMaxCoarse] = {}, M[kMaxCoarse] = {};
    const T *cx[kMaxCoarse] = {}, *cu[kMaxCoarse] = {}, *dinv[kMaxCoarse] = {};
    int offB[kMaxCoarse] = {}, offX[kMaxCoarse] = {};
    int offC[kMaxCoarse] = {};  // cx [M+2], cu [M+1], dinv [M+2] of each level, 3 (M+2) apart
    int offT = 0, Tsize = 0;
    T omega = T(0.8);
    int coarse_iters = 32;
    const T *cinv = nullptr;  // dense inverse of the coarsest level (or null)
    int cinv_n = 0;
    int offI = -1;            // its LDS copy when small (else -1: read from cinv)
};

template <typename T>
struct MGHier {
    int nlev = 0;
    MGLevel<T> lev[kMaxLevels];
    T omega = T(0.8);      // damped-Jacobi weight; two sweeps down, two up
    int coarse_iters = 32;  // Jacobi iterations on the coarsest level
    int coarse_from = 0;    // first level of the one-workgroup kernel
    size_t coarse_lds = 0;  // its LDS bytes (0: the levels live in cws)
    T *cws = nullptr;       // global-memory workspace of the one-workgroup
                            // kernel when even the coarsest level exceeds LDS
    // The coarsest level is solved exactly when it is small: its operator's
    // dense inverse (row-major, unknown t = (j-1) N + i), built once in FP64.
    T *cinv = nullptr;
    int cinv_n = 0;
    CoarseParams<T> cp;
};

// Choose coarse_from for an LDS budget (bytes); fills cp.  When no level
// fits, the coarsest runs from cws (allocated here).  Returns 0, or -1 if
// that allocation fails.
template <typename T>
int mg_plan(MGHier<T> &H, size_t lds_budget);

// One V-cycle on r (level-0 right-hand side) into zA (zA/zB are swapped as
// the result moves between buffers).  With init it also stores r.z in
// sc[P_RZ]; otherwise it leaves the r.z partials in row 0 of rs for the next
// k_pcg_pq and returns their count.  With skip_down0 the level-0 down step
// has already run (mg_down0_cg).  With w (the single-reduction CG), the
// level-0 up step also writes w = L z and leaves the partials of r.z and w.z
// in rows 0 and 1 of rs, and nothing is finalized, init or not.
template <typename T>
int mg_vcycle(MGHier<T> &H, const T *r, T *&zA, T *&zB, const RedSlot &rs, double *sc, bool init, hipStream_t st,
              bool skip_down0 = false, T *w = nullptr);

// Whether the hierarchy has the tiled level 0 that mg_down0_cg and the
// w-writing up step need, within max_blocks workgroups each.
template <typename T>
bool mg_cg_fusable(const MGHier<T> &H, int max_blocks);

// One iteration of the single-reduction (Chronopoulos-Gear) CG fused into
// the level-0 down step.  From gamma = r.z and delta = w.z (the partials the
// last V-cycle's level-0 up step left; z = M r, w = L z):
//   beta = gamma / gamma_old (0 on the first iteration),
//   alpha = gamma / (delta - beta gamma / alpha_old) (gamma / delta first),
//   p = z + beta p, q = w + beta q (= L p), x += alpha p, r -= alpha q,
// then the smoothing and restriction of the new r, and ||wD r||^2 per
// workgroup into hpart (host-mapped memory; the host adds the nblocks
// entries).  In exact arithmetic these are the iterates of standard PCG,
// with one grid-wide reduction an iteration instead of two.  Returns
// nblocks, or 0 when the hierarchy is not fusable.
template <typename T>
struct CGFuse {
    T *x;
    const T *rin;  // r before the update (read with halos by every workgroup)
    T *rout;       // r after it (another buffer: the reads above must see the old r)
    T *p;          // search direction, updated in place on owned points only
    const T *qin;  // L p before the update (read with halos)
    T *qout;       // after it (another buffer, as for r)
    const T *w;    // L z of the last V-cycle
    const T *wD;
    double *sc;
    int g_old, g_new, a_old, a_new;  // slots of gamma and alpha in sc, last and this iteration
    int first;
    RedSlot rs_gd;  // gamma partials in row 0, delta partials in row 1
    int ngd;
    double *hpart;
};
template <typename T>
int mg_down0_cg(MGHier<T> &H, T *zA, const CGFuse<T> &f, int max_blocks, hipStream_t st);

}  // namespace vvmd
