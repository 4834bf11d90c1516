// k_dynamics.hip -- advection, buoyancy, rain sedimentation, the negative
// value fixer, forcing and halos of one VVM timestep (phase A of
// vvm::Iteration::TimeMarching, src/Iteration.cpp:263-285), plus
// PoissonSolver::pubarTop_pt.
//
// FP64 throughout; no FMA contraction, so every expression rounds as the
// reference's does (the oracle and the compiled reference agree bit for bit
// on these operators; the GPU differs only where the device math library's
// pow/sqrt/exp do).
#pragma clang fp contract(off)

#include "vvm_internal.h"

namespace vvmd {

namespace {

constexpr int BX = 64, BY = 4;

__device__ inline double plus_(double v) { return 0.5 * (v + fabs(v)); }
__device__ inline double minu_(double v) { return 0.5 * (v - fabs(v)); }

// Third-order upwind-biased correction of one face flux, Advection.cpp:40-50
// (rho0/v0 at the face, rhom/vm one face upstream, rhop/vp one downstream;
// fm2..fp1 the four cell values around the face).
__device__ inline double upwind3(double rho0, double v0, double rhom, double vm, double rhop, double vp, double fm2,
                                 double fm1, double f0, double fp1) {
    const double p0 = plus_(v0), pm = plus_(vm), m0 = minu_(v0), mp = minu_(vp);
    return -1. / 3. *
           (rho0 * p0 * (f0 - fm1) - sqrt(rho0 * p0 * rhom * pm) * (fm1 - fm2) - rho0 * m0 * (f0 - fm1) -
            sqrt(fabs(rho0 * m0 * rhop * mp)) * (fp1 - f0));
}

__device__ inline bool corr_window(const Grid &G, int i, int k) {
    return i >= 2 && i <= G.nx - 3 && k >= 2 && k <= G.nz - 3;
}

// Scalar x-face flux at (i,k) (between i-1 and i), Advection.cpp:36,40-44.
__device__ inline double sflux_x(const Grid &G, const double *__restrict__ f, const double *__restrict__ u,
                                 const double *__restrict__ rhou, int i, int k) {
    const double r = rhou[k], ui = u[ix(G, i, k)];
    const double f0 = f[ix(G, i, k)], fm1 = f[ix(G, i - 1, k)];
    double fl = r * ui * (f0 + fm1);
    if (corr_window(G, i, k))
        fl += upwind3(r, ui, r, u[ix(G, i - 1, k)], r, u[ix(G, i + 1, k)], f[ix(G, i - 2, k)], fm1, f0,
                      f[ix(G, i + 1, k)]);
    return fl;
}

// Scalar z-face flux at (i,k) (between k-1 and k), Advection.cpp:37,46-50.
__device__ inline double sflux_z(const Grid &G, const double *__restrict__ f, const double *__restrict__ w,
                                 const double *__restrict__ rhow, int i, int k) {
    const double r = rhow[k], wk = w[ix(G, i, k)];
    const double f0 = f[ix(G, i, k)], fm1 = f[ix(G, i, k - 1)];
    double fl = r * wk * (f0 + fm1);
    if (corr_window(G, i, k))
        fl += upwind3(r, wk, rhow[k - 1], w[ix(G, i, k - 1)], rhow[k + 1], w[ix(G, i, k + 1)], f[ix(G, i, k - 2)],
                      fm1, f0, f[ix(G, i, k + 1)]);
    return fl;
}

// Corner velocities of Advection_zeta, Advection.cpp:90-91.
__device__ inline double Uw(const Grid &G, const double *u, const double *rhou, int i, int k) {
    return 0.25 * (rhou[k] * (u[ix(G, i + 1, k)] + u[ix(G, i, k)]) +
                   rhou[k - 1] * (u[ix(G, i + 1, k - 1)] + u[ix(G, i, k - 1)]));
}
__device__ inline double Wu(const Grid &G, const double *w, const double *rhow, int i, int k) {
    return 0.25 * (rhow[k + 1] * (w[ix(G, i, k + 1)] + w[ix(G, i - 1, k + 1)]) +
                   rhow[k] * (w[ix(G, i, k)] + w[ix(G, i - 1, k)]));
}

// Vorticity fluxes, Advection.cpp:107-118 (the density already sits in U_w/W_u).
__device__ inline double zflux_x(const Grid &G, const double *z, const double *u, const double *rhou, int i, int k) {
    const double U = Uw(G, u, rhou, i, k);
    const double z0 = z[ix(G, i, k)], z1 = z[ix(G, i + 1, k)];
    double fl = U * (z1 + z0);
    if (corr_window(G, i, k)) {
        const double Pu = plus_(U), Pum = plus_(Uw(G, u, rhou, i - 1, k));
        const double Mu = minu_(U), Mup = minu_(Uw(G, u, rhou, i + 1, k));
        fl += -1. / 3. *
              (Pu * (z1 - z0) - sqrt(Pu * Pum) * (z0 - z[ix(G, i - 1, k)]) - Mu * (z1 - z0) -
               sqrt(fabs(Mu * Mup)) * (z[ix(G, i + 2, k)] - z1));
    }
    return fl;
}
__device__ inline double zflux_z(const Grid &G, const double *z, const double *w, const double *rhow, int i, int k) {
    const double W = Wu(G, w, rhow, i, k);
    const double z0 = z[ix(G, i, k)], z1 = z[ix(G, i, k + 1)];
    double fl = W * (z1 + z0);
    if (corr_window(G, i, k)) {
        const double Pw = plus_(W), Pwm = plus_(Wu(G, w, rhow, i, k - 1));
        const double Mw = minu_(W), Mwp = minu_(Wu(G, w, rhow, i, k + 1));
        fl += -1. / 3. *
              (Pw * (z1 - z0) - sqrt(Pw * Pwm) * (z0 - z[ix(G, i, k - 1)]) - Mw * (z1 - z0) -
               sqrt(fabs(Mw * Mwp)) * (z[ix(G, i, k + 2)] - z1));
    }
    return fl;
}

__device__ inline double thv(const Grid &G, const double *th, const double *qv, int i, int k) {
    return th[ix(G, i, k)] * (1 + 0.608 * qv[ix(G, i, k)]);
}

// AB2 update of one history slot: writes d to slot sn (and to slot 0 on the
// reference's step-0 start, Advection.cpp:70) and returns the old slot.
__device__ inline double ab2_hist(double *const (&h)[2], size_t p, int step, double d) {
    // the slot pointers are picked by selects, not by a run-time index into
    // the kernel-argument array (which would read them through generic
    // memory accesses)
    const bool odd = (step & 1) != 0;
    double *const hn = odd ? h[0] : h[1], *const ho = odd ? h[1] : h[0];
    hn[p] = d;
    if (step == 0) h[0][p] = d;
    return ho[p];
}

// ------------------------------------------------------------------ zeta

// Iteration::pzeta_pt: Advection_zeta (Advection.cpp:84-149) + Bouyancy
// (Buoyancy.cpp:11-58) + BoundaryProcess2D_westdown(zetap).  As in
// k_thermo4, each vorticity flux is formed once per face: every thread forms
// the flux on its own east and upper faces, the first two waves the faces on
// the tile's west and lower edges, and after a barrier each point takes the
// differences.
constexpr int PZX = 64, PZY = 8;

// PARTS: PZ_ADV Advection_zeta, PZ_BUOY Bouyancy; both (the timestep) also
// apply BoundaryProcess2D_westdown.  Alone each writes what the reference
// function writes: the interior points k = 2..nz-2 (zetap = ... for the
// advection, zetap += ... for the buoyancy) and its AB2 history.
enum { PZ_ADV = 1, PZ_BUOY = 2, PZ_BOTH = 3 };

template <int PARTS>
__global__ void __launch_bounds__(PZX *PZY) k_pzeta(Grid G, StepPtrs S, int step) {
    __shared__ double Fx[PZY][PZX + 1];  // slot c: x-flux east of column i0 + c - 1
    __shared__ double Fz[PZY + 1][PZX];  // slot r: z-flux above row k0 + r - 1
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = blockIdx.x * PZX, k0 = blockIdx.y * PZY;
    const int i = i0 + tx, k = k0 + ty;
    const int nx = G.nx, nz = G.nz;
    const double *z = S.now[F_ZETA], *u = S.u, *w = S.w, *rhou = S.P.rhou, *rhow = S.P.rhow;
    if constexpr ((PARTS & PZ_ADV) != 0) {
        // flux_u at i-1 = 0 is the periodic image of nx-2 (BoundaryProcess2D_center
        // on flux_u): the ghost column 0 holds it for column 1
        if (k >= 2 && k <= nz - 2 && i <= nx - 2) Fx[ty][tx + 1] = zflux_x(G, z, u, rhou, (i == 0) ? nx - 2 : i, k);
        if (k >= 1 && k <= nz - 2 && i >= 1 && i <= nx - 2) Fz[ty + 1][tx] = zflux_z(G, z, w, rhow, i, k);
        if (ty == 0 && tx < PZY) {  // west edge: the face east of column i0 - 1
            const int kr = k0 + tx;
            if (i0 >= 1 && i0 <= nx - 2 && kr >= 2 && kr <= nz - 2) Fx[tx][0] = zflux_x(G, z, u, rhou, i0 - 1, kr);
        } else if (ty == 1) {  // lower edge: the face above row k0 - 1
            if (k0 >= 2 && k0 <= nz - 2 && i >= 1 && i <= nx - 2) Fz[0][tx] = zflux_z(G, z, w, rhow, i, k0 - 1);
        }
        __syncthreads();
    }
    if (i >= nx || k >= nz) return;
    double *zp = S.fut[F_ZETA];
    if (k == 0 || k == 1 || k == nz - 1) {
        if (PARTS == PZ_BOTH) zp[ix(G, i, k)] = 0.;
        return;
    }
    if (i < 1 || i > nx - 2) return;
    const size_t p = ix(G, i, k);
    double zeta_new;
    if constexpr ((PARTS & PZ_ADV) != 0) {
        const double fu0 = Fx[ty][tx + 1], fum = Fx[ty][tx];
        const double fw0 = Fz[ty + 1][tx], fwm = Fz[ty][tx];
        const double px = (fu0 - fum) * G.r2dx / rhow[k];
        const double pz = (fw0 - fwm) * G.r2dz / rhow[k];
        const double dadv = -px - pz;
        const double dold = ab2_hist(S.hist[H_ZETA_ADV], p, step, dadv);
        zeta_new = z[p] + 1.5 * G.dt * dadv - 0.5 * G.dt * dold;
        if constexpr (PARTS == PZ_ADV) {
            zp[p] = zeta_new;
            return;
        }
    } else {
        zeta_new = zp[p];
    }

    const double *th = S.now[F_TH], *qv = S.now[F_QV], *qc = S.now[F_QC], *qr = S.now[F_QR], *thvb = S.P.thvb;
    const double gth = G.GRAVITY / rhow[k] * 0.5 *
                       ((thv(G, th, qv, i, k) - thv(G, th, qv, i - 1, k)) / thvb[k] +
                        (thv(G, th, qv, i, k - 1) - thv(G, th, qv, i - 1, k - 1)) / thvb[k - 1]) *
                       G.rdx;
    const double gqc = G.GRAVITY / rhow[k] *
                       (0.5 * (qc[ix(G, i, k)] + qc[ix(G, i, k - 1)]) - 0.5 * (qc[ix(G, i - 1, k)] + qc[ix(G, i - 1, k - 1)])) *
                       G.rdx;
    const double gqr = G.GRAVITY / rhow[k] *
                       (0.5 * (qr[ix(G, i, k)] + qr[ix(G, i, k - 1)]) - 0.5 * (qr[ix(G, i - 1, k)] + qr[ix(G, i - 1, k - 1)])) *
                       G.rdx;
    double db = gth;
    db += -gqc - gqr;
    const double bold = ab2_hist(S.hist[H_TH_BUOY], p, step, db);
    zeta_new += (1.5 * G.dt * db) - (0.5 * G.dt * bold);
    if constexpr (PARTS == PZ_BOTH) store_westdown(zp, G, i, k, zeta_new);
    else zp[p] = zeta_new;
}

// ----------------------------------------------------------- scalars

// Advection_thermo (Advection.cpp:25-81) for one field at interior (i,k).
__device__ inline double thermo_tend(const Grid &G, const double *f, const double *u, const double *w,
                                     const double *rhou, const double *rhow, int i, int k) {
    const double fu1 = sflux_x(G, f, u, rhou, (i + 1 == G.nx - 1) ? 1 : i + 1, k);
    const double fu0 = sflux_x(G, f, u, rhou, i, k);
    const double fw1 = (k + 1 == G.nz - 1) ? 0. : sflux_z(G, f, w, rhow, i, k + 1);
    const double fw0 = sflux_z(G, f, w, rhow, i, k);
    const double px = (fu1 - fu0) * G.r2dx / rhou[k];
    const double pz = (fw1 - fw0) * G.r2dz / rhou[k];
    return -px - pz;
}

// The parts of a scalar face flux that depend only on the face (density and
// velocity), formed once and applied to every field advected across it.
// flux(f) rounds exactly as sflux_x / sflux_z do: the same products in the
// same order (upwind3's rho0 p0 (f0 - fm1) is (rho0 p0) (f0 - fm1), and so on).
struct FaceFlux {
    double rv;          // r v of the centred term
    double a, b, c, d;  // rho0 p0, sqrt(rho0 p0 rhom pm), rho0 m0, sqrt(|rho0 m0 rhop mp|)
    bool corr;          // inside the third-order correction window
    size_t s;           // index step across the face (1 for x faces, the row pitch for z faces)
    __device__ double flux(const double *__restrict__ f, size_t p) const {
        const double f0 = f[p], fm1 = f[p - s];
        double fl = rv * (f0 + fm1);
        if (corr) {
            const double fm2 = f[p - 2 * s], fp1 = f[p + s];
            fl += -1. / 3. * (a * (f0 - fm1) - b * (fm1 - fm2) - c * (f0 - fm1) - d * (fp1 - f0));
        }
        return fl;
    }
};
__device__ inline FaceFlux face_coeffs(double rho0, double v0, double rhom, double vm, double rhop, double vp,
                                       bool corr, size_t s) {
    FaceFlux F;
    F.rv = rho0 * v0;
    F.corr = corr;
    F.s = s;
    F.a = F.b = F.c = F.d = 0.;
    if (corr) {
        const double p0 = plus_(v0), pm = plus_(vm), m0 = minu_(v0), mp = minu_(vp);
        F.a = rho0 * p0;
        F.b = sqrt(F.a * rhom * pm);
        F.c = rho0 * m0;
        F.d = sqrt(fabs(F.c * rhop * mp));
    }
    return F;
}
// x face between i-1 and i / z face between k-1 and k, as sflux_x / sflux_z
__device__ inline FaceFlux face_x(const Grid &G, const double *u, const double *rhou, int i, int k) {
    const bool corr = corr_window(G, i, k);
    const double r = rhou[k];
    return face_coeffs(r, u[ix(G, i, k)], r, corr ? u[ix(G, i - 1, k)] : 0., r, corr ? u[ix(G, i + 1, k)] : 0., corr,
                       1);
}
__device__ inline FaceFlux face_z(const Grid &G, const double *w, const double *rhow, int i, int k) {
    const bool corr = corr_window(G, i, k);
    return face_coeffs(rhow[k], w[ix(G, i, k)], corr ? rhow[k - 1] : 0., corr ? w[ix(G, i, k - 1)] : 0.,
                       corr ? rhow[k + 1] : 0., corr ? w[ix(G, i, k + 1)] : 0., corr, G.pitch);
}

// Rain sedimentation flux of Advection_qrVT (Advection.cpp:165-166) at (i,k).
__device__ inline double vt_flux(const Grid &G, const double *qr, const double *rhow, const double *vtfac, int i,
                                 int k) {
    const double s = qr[ix(G, i, k)] + qr[ix(G, i, k - 1)];
    const double VT = 1E-2 * (3634 * pow(1E-3 * rhow[k] * 0.5 * s, 0.1346) * vtfac[k]);
    return rhow[k] * VT * s;
}

// pth_pt, pqv_pt, pqc_pt, pqr_pt (src/Iteration.cpp:36-40, 90-124): the four
// scalar advections sharing u and w, rain sedimentation on qr, surface
// precipitation, and the first pass of NegativeValueProcess (zero the
// negatives, sum positives and negatives) for qv, qc, qr.  k_finalize turns
// the sums into rescale ratios.
//
// Each face flux is formed once, as the reference forms its flux arrays:
// a workgroup covers T4X x T4Y points; first every thread forms the x-face
// flux on its west face and the z-face fluxes (four fields and the rain
// fall) on its lower face, and a few waves form the faces on the tile's east
// and top edges; after a barrier each point takes the differences.  (Forming
// both faces of every point would evaluate each flux, with its square roots
// and powers, twice.)
constexpr int T4X = 64, T4Y = 8;

// Four workgroups a CU (the LDS allows four), so that 1024 of them cover
// 2048 x 256 in one round: eight waves a SIMD, at most 64 VGPRs.
__global__ void __launch_bounds__(T4X *T4Y) __attribute__((amdgpu_waves_per_eu(8))) k_thermo4(Grid G, StepPtrs S, int step, RedSlot rs) {
    __shared__ double Fx[4][T4Y][T4X + 1];  // x-face flux at faces i0 .. i0+T4X (west faces)
    __shared__ double Fz[5][T4Y + 1][T4X];  // z-face flux (4 fields, rain fall) at faces k0 .. k0+T4Y
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = blockIdx.x * T4X, k0 = blockIdx.y * T4Y;
    const int i = i0 + tx, k = k0 + ty;
    const int nx = G.nx, nz = G.nz;
    const double *u = S.u, *w = S.w, *rhou = S.P.rhou, *rhow = S.P.rhow;
    const double *fl[4] = {S.now[F_TH], S.now[F_QV], S.now[F_QC], S.now[F_QR]};
    const bool row_in = (k >= 1 && k <= nz - 2), inside = row_in && i >= 1 && i <= nx - 2;
    // the ghost column nx-1 holds the face east of column nx-2: the
    // periodic image of face 1
    // the density-velocity part of each face is formed once for all four
    // fields (sflux_x / sflux_z per field would repeat its square roots)
    if (row_in && i >= 1 && i <= nx - 1) {
        const int ii = (i == nx - 1) ? 1 : i;
        const FaceFlux F = face_x(G, u, rhou, ii, k);
        const size_t p = ix(G, ii, k);
#pragma unroll
        for (int q = 0; q < 4; q++) Fx[q][ty][tx] = F.flux(fl[q], p);
    }
    if (inside) {
        const FaceFlux F = face_z(G, w, rhow, i, k);
        const size_t p = ix(G, i, k);
#pragma unroll
        for (int q = 0; q < 4; q++) Fz[q][ty][tx] = F.flux(fl[q], p);
        Fz[4][ty][tx] = vt_flux(G, fl[3], rhow, S.P.vtfac, i, k);
    }
    // edge faces: waves 0-1 the top faces (four fields each), wave 2 the
    // top rain-fall faces, wave 3 the east faces (one row each)
    {
        const int kt = k0 + T4Y, ie = i0 + T4X;
        if (ty == 0) {
            if (i >= 1 && i <= nx - 2 && kt <= nz - 2) {
                const FaceFlux F = face_z(G, w, rhow, i, kt);
                const size_t p = ix(G, i, kt);
#pragma unroll
                for (int q = 0; q < 4; q++) Fz[q][T4Y][tx] = F.flux(fl[q], p);
            }
        } else if (ty == 2) {
            if (i >= 1 && i <= nx - 2 && kt <= nz - 2) Fz[4][T4Y][tx] = vt_flux(G, fl[3], rhow, S.P.vtfac, i, kt);
        } else if (ty == 3 && tx < T4Y) {
            const int kr = k0 + tx;
            // the face east of column nx-2 is the periodic image of face 1
            if (ie - 1 <= nx - 2 && kr >= 1 && kr <= nz - 2) {
                const int ii = (ie == nx - 1) ? 1 : ie;
                const FaceFlux F = face_x(G, u, rhou, ii, kr);
                const size_t p = ix(G, ii, kr);
#pragma unroll
                for (int q = 0; q < 4; q++) Fx[q][tx][T4X] = F.flux(fl[q], p);
            }
        }
    }
    __syncthreads();
    double part[6] = {0., 0., 0., 0., 0., 0.};
    if (inside) {
        const size_t p = ix(G, i, k);
        const int hid[4] = {H_TH_ADV, H_QV_ADV, H_QC_ADV, H_QR_ADV};
        const bool top = (k + 1 == nz - 1);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // Advection_thermo's tendency (Advection.cpp:55-81)
            const double fu1 = Fx[q][ty][tx + 1], fu0 = Fx[q][ty][tx];
            const double fw1 = top ? 0. : Fz[q][ty + 1][tx];
            const double fw0 = Fz[q][ty][tx];
            const double px = (fu1 - fu0) * G.r2dx / rhou[k];
            const double pz = (fw1 - fw0) * G.r2dz / rhou[k];
            const double d = -px - pz;
            const double dold = ab2_hist(S.hist[hid[q]], p, step, d);
            double v = fl[q][p] + 1.5 * G.dt * d - 0.5 * G.dt * dold;
            if (q == 0) {
                S.fut[F_TH][p] = v;
                continue;
            }
            if (q == 3) {
                const double f1 = top ? 0. : Fz[4][ty + 1][tx];
                const double f0 = Fz[4][ty][tx];
                const double dvt = (f1 - f0) * G.r2dz / rhou[k];
                const double vold = ab2_hist(S.hist[H_QR_VT], p, step, dvt);
                v += 1.5 * G.dt * dvt - 0.5 * G.dt * vold;
            }
            if (v >= 0.) part[2 * (q - 1)] += v;
            else {
                part[2 * (q - 1) + 1] += v;
                v = 0.;
            }
            S.fut[F_QV + q - 1][p] = v;
        }
        if (k == 1) {
            // surface precipitation, src/Iteration.cpp:114-119 (now-qr, last w)
            const double r1 = rhou[1];
            const double q1 = S.now[F_QR][ix(G, i, 1)];
            const double VT = 1E-2 * (3634 * pow(1E-3 * r1 * q1, 0.1346) * S.s[S_PRECIP_FAC]);
            const double rain = -r1 * (0.5 * (w[ix(G, i, 2)] + 0.) - VT) * q1;
            S.precip[i] = (rain < 0) ? 0. : rain;
        }
    }
    block_partials<6>(part, rs);
}

// Second pass of NegativeValueProcess for qv, qc, qr (MicroPhysics.cpp:
// 129-137), AddForcing (AddForcing.cpp:6-28: Q1LS/Q2LS and the random
// theta kick) when `forcing`, and the center halos of thp, qvp, qcp, qrp.
// halos == 0 is AddForcing alone (ratios < 0): interior thp and qvp only.
__global__ void __launch_bounds__(BX *BY) k_scale_forcing(Grid G, StepPtrs S, int forcing, int tropical,
                                                          int random_kick, int halos) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const size_t p = ix(G, i, k);
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int f = F_QV + q;
        double v = S.fut[f][p];
        const double ratio = S.s[S_NEG_RATIO + q];
        if (ratio >= 0. && v > 0) v = v * ratio;
        if (forcing && tropical && f == F_QV) v += G.dt * S.P.Q2LS[k];
        if (halos) store_center(S.fut[f], G, i, k, v);
        else if (f == F_QV) S.fut[f][p] = v;
    }
    double th = S.fut[F_TH][p];
    if (forcing) {
        // AddForcing's order (AddForcing.cpp:15-25): large-scale forcing,
        // radiation heating [K/day], RADIATIONCOOLING, the random kick
        if (tropical) th += G.dt * S.P.Q1LS[k];
        if (S.rad) th += G.dt * S.rad[p] / 86400.;
        if (S.radcool) th += G.dt * (-2. / 86400.);
        if (random_kick) th += G.dt * S.forcing[p];
    }
    if (halos) store_center(S.fut[F_TH], G, i, k, th);
    else S.fut[F_TH][p] = th;
}

// Halo value of BoundaryProcess2D_center at a perimeter cell.
__device__ inline size_t center_src(const Grid &G, int i, int k) {
    const int is = (i == 0) ? G.nx - 2 : (i == G.nx - 1) ? 1 : i;
    const int ks = (k == 0) ? 1 : (k == G.nz - 1) ? G.nz - 2 : k;
    return ix(G, is, ks);
}

// BoundaryProcess2D_center / _westdown on the perimeter of a field.
__device__ inline void bc_center_cell(double *a, const Grid &G, int i, int k) {
    if (i == 0 || i == G.nx - 1 || k == 0 || k == G.nz - 1) a[ix(G, i, k)] = a[center_src(G, i, k)];
}
__device__ inline void bc_westdown_cell(double *a, const Grid &G, int i, int k) {
    if (k == 0 || k == 1 || k == G.nz - 1) a[ix(G, i, k)] = 0.;
    else if (i == 0) a[ix(G, i, k)] = a[ix(G, G.nx - 2, k)];
    else if (i == G.nx - 1) a[ix(G, i, k)] = a[ix(G, 1, k)];
}

__global__ void __launch_bounds__(BX *BY) k_bc_uw(Grid G, StepPtrs S) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i >= G.nx || k >= G.nz) return;
    bc_center_cell(S.u, G, i, k);
    bc_westdown_cell(S.w, G, i, k);
}

// BoundaryProcess2D_all (src/Boundary.cpp:33-52), Kessler build.
__global__ void __launch_bounds__(BX *BY) k_bc_all(Grid G, StepPtrs S) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i >= G.nx || k >= G.nz) return;
    bc_westdown_cell(S.fut[F_ZETA], G, i, k);
    bc_center_cell(S.fut[F_TH], G, i, k);
    bc_westdown_cell(S.w, G, i, k);
    bc_center_cell(S.u, G, i, k);
    bc_center_cell(S.fut[F_QV], G, i, k);
    bc_center_cell(S.fut[F_QC], G, i, k);
    bc_center_cell(S.fut[F_QR], G, i, k);
}

// ------------------------------------------------ single operators alone
// (the statics of class vvm one at a time; the timestep uses the fused
// kernels above).  Each writes exactly what its reference function writes.

// Advection_thermo (Advection.cpp:25-81) of field q (F_TH..F_QR): the
// interior of its p-field and its AB2 history.
__global__ void __launch_bounds__(BX *BY) k_adv_scalar(Grid G, StepPtrs S, int step, int q) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const int hid[NPROG] = {H_ZETA_ADV, H_TH_ADV, H_QV_ADV, H_QC_ADV, H_QR_ADV};
    const double *f = S.now[q];
    const size_t p = ix(G, i, k);
    const double d = thermo_tend(G, f, S.u, S.w, S.P.rhou, S.P.rhow, i, k);
    const double dold = ab2_hist(S.hist[hid[q]], p, step, d);
    S.fut[q][p] = f[p] + 1.5 * G.dt * d - 0.5 * G.dt * dold;
}

// Advection_qrVT (Advection.cpp:153-192): the rain fall added to the
// interior of qrp, and its AB2 history.
__global__ void __launch_bounds__(BX *BY) k_qrvt(Grid G, StepPtrs S, int step) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i < 1 || i > G.nx - 2 || k < 1 || k > G.nz - 2) return;
    const double *qr = S.now[F_QR], *rhow = S.P.rhow;
    const size_t p = ix(G, i, k);
    const double f1 = (k + 1 == G.nz - 1) ? 0. : vt_flux(G, qr, rhow, S.P.vtfac, i, k + 1);
    const double f0 = vt_flux(G, qr, rhow, S.P.vtfac, i, k);
    const double dvt = (f1 - f0) * G.r2dz / S.P.rhou[k];
    const double vold = ab2_hist(S.hist[H_QR_VT], p, step, dvt);
    S.fut[F_QR][p] += 1.5 * G.dt * dvt - 0.5 * G.dt * vold;
}

// Surface precipitation of pqr_pt (src/Iteration.cpp:114-119).
__global__ void k_precip(Grid G, StepPtrs S) {
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i > G.nx - 2) return;
    const double r1 = S.P.rhou[1], q1 = S.now[F_QR][ix(G, i, 1)];
    const double VT = 1E-2 * (3634 * pow(1E-3 * r1 * q1, 0.1346) * S.s[S_PRECIP_FAC]);
    const double rain = -r1 * (0.5 * (S.w[ix(G, i, 2)] + 0.) - VT) * q1;
    S.precip[i] = (rain < 0) ? 0. : rain;
}

// BoundaryProcess2D_center (westdown = 0) or _westdown on one field.
__global__ void __launch_bounds__(BX *BY) k_bc_one(Grid G, double *a, int westdown) {
    const int i = blockIdx.x * BX + threadIdx.x, k = blockIdx.y * BY + threadIdx.y;
    if (i >= G.nx || k >= G.nz) return;
    if (westdown) bc_westdown_cell(a, G, i, k);
    else bc_center_cell(a, G, i, k);
}

// PoissonSolver::pubarTop_pt (src/PoissonSolver.cpp:655-676): domain-mean
// vertical momentum flux and K_M shear at the top, AB2 for ubarTopp.
__global__ void __launch_bounds__(1024) k_pubartop(Grid G, StepPtrs S, int step) {
    const int nx = G.nx, nz = G.nz;
    const double *u = S.u, *w = S.w, *K = S.RKM, *rhou = S.P.rhou, *rhow = S.P.rhow;
    double acc[3] = {0., 0., 0.};
    for (int i = 1 + threadIdx.x; i < nx - 1; i += blockDim.x) {
        acc[0] += 0.25 * (rhou[nz - 2] * u[ix(G, i, nz - 2)] + rhou[nz - 3] * u[ix(G, i, nz - 3)]) *
                  (w[ix(G, i, nz - 2)] + w[ix(G, i - 1, nz - 2)]);
        acc[1] += rhow[nz - 1] * 0.25 *
                  (K[ix(G, i, nz - 1)] + K[ix(G, i, nz - 2)] + K[ix(G, i - 1, nz - 1)] + K[ix(G, i - 1, nz - 2)]) *
                  (G.rdz * (u[ix(G, i, nz - 1)] - u[ix(G, i, nz - 2)]));
        acc[2] += rhow[nz - 2] * 0.25 *
                  (K[ix(G, i, nz - 2)] + K[ix(G, i, nz - 3)] + K[ix(G, i - 1, nz - 2)] + K[ix(G, i - 1, nz - 3)]) *
                  (G.rdx * (w[ix(G, i, nz - 2)] - w[ix(G, i - 1, nz - 2)]) +
                   G.rdz * (u[ix(G, i, nz - 2)] - u[ix(G, i, nz - 3)]));
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        const double n = (double)(nx - 2.);
        const double down = acc[0] / n, kup = acc[1] / n, kdown = acc[2] / n;
        const double p1 = (-down) * G.rdz / rhou[nz - 2];
        const double p2 = (kup - kdown) * G.rdz / rhou[nz - 2];
        const int sn = (step + 1) % 2, so = step % 2;
        double *dub = S.s + S_DUBAR0;
        dub[sn] = -p1 + p2;
        if (step == 0) dub[0] = dub[1];
        S.s[S_UBARTOPP] = S.s[S_UBARTOP] + 1.5 * G.dt * dub[sn] - 0.5 * G.dt * dub[so];
    }
}

inline dim3 grid2d(const Grid &G) { return dim3((G.nx + BX - 1) / BX, (G.nz + BY - 1) / BY); }

}  // namespace

void launch_pzeta(const Grid &G, const StepPtrs &S, int step, hipStream_t st) {
    const dim3 g((G.nx + PZX - 1) / PZX, (G.nz + PZY - 1) / PZY);
    hipLaunchKernelGGL(k_pzeta<PZ_BOTH>, g, dim3(PZX, PZY), 0, st, G, S, step);
}
void launch_advection_zeta(const Grid &G, const StepPtrs &S, int step, hipStream_t st) {
    const dim3 g((G.nx + PZX - 1) / PZX, (G.nz + PZY - 1) / PZY);
    hipLaunchKernelGGL(k_pzeta<PZ_ADV>, g, dim3(PZX, PZY), 0, st, G, S, step);
}
void launch_buoyancy(const Grid &G, const StepPtrs &S, int step, hipStream_t st) {
    const dim3 g((G.nx + PZX - 1) / PZX, (G.nz + PZY - 1) / PZY);
    hipLaunchKernelGGL(k_pzeta<PZ_BUOY>, g, dim3(PZX, PZY), 0, st, G, S, step);
}
void launch_advection_scalar(const Grid &G, const StepPtrs &S, int step, int field, hipStream_t st) {
    hipLaunchKernelGGL(k_adv_scalar, grid2d(G), dim3(BX, BY), 0, st, G, S, step, field);
}
void launch_advection_qrvt(const Grid &G, const StepPtrs &S, int step, hipStream_t st) {
    hipLaunchKernelGGL(k_qrvt, grid2d(G), dim3(BX, BY), 0, st, G, S, step);
}
void launch_precip(const Grid &G, const StepPtrs &S, hipStream_t st) {
    hipLaunchKernelGGL(k_precip, dim3((G.nx + 255) / 256), dim3(256), 0, st, G, S);
}
void launch_bc_one(const Grid &G, double *a, bool westdown, hipStream_t st) {
    hipLaunchKernelGGL(k_bc_one, grid2d(G), dim3(BX, BY), 0, st, G, a, int(westdown));
}
void launch_thermo4(const Grid &G, const StepPtrs &S, int step, const RedSlot &rs, hipStream_t st) {
    const dim3 g((G.nx + T4X - 1) / T4X, (G.nz + T4Y - 1) / T4Y);
    hipLaunchKernelGGL(k_thermo4, g, dim3(T4X, T4Y), 0, st, G, S, step, rs);
    finalize<6>(rs, nblocks(g), S.s, RED_NEGVAL, S_NEG_RATIO, st);
}
void launch_scale_forcing(const Grid &G, const StepPtrs &S, bool forcing, bool tropical, bool random_kick,
                          hipStream_t st, bool halos) {
    hipLaunchKernelGGL(k_scale_forcing, grid2d(G), dim3(BX, BY), 0, st, G, S, int(forcing), int(tropical),
                       int(random_kick), int(halos));
}
void launch_bc_uw(const Grid &G, const StepPtrs &S, hipStream_t st) {
    hipLaunchKernelGGL(k_bc_uw, grid2d(G), dim3(BX, BY), 0, st, G, S);
}
void launch_bc_all_prog(const Grid &G, const StepPtrs &S, hipStream_t st) {
    hipLaunchKernelGGL(k_bc_all, grid2d(G), dim3(BX, BY), 0, st, G, S);
}
void launch_pubartop(const Grid &G, const StepPtrs &S, int step, hipStream_t st) {
    hipLaunchKernelGGL(k_pubartop, dim3(1), dim3(1024), 0, st, G, S, step);
}

}  // namespace vvmd
