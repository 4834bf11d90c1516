// dev_common.h -- device layout and helpers shared by the VVM HIP kernels.
//
// Every 2-D model field lives x-contiguous on the device: element (i,k) of
// the reference's [i][k] array is at k*pitch + i, with i = 0..nx-1 and
// k = 0..nz-1 including the ghost cells.  Rows are padded to `pitch`
// elements (a multiple of 64) so every row starts 512-byte aligned.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace vvmd {

struct Grid {
    int nx, nz, pitch;
    double dx, dz, dt;
    double rdx, r2dx, rdz, r2dz, rdx2, rdz2;
    double GRAVITY, Cp, Cv, Rd, Lv, P0;
};

__host__ __device__ inline size_t ix(const Grid &G, int i, int k) { return size_t(k) * G.pitch + i; }

// Stores v at interior point (i,k) and at every ghost cell that
// BoundaryProcess2D_center (src/Boundary.cpp:9-19) fills from it: the
// periodic column image for k in 1..nz-2, the zero-gradient row image at
// k = 1 / nz-2, and the corner that takes both.
__device__ inline void store_center(double *a, const Grid &G, int i, int k, double v) {
    a[ix(G, i, k)] = v;
    const bool west_img = (i == G.nx - 2), east_img = (i == 1);
    if (west_img) a[ix(G, 0, k)] = v;
    if (east_img) a[ix(G, G.nx - 1, k)] = v;
    if (k == 1) {
        a[ix(G, i, 0)] = v;
        if (west_img) a[ix(G, 0, 0)] = v;
        if (east_img) a[ix(G, G.nx - 1, 0)] = v;
    }
    if (k == G.nz - 2) {
        a[ix(G, i, G.nz - 1)] = v;
        if (west_img) a[ix(G, 0, G.nz - 1)] = v;
        if (east_img) a[ix(G, G.nx - 1, G.nz - 1)] = v;
    }
}

// Stores v at (i,k), k in 2..nz-2, plus its periodic column image, as
// BoundaryProcess2D_westdown (src/Boundary.cpp:21-31) does; rows 0, 1 and
// nz-1 are zero and are written by the caller.
__device__ inline void store_westdown(double *a, const Grid &G, int i, int k, double v) {
    a[ix(G, i, k)] = v;
    if (i == G.nx - 2) a[ix(G, 0, k)] = v;
    if (i == 1) a[ix(G, G.nx - 1, k)] = v;
}

// ---------------------------------------------------------------- sums

constexpr int kWave = 64;

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sums of NV values; the totals are returned to every thread.
// blockDim must be a multiple of 64 and at most 1024.
template <int NV, typename T>
__device__ inline void block_sum(T (&v)[NV]) {
    __shared__ T sh[NV][16];
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid % kWave, wid = tid / kWave, nw = nthr / kWave;
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = wave_sum(v[j]);
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) sh[j][wid] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; j++) {
        T s = 0;
        for (int w = 0; w < nw; w++) s += sh[j][w];
        v[j] = s;
    }
    __syncthreads();
}

// Grid-wide FP64 sums are two launches: every block of the producing kernel
// writes its block sums with plain stores (block_partials), then one
// workgroup (k_finalize) adds them in block order and derives the scalars
// the next kernels read.  The kernel boundary orders the two; no fence or
// atomic is needed inside the producer.  (A last-block-done pattern needs an
// agent-scope release per block, which on a GPU with one L2 per XCD costs an
// L2 write-back each time.)  The order is fixed, so results are
// deterministic from run to run.
struct RedSlot {
    double *partial;  // [8][max_blocks]
    int max_blocks;
};

template <int NV>
__device__ inline void block_partials(double (&v)[NV], const RedSlot &rs) {
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<NV>(v);
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) rs.partial[size_t(j) * rs.max_blocks + bid] = v[j];
    }
}

// Scalars of the preconditioned CG (doubles in its scalar block).
// P_RZ / P_RZ2: r.z of odd / even CG iterations, P_ALPHA / P_ALPHA2 the
// step lengths of the single-reduction form (two slots each, so that the
// kernel forming the new value never overwrites the one its other
// workgroups still read).  Slot 1 is unused.
enum { P_RZ = 0, P_SPARE, P_ALPHA, P_ALPHA2, P_RNORM2, P_BNORM2, P_RNORM2_0, P_RZ2, NPSCAL = 8 };

// What k_finalize does with the totals.
enum RedMode {
    RED_STORE = 0,   // sc[dst+j] = tot[j] (and sc[dst2+j] when dst2 >= 0)
    RED_NEGVAL,      // NegativeValueProcess ratios of 3 fields into sc[dst+q]
    RED_RZ_INIT      // CG: sc[P_RZ] = rz
};

namespace {
template <int NV>
__global__ void __launch_bounds__(256) k_finalize(RedSlot rs, int nblk, double *sc, int mode, int dst, int dst2,
                                                  double *hout) {
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0.;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
        for (int j = 0; j < NV; j++) acc[j] += rs.partial[size_t(j) * rs.max_blocks + b];
    }
    block_sum<NV>(acc);
    if (threadIdx.x != 0) return;
    switch (mode) {
        case RED_STORE:
            for (int j = 0; j < NV; j++) {
                sc[dst + j] = acc[j];
                if (dst2 >= 0) sc[dst2 + j] = acc[j];
                // host-mapped coherent memory: the host reads it after an
                // event, with no copy in the stream
                if (hout) __hip_atomic_store(hout + j, acc[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            break;
        case RED_NEGVAL:
            // MicroPhysics.cpp:129-131: no rescale when nothing is positive or
            // the negatives outweigh the positives (marked by a ratio < 0).
            for (int q = 0; q < NV / 2; q++) {
                const double P = acc[2 * q], N = acc[2 * q + 1];
                sc[dst + q] = (P == 0. || fabs(N) > P) ? -1. : 1. - fabs(N / P);
            }
            break;
        case RED_RZ_INIT:
            sc[P_RZ] = acc[0];
            break;
    }
}
}  // namespace

template <int NV>
inline void finalize(const RedSlot &rs, int nblk, double *sc, int mode, int dst, hipStream_t st, int dst2 = -1,
                     double *hout = nullptr) {
    hipLaunchKernelGGL(k_finalize<NV>, dim3(1), dim3(256), 0, st, rs, nblk, sc, mode, dst, dst2, hout);
}

// Sum of a grid's block partials (row 0 of rs, nblk of them) inside every
// block of the consuming kernel, in a fixed order, so that all blocks agree
// bit for bit; this saves a one-block k_finalize launch between producer
// and consumer when nblk is small.
__device__ inline double block_total(const RedSlot &rs, int nblk) {
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    double v[1] = {0.};
    for (int b = tid; b < nblk; b += nthr) v[0] += rs.partial[b];
    block_sum<1>(v);
    return v[0];
}

inline int nblocks(dim3 g) { return int(g.x * g.y * g.z); }

}  // namespace vvmd
