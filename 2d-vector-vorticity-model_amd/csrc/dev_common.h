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

namespace vvm {

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

// Scratch of one grid-wide reduction: per-block partials and an arrival
// counter that the last block resets to 0.
struct RedSlot {
    double *partial;    // [NV][max_blocks]
    unsigned *counter;  // one word
    int max_blocks;
};

// Grid-wide FP64 sums in a fixed order (deterministic): every block writes
// its partials; the block that arrives last sums all partials in block order
// and gets `true` with the totals in tot[] (in all its threads).  Every
// other block gets `false`.
template <int NV>
__device__ inline bool grid_sum_last(double (&v)[NV], const RedSlot &rs, double (&tot)[NV]) {
    __shared__ bool is_last;
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    const int nblk = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<NV>(v);
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) rs.partial[size_t(j) * rs.max_blocks + bid] = v[j];
        __threadfence();
        unsigned prev = atomicAdd(rs.counter, 1u);
        is_last = (prev == unsigned(nblk - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0.;
    // Fixed-order partial sums: thread t sums blocks t, t+nthr, ...
    for (int b = tid; b < nblk; b += nthr) {
#pragma unroll
        for (int j = 0; j < NV; j++)
            acc[j] += __hip_atomic_load(rs.partial + size_t(j) * rs.max_blocks + b, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(acc);
#pragma unroll
    for (int j = 0; j < NV; j++) tot[j] = acc[j];
    if (tid == 0) *rs.counter = 0u;
    return true;
}

}  // namespace vvm
