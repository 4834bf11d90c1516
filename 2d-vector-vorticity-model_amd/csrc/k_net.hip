// k_net.hip -- the learned initial-guess network on the device: 3x3
// convolutions as implicit GEMMs on MFMA (bf16 operands, fp32
// accumulation), GELU, the weight gradient, Adam.
//
// Tensors are pixel-major with 32 bf16 channels per pixel ([b][y][x][32],
// 64 bytes a pixel); the network input uses channels 0..2 and keeps the
// rest zero, and layer 6 produces channel 0 only (its other weights stay
// zero), so every layer is the same 32 -> 32 convolution.
//
// Forward / backward-data kernel (k_conv): a workgroup of 4 wavefronts
// computes a 64 x 2, 4 or 8 pixel tile; each wavefront owns one to four
// v_mfma_f32_32x32x16_bf16 output blocks of 32 output channels x 32 pixels,
// and walks K = 9 taps x 32 input channels in 18 steps of 16.  The input
// tile with its one-pixel halo is staged once in LDS and read as the B
// operand.  The A operand (weights) is pre-packed in the exact lane order
// of the instruction, so a lane fetches its 8 weights as one 16-byte load
// and uses them for all its blocks.  With channels as the rows of the
// result, a lane ends with 16 channels of one pixel and stores them as four
// 8-byte runs.
//
// Weight-gradient kernel (k_conv_wgrad): one wavefront per tap; the GEMM
// runs over pixels (K) with the input tile and the output gradient staged
// channel-major in LDS; partial sums per workgroup are reduced by
// k_wgrad_reduce.
#include <algorithm>
#include <mutex>
#include <cstdint>

#include "net.h"

namespace vvmd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr int TW = 64, TH = 2;                 // weight-gradient tile (columns x rows)
constexpr int HW = TW + 2, HH = TH + 2;        // its haloed tile
constexpr int FTW = 64;                        // forward / backward-data tile width; its height
constexpr int FHW = FTW + 2;                   // (2, 4 or 8 rows) is a template parameter
__device__ inline double inv_rms(double s2, double n) { return s2 > 0. ? 1. / sqrt(s2 / n) : 0.; }
constexpr int PXS = 40;                        // LDS pixel stride in bf16 (80 bytes)
constexpr float GC = 0.7978845608028654f;      // sqrt(2/pi)
constexpr float GA = 0.044715f;

// GELU (tanh form) written as z * sigmoid(2u), u = sqrt(2/pi) (z + 0.044715
// z^3), since 0.5 (1 + tanh u) = 1 / (1 + e^{-2u}): one hardware exp2 and
// one hardware reciprocal (v_exp_f32, v_rcp_f32, ~1 ulp each) and three
// multiply-adds.  The epilogue evaluates it for every output channel of
// every pixel, which made the library's full-accuracy tanhf and IEEE
// division the costliest part of a convolution; the ~1e-7 relative error
// here is far below the bf16 rounding the activations get next.  Both ends
// saturate: e -> inf gives z * 0, e -> 0 gives z.
constexpr float LOG2E = 1.4426950408889634f;
constexpr float S1 = -2.f * GC * LOG2E, S3 = S1 * GA;  // -2u log2(e) = z (S1 + S3 z^2)
__device__ inline float sigm2u(float z) {
    const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(S3, z * z, S1));
    return __builtin_amdgcn_rcpf(1.f + e);
}
__device__ inline float gelu(float z) { return z * sigm2u(z); }
// d/dz [z s(z)] with s = sigmoid(2u): s + z s (1 - s) 2u'
__device__ inline float gelu_d(float z) {
    const float sg = sigm2u(z);
    return sg + z * sg * (1.f - sg) * (2.f * GC) * __builtin_fmaf(3.f * GA, z * z, 1.f);
}

__device__ inline __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }
__device__ inline float from_bf16(__bf16 x) { return static_cast<float>(x); }

// global pixel index of haloed tile position (ry, rx) or -1 for padding
__device__ inline long halo_src(const NetGeo &g, int b, int y0, int x0, int ry, int rx) {
    const int y = y0 + ry - 1;
    int x = x0 + rx - 1;
    if (y < 0 || y >= g.h) return -1;
    if (x < 0 || x >= g.w) {
        if (!g.periodic) return -1;
        x = (x + g.w) % g.w;
    }
    return (long(b) * g.h + y) * g.w + x;
}

// MODE 0: hidden layer (bias, GELU; writes A and Z), 1: last layer (bias;
// channel 0 to fp32), 2: backward data (times gelu'(Z of the layer below);
// writes that layer's pre-activation gradient).
//
// MODE 3: last layer writing the solver's guess (NetApply).  IN 1: the
// input is the three raw FP32 feature planes (NetFeatIn), scaled and
// rounded to bf16 while staged; the other 29 channels are zero.
//
// A workgroup of 4 wavefronts computes 64 x FTH pixel tiles from the
// (66 x FTH+2)-pixel haloed input staged in LDS; a tile holds 2 FTH MFMA
// output blocks of 32 channels x 32 pixels, FTH / 2 per wavefront, and a
// wavefront applies each weight fragment to all of its blocks.  The
// workgroups stay resident and walk the tiles with a stride of the grid:
// a lane loads its 18 weight fragments once, and the next tile's input is
// fetched into registers while the current tile's products run, so neither
// the weights' nor the input's memory latency sits between two tiles.
template <int MODE, int IN = 0, int FTH = 4>
__global__ void __launch_bounds__(256) k_conv(NetGeo g, const __bf16 *__restrict__ in,
                                              const bf16x8 *__restrict__ wpk, const float *__restrict__ bias,
                                              __bf16 *__restrict__ outA, __bf16 *__restrict__ outZ,
                                              float *__restrict__ outF, const __bf16 *__restrict__ zbelow,
                                              NetFeatIn fin, NetApply ap) {
    constexpr int FHH = FTH + 2, BPW = FTH / 2;  // haloed rows, output blocks per wavefront
    constexpr int NCH = FHH * FHW * 4;           // 16-byte chunks of a haloed tile (IN 0)
    constexpr int NPX = FHH * FHW;               // its pixels (IN 1: one thread a pixel)
    constexpr int NPF = IN == 0 ? (NCH + 255) / 256 : (NPX + 255) / 256;
    __shared__ __attribute__((aligned(16))) __bf16 tile[FHH * FHW * PXS];
    const int gx = (g.w + FTW - 1) / FTW, gy = (g.h + FTH - 1) / FTH;
    const int ntile = gx * gy * g.nb;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 31, gk = lane >> 5;
    // IN 1: channels 3..31 of the input are zero, so the steps over input
    // channels 16..31 (odd s) add nothing and are skipped
    bf16x8 wr[18];
#pragma unroll
    for (int s = 0; s < 18; s++)
        if (IN == 0 || (s & 1) == 0) wr[s] = wpk[s * 64 + lane];
    float bv[16];
    if (MODE == 0) {
#pragma unroll
        for (int r = 0; r < 16; r++) bv[r] = bias[8 * (r >> 2) + 4 * gk + (r & 3)];
    }
    double sc[3];
    if (IN == 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) sc[c] = inv_rms(fin.sumsq[c], fin.npts);
    }
    const long hw = long(g.h) * g.w;
    bf16x8 pf[IN == 0 ? NPF : 1];     // IN 0: prefetched chunks
    float pff[IN == 1 ? NPF : 1][3];  // IN 1: prefetched feature values
    auto fetch = [&](int t) {
        const int b = t / (gx * gy), y0 = (t / gx) % gy * FTH, x0 = t % gx * FTW;
#pragma unroll
        for (int k = 0; k < NPF; k++) {
            if (IN == 0) {
                const int c = threadIdx.x + 256 * k, px = c >> 2, part = c & 3;
                const long src = c < NCH ? halo_src(g, b, y0, x0, px / FHW, px % FHW) : -1;
                pf[k] = src >= 0 ? *reinterpret_cast<const bf16x8 *>(in + src * NET_C + part * 8) : bf16x8{};
            } else {
                const int px = threadIdx.x + 256 * k;
                const long src = px < NPX ? halo_src(g, b, y0, x0, px / FHW, px % FHW) : -1;
                const long q = src - long(b) * hw;
#pragma unroll
                for (int c = 0; c < 3; c++) pff[k][c] = src >= 0 ? fin.feat[(long(b) * 3 + c) * hw + q] : 0.f;
            }
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int k = 0; k < NPF; k++) {
            if (IN == 0) {
                const int c = threadIdx.x + 256 * k;
                if (c < NCH) *reinterpret_cast<bf16x8 *>(tile + (c >> 2) * PXS + (c & 3) * 8) = pf[k];
            } else {
                const int px = threadIdx.x + 256 * k;
                if (px < NPX) {
                    bf16x8 lo = {}, z = {};
#pragma unroll
                    for (int c = 0; c < 3; c++) lo[c] = to_bf16(float(double(pff[k][c]) * sc[c]));
                    bf16x8 *dst = reinterpret_cast<bf16x8 *>(tile + px * PXS);
                    dst[0] = lo;
                    dst[1] = z;
                    dst[2] = z;
                    dst[3] = z;
                }
            }
        }
    };
    int t = blockIdx.x;
    if (t < ntile) fetch(t);
    for (; t < ntile; t += gridDim.x) {
        const int b = t / (gx * gy), y0 = (t / gx) % gy * FTH, x0 = t % gx * FTW;
        __syncthreads();  // the previous tile's reads of the LDS tile are done
        put();
        __syncthreads();
        if (t + int(gridDim.x) < ntile) fetch(t + gridDim.x);
        f32x16 acc[BPW] = {};  // block gq = BPW wv + q: row gq / 2, columns 32 (gq % 2) ..
#pragma unroll
        for (int s = 0; s < 18; s++) {
            if (IN == 1 && (s & 1)) continue;
            const int tap = s >> 1, hh = s & 1;
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
            for (int q = 0; q < BPW; q++) {
                const int gq = BPW * wv + q, ty = gq >> 1, cb = (gq & 1) * 32;
                const int px = (ty + 1 + dy) * FHW + (cb + i + 1 + dx);
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(tile + px * PXS + hh * 16 + gk * 8);
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wr[s], a, acc[q], 0, 0, 0);
            }
        }
        // weights as the A operand, pixels as B: D[m][n] with n = lane % 32
        // the pixel and m = 8 (r / 4) + 4 (lane / 32) + r % 4 the output
        // channel, so a lane holds four runs of 4 consecutive channels of one
        // pixel
#pragma unroll
        for (int q = 0; q < BPW; q++) {
            const int gq = BPW * wv + q, y = y0 + (gq >> 1), x = x0 + (gq & 1) * 32 + i;
            if (y >= g.h || x >= g.w) continue;
            const long p = (long(b) * g.h + y) * g.w + x;
            if (MODE == 0 || MODE == 2) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const long e = p * NET_C + 8 * j + 4 * gk;  // channels 8 j + 4 gk .. + 3
                    bf16x4 o;
                    if (MODE == 0) {
                        bf16x4 zq;
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            const float z = acc[q][4 * j + c] + bv[4 * j + c];
                            zq[c] = to_bf16(z);
                            o[c] = to_bf16(gelu(z));
                        }
                        if (outZ) *reinterpret_cast<bf16x4 *>(outZ + e) = zq;  // kept for training only
                    } else {
                        const bf16x4 zb = *reinterpret_cast<const bf16x4 *>(zbelow + e);
#pragma unroll
                        for (int c = 0; c < 4; c++)
                            o[c] = to_bf16(acc[q][4 * j + c] * gelu_d(from_bf16(zb[c])));
                    }
                    *reinterpret_cast<bf16x4 *>(outA + e) = o;
                }
            } else if (gk == 0) {  // channel 0 is m = 0: register 0 of lanes 0..31
                if (MODE == 1) {
                    outF[p] = acc[q][0] + bias[0];
                } else {
#pragma clang fp contract(off)
                    const float o = acc[q][0] + bias[0];
                    const size_t a = size_t(y + 1) * ap.N + x;
                    ap.x[a] = extrap_at(ap.x1, ap.x2, ap.x3, ap.ord, a) + sqrt(ap.sumsq[1] / ap.npts) * double(o);
                }
            }
        }
    }
}

// Two hidden layers in one pass (inference): layer a on a 64 x 6 pixel
// region kept in LDS, then layer b on the 62 x 4 pixel tile inside it, so
// the activations between them never go to memory.  The output tile is 62
// columns wide so that layer a's region is exactly two 32-pixel MFMA blocks
// a row (12 blocks, 3 a wavefront); layer b computes 64 columns and keeps
// 62 (8 blocks, 2 a wavefront).  Layer a's values outside the image are the
// zero padding of layer b's input.
constexpr int PTW = 62, PTH = 4;                  // output tile
constexpr int PIW = PTW + 4, PIH = PTH + 4;       // input tile (two-pixel halo)
constexpr int PMW = PTW + 4, PMH = PTH + 2;       // layer-a region (64 columns used, 2 spare)

__global__ void __launch_bounds__(256) k_conv_pair(NetGeo g, const __bf16 *__restrict__ in,
                                                   const bf16x8 *__restrict__ wa, const float *__restrict__ ba,
                                                   const bf16x8 *__restrict__ wb, const float *__restrict__ bb,
                                                   __bf16 *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) __bf16 tin[PIH * PIW * PXS];
    __shared__ __attribute__((aligned(16))) __bf16 mid[PMH * PMW * PXS];
    const int b = blockIdx.z, y0 = blockIdx.y * PTH, x0 = blockIdx.x * PTW;
    bf16x8 wr[18];  // layer a's weight fragments, requested before the staging
#pragma unroll
    for (int s = 0; s < 18; s++) wr[s] = wa[s * 64 + (threadIdx.x & 63)];
    // input rows y0-2 .., columns x0-2 .. (halo_src takes the tile origin one
    // pixel inside the halo)
    for (int c = threadIdx.x; c < PIH * PIW * 4; c += blockDim.x) {
        const int px = c >> 2, part = c & 3;
        const long src = halo_src(g, b, y0 - 1, x0 - 1, px / PIW, px % PIW);
        bf16x8 v = {};
        if (src >= 0) v = *reinterpret_cast<const bf16x8 *>(in + src * NET_C + part * 8);
        *reinterpret_cast<bf16x8 *>(tin + px * PXS + part * 8) = v;
    }
    for (int c = threadIdx.x; c < PMH * 2 * 4; c += blockDim.x) {  // spare columns 64, 65 of mid: zero
        const int row = c / 8, col = 64 + (c / 4) % 2, part = c % 4;
        *reinterpret_cast<bf16x8 *>(mid + (row * PMW + col) * PXS + part * 8) = bf16x8{};
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 31, gk = lane >> 5;
    {  // layer a: blocks 3 wv .. 3 wv + 2 of the 12 (row gq / 2, half gq % 2)
        f32x16 acc[3] = {};
#pragma unroll
        for (int s = 0; s < 18; s++) {
            const int t = s >> 1, hh = s & 1;
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            const bf16x8 w = wr[s];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int gq = 3 * wv + q, ty = gq >> 1, cb = (gq & 1) * 32;
                const int px = (ty + 1 + dy) * PIW + (cb + i + 1 + dx);
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(tin + px * PXS + hh * 16 + gk * 8);
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, a, acc[q], 0, 0, 0);
            }
        }
#pragma unroll
        for (int s = 0; s < 18; s++) wr[s] = wb[s * 64 + lane];  // layer b's, during the epilogue
        float bv[16];  // D[channel][pixel] as in k_conv
#pragma unroll
        for (int r = 0; r < 16; r++) bv[r] = ba[8 * (r >> 2) + 4 * gk + (r & 3)];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int gq = 3 * wv + q, ty = gq >> 1, cb = (gq & 1) * 32;
            const int y = y0 - 1 + ty, x = x0 - 1 + cb + i;
            const bool in_img = y >= 0 && y < g.h && (g.periodic || (x >= 0 && x < g.w));
#pragma unroll
            for (int j = 0; j < 4; j++) {
                bf16x4 o;
#pragma unroll
                for (int c = 0; c < 4; c++) o[c] = to_bf16(in_img ? gelu(acc[q][4 * j + c] + bv[4 * j + c]) : 0.f);
                *reinterpret_cast<bf16x4 *>(mid + (ty * PMW + cb + i) * PXS + 8 * j + 4 * gk) = o;
            }
        }
    }
    __syncthreads();
    {  // layer b: blocks 2 wv, 2 wv + 1 of the 8
        f32x16 acc[2] = {};
#pragma unroll
        for (int s = 0; s < 18; s++) {
            const int t = s >> 1, hh = s & 1;
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            const bf16x8 w = wr[s];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int gq = 2 * wv + q, ty = gq >> 1, cb = (gq & 1) * 32;
                const int px = (ty + 1 + dy) * PMW + (cb + i + 1 + dx);
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(mid + px * PXS + hh * 16 + gk * 8);
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, a, acc[q], 0, 0, 0);
            }
        }
        float bv[16];
#pragma unroll
        for (int r = 0; r < 16; r++) bv[r] = bb[8 * (r >> 2) + 4 * gk + (r & 3)];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int gq = 2 * wv + q, y = y0 + (gq >> 1), xo = (gq & 1) * 32 + i, x = x0 + xo;
            if (y >= g.h || xo >= PTW || x >= g.w) continue;
            const long p = (long(b) * g.h + y) * g.w + x;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                bf16x4 o;
#pragma unroll
                for (int c = 0; c < 4; c++) o[c] = to_bf16(gelu(acc[q][4 * j + c] + bv[4 * j + c]));
                *reinterpret_cast<bf16x4 *>(out + p * NET_C + 8 * j + 4 * gk) = o;
            }
        }
    }
}

// Weight gradient of one layer: dW[co][ci][t] = sum_p dZ[p][co] X[p + off_t][ci]
// and db[co] = sum_p dZ[p][co].  Workgroup = 9 wavefronts (one per tap),
// grid-stride over 64 x 2 tiles; per-workgroup partials [blk][9][32 ci][32 co]
// then 32 bias sums.
__global__ void __launch_bounds__(576) k_conv_wgrad(NetGeo g, const __bf16 *__restrict__ x,
                                                    const __bf16 *__restrict__ dz, float *__restrict__ partial) {
    __shared__ __attribute__((aligned(16))) __bf16 xt[NET_C][HH * HW + 8];  // channel-major haloed input
    __shared__ __attribute__((aligned(16))) __bf16 dzt[NET_C][TH * TW];     // channel-major output gradient
    const int lane = threadIdx.x & 63, t = threadIdx.x >> 6;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int gk = lane >> 5, ch = lane & 31;
    const int tx_n = (g.w + TW - 1) / TW, ty_n = (g.h + TH - 1) / TH;
    const long ntile = long(tx_n) * ty_n * g.nb;
    f32x16 acc = {};
    float bsum = 0.f;
    for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int b = int(tile / (long(tx_n) * ty_n));
        const int rem = int(tile % (long(tx_n) * ty_n));
        const int y0 = (rem / tx_n) * TH, x0 = (rem % tx_n) * TW;
        __syncthreads();  // the previous tile's reads are done
        for (int c = threadIdx.x; c < HH * HW * NET_C; c += blockDim.x) {
            const int px = c / NET_C, ci = c % NET_C;
            const long src = halo_src(g, b, y0, x0, px / HW, px % HW);
            xt[ci][px] = (src >= 0) ? x[src * NET_C + ci] : to_bf16(0.f);
        }
        for (int c = threadIdx.x; c < TH * TW * NET_C; c += blockDim.x) {
            const int px = c / NET_C, co = c % NET_C;
            const int yy = y0 + px / TW, xx = x0 + px % TW;
            __bf16 v = to_bf16(0.f);
            if (yy < g.h && xx < g.w) v = dz[((long(b) * g.h + yy) * g.w + xx) * NET_C + co];
            dzt[co][px] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TH * TW / 16; kk++) {
            const int ty = kk / (TW / 16), c0 = (kk % (TW / 16)) * 16 + 8 * gk;
            bf16x8 a, bb;
            const int base = (ty + 1 + dy) * HW + c0 + 1 + dx;
#pragma unroll
            for (int e = 0; e < 8; e++) a[e] = xt[ch][base + e];
            bb = *reinterpret_cast<const bf16x8 *>(&dzt[ch][ty * TW + c0]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
        }
        if (t == 0 && lane < NET_C)
            for (int px = 0; px < TH * TW; px++) bsum += from_bf16(dzt[lane][px]);
    }
    // D[m][n]: m = ci, n = co
    float *dst = partial + size_t(blockIdx.x) * NET_WGRAD_PARTIAL;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int ci = 8 * (r >> 2) + 4 * gk + (r & 3), co = lane & 31;
        dst[(t * NET_C + ci) * NET_C + co] = acc[r];
    }
    if (t == 0 && lane < NET_C) dst[NET_T * NET_C * NET_C + lane] = bsum;
}

// grad[co][ci][t] and grad bias from nblk partials; masked entries (frozen
// weights) stay zero.
__global__ void k_wgrad_reduce(const float *__restrict__ partial, int nblk, float *__restrict__ grad, int layer) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NET_LAYER_PARAMS) return;
    int src;
    bool live;
    if (e < NET_C * NET_C * NET_T) {
        const int co = e / (NET_C * NET_T), ci = (e / NET_T) % NET_C, t = e % NET_T;
        src = (t * NET_C + ci) * NET_C + co;
        live = !(layer == 0 && ci >= NET_CIN) && !(layer == NET_L - 1 && co >= 1);
    } else {
        const int co = e - NET_C * NET_C * NET_T;
        src = NET_T * NET_C * NET_C + co;
        live = !(layer == NET_L - 1 && co >= 1);
    }
    double s = 0.;
    for (int k = 0; k < nblk; k++) s += partial[size_t(k) * NET_WGRAD_PARTIAL + src];
    grad[size_t(layer) * NET_LAYER_PARAMS + e] = live ? float(s) : 0.f;
}

// Pack bf16 weights in the lane order of the MFMA B operand.
//   forward:  step s = 2 t + hh, lane (n = co, gk): ci = 16 hh + 8 gk + e
//   backward: n = ci of the layer, k = co, tap flipped (8 - t)
__global__ void k_pack(const float *__restrict__ params, __bf16 *__restrict__ fwd, __bf16 *__restrict__ bwd) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over layers x 18 x 64 x 8
    if (idx >= NET_L * NET_PACK) return;
    const int l = idx / NET_PACK, r = idx % NET_PACK;
    const int s = r / 512, lane = (r / 8) % 64, e = r % 8;
    const int t = s >> 1, hh = s & 1, n = lane & 31, gk = lane >> 5, k = 16 * hh + 8 * gk + e;
    const float *W = params + size_t(l) * NET_LAYER_PARAMS;
    fwd[idx] = to_bf16(W[(n * NET_C + k) * NET_T + t]);
    bwd[idx] = to_bf16(W[(k * NET_C + n) * NET_T + (8 - t)]);
}

__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ grad, int n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// network input from channel-major fp32 features [b][3][h][w]
__global__ void k_pack_input(NetGeo g, const float *__restrict__ feat, __bf16 *__restrict__ x) {
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    const long hw = long(g.h) * g.w, n = hw * g.nb;
    if (p >= n) return;
    const long b = p / hw, q = p % hw;
    bf16x8 lo = {}, z = {};
    lo[0] = to_bf16(feat[(b * 3 + 0) * hw + q]);
    lo[1] = to_bf16(feat[(b * 3 + 1) * hw + q]);
    lo[2] = to_bf16(feat[(b * 3 + 2) * hw + q]);
    bf16x8 *dst = reinterpret_cast<bf16x8 *>(x + p * NET_C);
    dst[0] = lo;
    dst[1] = z;
    dst[2] = z;
    dst[3] = z;
}

// d loss / d out for loss = sum mask (out - target)^2 / msum, as the
// channel-0 pre-activation gradient of layer 6; loss partial per block.
__global__ void __launch_bounds__(256) k_loss(long n, const float *__restrict__ out, const float *__restrict__ target,
                                              const float *__restrict__ mask, float inv_msum, __bf16 *__restrict__ dz,
                                              double *__restrict__ lpart) {
    __shared__ double red[256];
    const long p = blockIdx.x * long(blockDim.x) + threadIdx.x;
    double l = 0.;
    if (p < n) {
        const float e = out[p] - target[p];
        l = double(mask[p]) * e * e;
        bf16x8 z = {};
        bf16x8 *dst = reinterpret_cast<bf16x8 *>(dz + p * NET_C);
        bf16x8 v = z;
        v[0] = to_bf16(2.f * mask[p] * e * inv_msum);
        dst[0] = v;
        dst[1] = z;
        dst[2] = z;
        dst[3] = z;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// Training crops: crop b of the batch is rows y0.. (no wrap) and columns
// x0.. (periodic) of one recorded sample.  The loss mask leaves out a
// border of `margin` pixels wherever the crop edge is not a domain edge
// (the network sees zeros there instead of the neighbouring values).
__global__ void k_gather(int ch, int cw, int margin, const CropDesc *__restrict__ d, float *__restrict__ feat,
                         float *__restrict__ tgt, float *__restrict__ mask) {
    const int b = blockIdx.y;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ch * cw) return;
    const int py = p / cw, px = p % cw;
    const CropDesc c = d[b];
    const int yy = c.y0 + py, xx = (c.x0 + px) % c.N;
    const size_t plane = size_t(c.M) * c.N, q = size_t(yy) * c.N + xx;
    const size_t hw = size_t(ch) * cw, o = size_t(b) * hw + p;
#pragma unroll
    for (int k = 0; k < 3; k++) feat[(size_t(b) * 3 + k) * hw + p] = c.src[k * plane + q];
    tgt[o] = c.src[3 * plane + q];
    const bool in_x = px >= margin && px < cw - margin;
    const bool in_lo = (c.y0 == 0) || py >= margin;
    const bool in_hi = (c.y0 + ch == c.M) || py < ch - margin;
    mask[o] = (in_x && in_lo && in_hi) ? 1.f : 0.f;
}

}  // namespace

// ------------------------------------------------------------- launchers

void net_gather(int nb, int ch, int cw, int margin, const CropDesc *d, float *feat, float *tgt, float *mask,
                hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3((ch * cw + 255) / 256, nb), dim3(256), 0, st, ch, cw, margin, d, feat, tgt,
                       mask);
}

void net_pack(const float *params, __bf16 *fwd, __bf16 *bwd, hipStream_t st) {
    const int n = NET_L * NET_PACK;
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, st, params, fwd, bwd);
}

void net_pack_input(const NetGeo &g, const float *feat, __bf16 *x, hipStream_t st) {
    const long n = long(g.nb) * g.h * g.w;
    hipLaunchKernelGGL(k_pack_input, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, g, feat, x);
}

// tile height of the forward / backward-data kernels: 2, 4 or 8 rows
static int g_conv_rows = 2;
void net_set_conv_rows(int rows) { g_conv_rows = rows; }
int net_conv_rows() { return g_conv_rows; }
// inference passes run hidden layers two at a time (k_conv_pair); off by
// default: at 2048 x 256 the pair kernel (60 us) is slower than two
// single-layer launches (2 x 26 us), its 74 KB of LDS allowing only two
// workgroups a CU
static bool g_layer_pairs = false;
void net_set_layer_pairs(bool on) { g_layer_pairs = on; }
bool net_layer_pairs() { return g_layer_pairs; }

// Workgroups of one k_conv instantiation that fit on the current device at
// once, queried once per device and instantiation.  Several host threads
// (ensemble members sharing one network) may launch at the same time, so the
// table is filled under a lock.
template <int MODE, int IN, int FTH>
static long conv_resident() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static long cached[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (cached[dev] == 0) {
        int ncu = 0, per_cu = 0;
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_conv<MODE, IN, FTH>),
                                                     256, 0);
        cached[dev] = long(std::max(per_cu, 1)) * std::max(ncu, 1);
    }
    return cached[dev];
}

template <int MODE, int IN, int FTH>
static void conv_launch(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, __bf16 *A,
                        __bf16 *Z, float *F, const __bf16 *zb, const NetFeatIn &fin, const NetApply &ap,
                        hipStream_t st) {
    // resident workgroups: as many as fit on the device at once, at most one a tile
    const long ntile = long((g.w + FTW - 1) / FTW) * ((g.h + FTH - 1) / FTH) * g.nb;
    const dim3 grid(unsigned(std::min<long>(ntile, conv_resident<MODE, IN, FTH>())));
    hipLaunchKernelGGL((k_conv<MODE, IN, FTH>), grid, dim3(256), 0, st, g, in,
                       reinterpret_cast<const bf16x8 *>(wpk), bias, A, Z, F, zb, fin, ap);
}

template <int MODE, int IN = 0>
static void conv(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, __bf16 *A, __bf16 *Z,
                 float *F, const __bf16 *zb, const NetFeatIn &fin, const NetApply &ap, hipStream_t st) {
    if (g_conv_rows == 2)
        conv_launch<MODE, IN, 2>(g, in, wpk, bias, A, Z, F, zb, fin, ap, st);
    else if (g_conv_rows == 8)
        conv_launch<MODE, IN, 8>(g, in, wpk, bias, A, Z, F, zb, fin, ap, st);
    else
        conv_launch<MODE, IN, 4>(g, in, wpk, bias, A, Z, F, zb, fin, ap, st);
}

void net_conv_hidden(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, __bf16 *A, __bf16 *Z,
                     hipStream_t st) {
    conv<0>(g, in, wpk, bias, A, Z, nullptr, nullptr, NetFeatIn{}, NetApply{}, st);
}

void net_conv_last(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, float *out,
                   hipStream_t st) {
    conv<1>(g, in, wpk, bias, nullptr, nullptr, out, nullptr, NetFeatIn{}, NetApply{}, st);
}

void net_conv_first_features(const NetGeo &g, const NetFeatIn &fin, const __bf16 *wpk, const float *bias, __bf16 *A,
                             hipStream_t st) {
    conv<0, 1>(g, nullptr, wpk, bias, A, nullptr, nullptr, nullptr, fin, NetApply{}, st);
}

void net_conv_last_apply(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, const NetApply &ap,
                         hipStream_t st) {
    conv<3>(g, in, wpk, bias, nullptr, nullptr, nullptr, nullptr, NetFeatIn{}, ap, st);
}

void net_conv_pair(const NetGeo &g, const __bf16 *in, const __bf16 *wa, const float *ba, const __bf16 *wb,
                   const float *bb, __bf16 *out, hipStream_t st) {
    const dim3 grid((g.w + PTW - 1) / PTW, (g.h + PTH - 1) / PTH, g.nb);
    hipLaunchKernelGGL(k_conv_pair, grid, dim3(256), 0, st, g, in, reinterpret_cast<const bf16x8 *>(wa), ba,
                       reinterpret_cast<const bf16x8 *>(wb), bb, out);
}

void net_conv_bwd(const NetGeo &g, const __bf16 *dz, const __bf16 *wpk_bwd, const __bf16 *zbelow, __bf16 *dzbelow,
                  hipStream_t st) {
    conv<2>(g, dz, wpk_bwd, nullptr, dzbelow, nullptr, nullptr, zbelow, NetFeatIn{}, NetApply{}, st);
}

void net_wgrad(const NetGeo &g, const __bf16 *x, const __bf16 *dz, float *partial, int nblk, float *grad, int layer,
               hipStream_t st) {
    hipLaunchKernelGGL(k_conv_wgrad, dim3(nblk), dim3(576), 0, st, g, x, dz, partial);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((NET_LAYER_PARAMS + 255) / 256), dim3(256), 0, st, partial, nblk, grad,
                       layer);
}

void net_loss(long n, const float *out, const float *target, const float *mask, float inv_msum, __bf16 *dz,
              double *lpart, hipStream_t st) {
    hipLaunchKernelGGL(k_loss, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, n, out, target, mask, inv_msum, dz,
                       lpart);
}

void net_adam(float *w, float *m, float *v, const float *grad, int n, float lr, float b1, float b2, float eps,
              float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, w, m, v, grad, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace vvmd
