// net.h -- the learned initial-guess network (kernels in k_net.hip, host
// side in vvm_net.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "vvm_net.h"

namespace vvmd {

class GuessNet;
// the network behind a C handle, and its device
GuessNet *guess_net_of(vvm_net_t *n);
int guess_net_device(vvm_net_t *n);
// recorded training samples of a timestep handle ([count][4][M][N] fp32:
// three features and the target), after its stream has drained
int dev_samples(vvm_dev_t *h, const float **buf, int *N, int *M, int *count, int *device);

constexpr int NET_L = 6;       // layers
constexpr int NET_C = 32;      // channels stored per pixel
constexpr int NET_T = 9;       // 3x3 taps
constexpr int NET_CIN = 3;     // input features
constexpr int NET_LAYER_PARAMS = NET_C * NET_C * NET_T + NET_C;  // W[co][ci][t], b[co]
constexpr int NET_PARAMS = NET_L * NET_LAYER_PARAMS;
constexpr int NET_PACK = 18 * 64 * 8;  // packed bf16 weights of one layer (MFMA lane order)
constexpr int NET_WGRAD_PARTIAL = NET_T * NET_C * NET_C + NET_C;

// nb samples of h rows x w columns; columns wrap when periodic (the model
// domain), else read as zero (training crops); rows outside read as zero.
struct NetGeo {
    int nb, h, w, periodic;
};

// one training crop: rows y0.. and (periodic) columns x0.. of a recorded
// sample src ([4][M][N]: three features, target)
struct CropDesc {
    const float *src;
    int N, M, y0, x0;
};
void net_gather(int nb, int ch, int cw, int margin, const CropDesc *d, float *feat, float *tgt, float *mask,
                hipStream_t st);
// Inference ends of the network fused with the guess: the first layer
// reads the three raw FP32 feature planes and scales feature c by
// 1/RMS = 1/sqrt(sumsq[c]/npts); the last layer writes the guess
// x = xe + sqrt(sumsq[1]/npts) * out straight into the solver's unknown
// (row y of the network is row y+1 of x, N columns), xe the extrapolation
// of the previous solutions (extrap_at).
struct NetFeatIn {
    const float *feat;
    const double *sumsq;
    double npts;
};
struct NetApply {
    const double *x1, *x2, *x3, *sumsq;
    double *x;
    int N;
    double npts;
    int ord;
};

// Polynomial extrapolation of the last solutions x1 = x_{t-1}, x2, x3 at
// element a: order 1 is 2 x1 - x2, order 2 is 3 (x1 - x2) + x3.
__host__ __device__ inline double extrap_at(const double *x1, const double *x2, const double *x3, int ord, size_t a) {
    return (ord == 2) ? 3. * (x1[a] - x2[a]) + x3[a] : 2. * x1[a] - x2[a];
}
void net_conv_first_features(const NetGeo &g, const NetFeatIn &fin, const __bf16 *wpk, const float *bias, __bf16 *A,
                             hipStream_t st);
void net_conv_last_apply(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, const NetApply &ap,
                         hipStream_t st);
void net_pack(const float *params, __bf16 *fwd, __bf16 *bwd, hipStream_t st);
// tile height (2, 4 or 8 rows) of the convolution kernels, process-wide
void net_set_conv_rows(int rows);
int net_conv_rows();
void net_set_layer_pairs(bool on);
bool net_layer_pairs();
void net_pack_input(const NetGeo &g, const float *feat, __bf16 *x, hipStream_t st);
void net_conv_hidden(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, __bf16 *A, __bf16 *Z,
                     hipStream_t st);
void net_conv_last(const NetGeo &g, const __bf16 *in, const __bf16 *wpk, const float *bias, float *out,
                   hipStream_t st);
void net_conv_bwd(const NetGeo &g, const __bf16 *dz, const __bf16 *wpk_bwd, const __bf16 *zbelow, __bf16 *dzbelow,
                  hipStream_t st);
// two hidden layers in one pass (inference only: no pre-activations kept)
void net_conv_pair(const NetGeo &g, const __bf16 *in, const __bf16 *wa, const float *ba, const __bf16 *wb,
                   const float *bb, __bf16 *out, hipStream_t st);
void net_wgrad(const NetGeo &g, const __bf16 *x, const __bf16 *dz, float *partial, int nblk, float *grad, int layer,
               hipStream_t st);
void net_loss(long n, const float *out, const float *target, const float *mask, float inv_msum, __bf16 *dz,
              double *lpart, hipStream_t st);
void net_adam(float *w, float *m, float *v, const float *grad, int n, float lr, float b1, float b2, float eps,
              float c1, float c2, hipStream_t st);

// Activation buffers of one inference caller (each timestep handle has its
// own, so handles that share a network never share activations).
struct NetScratch {
    __bf16 *a0 = nullptr, *a1 = nullptr, *z = nullptr;
    long cap = 0;
    int reserve(long npx);
    void release();
};

// Parameters, optimiser state and activation buffers of one network.
class GuessNet {
   public:
    // the solver's guess in one pass: raw features in, guess written to ap.x
    int infer_guess(const NetGeo &g, const NetFeatIn &fin, const NetApply &ap, NetScratch &sc,
                    hipStream_t s) const;
    int hidden_step(const NetGeo &g, int l, const __bf16 *in, __bf16 *out, hipStream_t s,
                    bool pair_ok = true) const;
    int init(int device, hipStream_t st);
    void destroy();
    // activation buffers for up to npx pixels (grown on demand)
    int reserve(long npx, bool training);
    // out[nb][h][w] (device fp32) from feat[nb][3][h][w] (device fp32), on
    // stream s (the network's own stream when null)
    int forward(const NetGeo &g, const float *feat, float *out, bool keep, hipStream_t s = nullptr);
    // loss and parameter gradient (into grad) for the last forward with keep
    int backward(const NetGeo &g, const float *out, const float *target, const float *mask, double msum,
                 double *loss);
    void adam_step(float lr, float b1 = 0.9f, float b2 = 0.999f, float eps = 1e-8f);
    void repack();  // bf16 copies after the fp32 parameters changed
    void reset_adam();

    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;  // NET_PARAMS each
    __bf16 *wf = nullptr, *wb = nullptr;                                  // NET_L * NET_PACK each
    int adam_t = 0;
    hipStream_t st = nullptr;

   private:
    long cap = 0;
    bool cap_train = false;
    __bf16 *act[NET_L] = {};   // X0 (input), A1..A5
    __bf16 *zpre[NET_L] = {};  // Z1..Z5 (training)
    __bf16 *dz[2] = {};        // pre-activation gradients (ping-pong)
    float *partial = nullptr;  // weight-gradient partials
    double *lpart = nullptr;   // loss partials
    double *hloss = nullptr;   // pinned
    int nblk_wgrad = 128;
};

}  // namespace vvmd
