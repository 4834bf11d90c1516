/*
 * convnet.c -- CPU restatement of the learned initial-guess network.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the device
 * network's forward and backward kernels: only tests/ may load it.  The
 * product never links, loads or calls it.
 *
 * Network (all layers 3x3, 32 channels; layer 1 reads 3 input channels,
 * layer 6 writes 1 output channel, GELU between layers):
 *   X0 = bf16(input), Z_l = b_l + W_l * X_{l-1},
 *   A_l = bf16(gelu(Z_l)) for l < 6, out = Z_6 (channel 0, fp32).
 * Parameters, flat fp32: for l = 1..6, W_l[co][ci][t] (32x32x9, tap
 * t = 3*(dz+1) + (dx+1)) then b_l[32].  Storage roundings follow the
 * device: weights, activations and back-propagated gradients enter the
 * products as bf16 (round to nearest even); sums here are in double.
 * Pre-activations Z_l are kept as bf16 for gelu' in the backward pass.
 *
 * Geometry: nb samples of h rows (z) x w columns (x), channel-major
 * [b][c][y][x].  Rows outside [0, h) read as zero (the rigid lid); columns
 * wrap when `periodic`, else read as zero.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NL 6
#define NC 32
#define NT 9
#define LAYER_PARAMS (NC * NC * NT + NC)

/* bf16 storage rounding on (the device's arithmetic) or off (exact fp32
 * storage, for finite-difference checks of the gradient formulas) */
static int g_round = 1;
void vo_net_set_rounding(int on) { g_round = on; }

static float bf16r(float x) {
    if (!g_round) return x;
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x; /* inf / nan unchanged */
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    r &= 0xffff0000u;
    float y;
    memcpy(&y, &r, 4);
    return y;
}

static const double GC = 0.7978845608028654; /* sqrt(2/pi) */
static const double GA = 0.044715;

static double gelu(double z) { return 0.5 * z * (1.0 + tanh(GC * (z + GA * z * z * z))); }
static double gelu_d(double z) {
    double t = tanh(GC * (z + GA * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GC * (1.0 + 3.0 * GA * z * z);
}

typedef struct {
    int nb, h, w, periodic;
} geo_t;

/* index of the input pixel at offset tap t from (y, x); -1 when outside */
static long src_index(const geo_t *g, int y, int x, int t) {
    int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= g->h) return -1;
    if (xx < 0 || xx >= g->w) {
        if (!g->periodic) return -1;
        xx = (xx + g->w) % g->w;
    }
    return (long)yy * g->w + xx;
}

static int cin_of(int l) { return l == 1 ? 3 : NC; }
static int cout_of(int l) { return l == NL ? 1 : NC; }

/* acts: (NL) blocks of nb*NC*h*w floats: X0 (bf16 input), A_1..A_5 */
/* zs:   (NL-1) blocks: bf16(Z_1)..bf16(Z_5) */
static void forward(const geo_t *g, const float *params, const float *input, float *out, float *acts, float *zs) {
    const long hw = (long)g->h * g->w, blk = (long)g->nb * NC * hw;
    memset(acts, 0, sizeof(float) * blk * NL);
    for (int b = 0; b < g->nb; b++)
        for (int c = 0; c < 3; c++)
            for (long p = 0; p < hw; p++) acts[((long)b * NC + c) * hw + p] = bf16r(input[((long)b * 3 + c) * hw + p]);
    for (int l = 1; l <= NL; l++) {
        const float *W = params + (long)(l - 1) * LAYER_PARAMS, *bias = W + NC * NC * NT;
        const float *X = acts + (l - 1) * blk;
        for (int b = 0; b < g->nb; b++)
            for (int co = 0; co < cout_of(l); co++)
                for (int y = 0; y < g->h; y++)
                    for (int x = 0; x < g->w; x++) {
                        double z = bias[co];
                        for (int ci = 0; ci < cin_of(l); ci++)
                            for (int t = 0; t < NT; t++) {
                                long q = src_index(g, y, x, t);
                                if (q < 0) continue;
                                z += (double)bf16r(W[(co * NC + ci) * NT + t]) * X[((long)b * NC + ci) * hw + q];
                            }
                        long p = (long)y * g->w + x;
                        if (l < NL) {
                            zs[(l - 1) * blk + ((long)b * NC + co) * hw + p] = bf16r((float)z);
                            acts[l * blk + ((long)b * NC + co) * hw + p] = bf16r((float)gelu((float)z));
                        } else {
                            out[(long)b * hw + p] = (float)z;
                        }
                    }
    }
}

/*
 * vo_net_forward: out[nb][h][w] from input[nb][3][h][w].
 */
int vo_net_forward(int nb, int h, int w, int periodic, const float *params, const float *input, float *out) {
    geo_t g = {nb, h, w, periodic};
    const long blk = (long)nb * NC * h * w;
    float *acts = malloc(sizeof(float) * blk * NL), *zs = malloc(sizeof(float) * blk * (NL - 1));
    if (!acts || !zs) {
        free(acts);
        free(zs);
        return -1;
    }
    forward(&g, params, input, out, acts, zs);
    free(acts);
    free(zs);
    return 0;
}

/*
 * vo_net_grad: loss = sum over mask of (out - target)^2 / sum(mask), and
 * its gradient with respect to every parameter (grad, same layout).
 * Returns the loss through *loss.
 */
int vo_net_grad(int nb, int h, int w, int periodic, const float *params, const float *input, const float *target,
                const float *mask, float *grad, double *loss) {
    geo_t g = {nb, h, w, periodic};
    const long hw = (long)h * w, blk = (long)nb * NC * hw;
    float *acts = malloc(sizeof(float) * blk * NL), *zs = malloc(sizeof(float) * blk * (NL - 1));
    float *out = malloc(sizeof(float) * nb * hw);
    float *dz = malloc(sizeof(float) * blk);
    double *dprev = malloc(sizeof(double) * blk);
    if (!acts || !zs || !out || !dz || !dprev) {
        free(acts), free(zs), free(out), free(dz), free(dprev);
        return -1;
    }
    forward(&g, params, input, out, acts, zs);
    double msum = 0., l2 = 0.;
    for (long p = 0; p < nb * hw; p++) msum += mask[p];
    if (msum <= 0.) msum = 1.;
    memset(dz, 0, sizeof(float) * blk);
    for (int b = 0; b < nb; b++)
        for (long p = 0; p < hw; p++) {
            double e = (double)out[b * hw + p] - target[b * hw + p];
            l2 += mask[b * hw + p] * e * e;
            dz[((long)b * NC + 0) * hw + p] = bf16r((float)(2.0 * mask[b * hw + p] * e / msum));
        }
    *loss = l2 / msum;
    memset(grad, 0, sizeof(float) * LAYER_PARAMS * NL);
    for (int l = NL; l >= 1; l--) {
        const float *W = params + (long)(l - 1) * LAYER_PARAMS;
        float *gW = grad + (long)(l - 1) * LAYER_PARAMS, *gb = gW + NC * NC * NT;
        const float *X = acts + (l - 1) * blk;
        for (int co = 0; co < cout_of(l); co++) {
            double sb = 0.;
            for (int b = 0; b < nb; b++)
                for (long p = 0; p < hw; p++) sb += dz[((long)b * NC + co) * hw + p];
            gb[co] = (float)sb;
            for (int ci = 0; ci < cin_of(l); ci++)
                for (int t = 0; t < NT; t++) {
                    double s = 0.;
                    for (int b = 0; b < nb; b++)
                        for (int y = 0; y < h; y++)
                            for (int x = 0; x < w; x++) {
                                long q = src_index(&g, y, x, t);
                                if (q < 0) continue;
                                s += (double)dz[((long)b * NC + co) * hw + (long)y * w + x] *
                                     X[((long)b * NC + ci) * hw + q];
                            }
                    gW[(co * NC + ci) * NT + t] = (float)s;
                }
        }
        if (l == 1) break;
        /* gradient into A_{l-1}, then through gelu' into Z_{l-1} */
        memset(dprev, 0, sizeof(double) * blk);
        for (int b = 0; b < nb; b++)
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    const long p = (long)y * w + x;
                    for (int t = 0; t < NT; t++) {
                        long q = src_index(&g, y, x, t);
                        if (q < 0) continue;
                        for (int ci = 0; ci < NC; ci++) {
                            double s = 0.;
                            for (int co = 0; co < cout_of(l); co++)
                                s += (double)bf16r(W[(co * NC + ci) * NT + t]) * dz[((long)b * NC + co) * hw + p];
                            dprev[((long)b * NC + ci) * hw + q] += s;
                        }
                    }
                }
        const float *Z = zs + (l - 2) * blk;
        for (long e = 0; e < blk; e++) dz[e] = bf16r((float)(dprev[e] * gelu_d(Z[e])));
    }
    free(acts), free(zs), free(out), free(dz), free(dprev);
    return 0;
}

int vo_net_param_count(void) { return LAYER_PARAMS * NL; }

float vo_bf16_round(float x) { return bf16r(x); }
