/* FP64 CPU oracle of the warm-start MLP (test infrastructure only).
 *
 * The network has no counterpart in the reference; this file states its mathematics plainly:
 *   h0 = x (zero padded to 64 columns)
 *   h_l = tanh(h_{l-1} W_l + b_l), l = 1..3      (W_l: [in_pad][out_pad] row-major)
 *   y   = h3 W_4 + b_4                            (first n_out columns are the output)
 *   loss = mean over rows and the n_out columns of (y - t)^2
 * and its gradient by the chain rule.  Parameters use the padded layout of the device
 * library: per layer W [in_pad][out_pad] followed by b [out_pad], with (in_pad, out_pad) =
 * (64,256), (256,256), (256,256), (256,32).
 *
 * With emulate_bf16 set, every operand the device's MFMA products read is rounded the way the
 * device rounds it (binary64 -> binary32 -> bfloat16, round to nearest even): the inputs,
 * the weights, the hidden activations, and the loss/tanh derivatives fed to the weight and
 * input gradient products.  Everything else (accumulation, bias, tanh, bias gradients) stays
 * binary64, so the device differs from it only by binary32 accumulation and rare bfloat16
 * rounding-boundary flips. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define L_IN 64
#define L_W 256
#define L_OUT 32

static const int KP[4] = {L_IN, L_W, L_W, L_W};
static const int NP[4] = {L_W, L_W, L_W, L_OUT};

size_t orc_mlp_n_params(void) {
    size_t n = 0;
    for (int l = 0; l < 4; ++l) n += (size_t)KP[l] * NP[l] + NP[l];
    return n;
}

static void offsets(size_t *w, size_t *b) {
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        w[l] = off;
        off += (size_t)KP[l] * NP[l];
        b[l] = off;
        off += NP[l];
    }
}

/* binary64 -> binary32 -> bfloat16 (round to nearest even) -> binary64 */
double orc_bf16_round(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v; /* inf / nan unchanged */
    const uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double rnd(double v, int emu) { return emu ? orc_bf16_round(v) : v; }

/* Activations of one forward pass: H[0] = padded input, H[1..3] = hidden, Y = output [M][32]. */
typedef struct {
    double *H[4];
    double *Y;
} fwd_t;

static void forward(const double *p, int n_in, const double *x, int M, int emu, fwd_t *f) {
    size_t wo[4], bo[4];
    offsets(wo, bo);
    for (int l = 0; l < 4; ++l) f->H[l] = calloc((size_t)M * KP[l], sizeof(double));
    f->Y = calloc((size_t)M * L_OUT, sizeof(double));
    for (int r = 0; r < M; ++r)
        for (int c = 0; c < n_in; ++c) f->H[0][(size_t)r * L_IN + c] = rnd(x[(size_t)r * n_in + c], emu);
    double *Wr = malloc((size_t)L_W * L_W * sizeof(double));
    for (int l = 0; l < 4; ++l) {
        const int K = KP[l], N = NP[l];
        for (size_t e = 0; e < (size_t)K * N; ++e) Wr[e] = rnd(p[wo[l] + e], emu);
        double *out = l < 3 ? f->H[l + 1] : f->Y;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < M; ++r) {
            const double *a = f->H[l] + (size_t)r * K;
            for (int n = 0; n < N; ++n) {
                double s = 0;
                for (int k = 0; k < K; ++k) s += a[k] * Wr[(size_t)k * N + n];
                s += p[bo[l] + n];
                out[(size_t)r * N + n] = l < 3 ? rnd(tanh(s), emu) : s;
            }
        }
    }
    free(Wr);
}

static void fwd_free(fwd_t *f) {
    for (int l = 0; l < 4; ++l) free(f->H[l]);
    free(f->Y);
}

int orc_mlp_forward(const double *params, int n_in, int n_out, const double *x, int M, int emulate_bf16,
                    double *y) {
    if (n_in < 1 || n_in > L_IN || n_out < 1 || n_out > L_OUT || M < 0) return 1;
    fwd_t f;
    forward(params, n_in, x, M, emulate_bf16, &f);
    for (int r = 0; r < M; ++r)
        for (int c = 0; c < n_out; ++c) y[(size_t)r * n_out + c] = f.Y[(size_t)r * L_OUT + c];
    fwd_free(&f);
    return 0;
}

int orc_mlp_loss_grad(const double *params, int n_in, int n_out, const double *x, const double *t, int M,
                      int emulate_bf16, double *loss, double *grad) {
    if (n_in < 1 || n_in > L_IN || n_out < 1 || n_out > L_OUT || M < 1) return 1;
    const int emu = emulate_bf16;
    size_t wo[4], bo[4];
    offsets(wo, bo);
    fwd_t f;
    forward(params, n_in, x, M, emu, &f);
    memset(grad, 0, orc_mlp_n_params() * sizeof(double));
    /* dY (unrounded, for the bias gradient) and its rounded copy for the products */
    double *d = calloc((size_t)M * L_W, sizeof(double));
    double *dr = calloc((size_t)M * L_W, sizeof(double));
    double *dh = calloc((size_t)M * L_W, sizeof(double));
    double se = 0;
    const double scale = 2.0 / ((double)M * n_out);
    for (int r = 0; r < M; ++r)
        for (int c = 0; c < n_out; ++c) {
            const double diff = f.Y[(size_t)r * L_OUT + c] - t[(size_t)r * n_out + c];
            se += diff * diff;
            const double g = emu ? (double)(float)(scale * diff) : scale * diff;
            d[(size_t)r * L_OUT + c] = g;
            dr[(size_t)r * L_OUT + c] = rnd(g, emu);
        }
    *loss = se / ((double)M * n_out);
    double *Wr = malloc((size_t)L_W * L_W * sizeof(double));
    for (int l = 3; l >= 0; --l) {
        const int K = KP[l], N = NP[l];
        /* dW = H_l^T dZ (rounded), db = column sums of dZ */
        double *gw = grad + wo[l], *gb = grad + bo[l];
#pragma omp parallel for schedule(static)
        for (int k = 0; k < K; ++k)
            for (int r = 0; r < M; ++r) {
                const double a = f.H[l][(size_t)r * K + k];
                if (a == 0) continue;
                for (int n = 0; n < N; ++n) gw[(size_t)k * N + n] += a * dr[(size_t)r * N + n];
            }
        for (int r = 0; r < M; ++r)
            for (int n = 0; n < N; ++n) gb[n] += d[(size_t)r * N + n];
        if (l == 0) break;
        /* dH_{l} = dZ W_l^T, then dZ_{l-1} = dH (1 - h^2) */
        for (size_t e = 0; e < (size_t)K * N; ++e) Wr[e] = rnd(params[wo[l] + e], emu);
#pragma omp parallel for schedule(static)
        for (int r = 0; r < M; ++r)
            for (int k = 0; k < K; ++k) {
                double s = 0;
                for (int n = 0; n < N; ++n) s += dr[(size_t)r * N + n] * Wr[(size_t)k * N + n];
                dh[(size_t)r * K + k] = s;
            }
        for (size_t e = 0; e < (size_t)M * K; ++e) {
            const double h = f.H[l][e];
            double g = dh[e] * (1.0 - h * h);
            if (emu) g = (double)(float)g;
            d[e] = g;
            dr[e] = rnd(g, emu);
        }
    }
    free(Wr);
    free(d);
    free(dr);
    free(dh);
    fwd_free(&f);
    return 0;
}
