/*
 * prior_oracle.c -- CPU restatement of the learned edge prior.
 *
 * TEST INFRASTRUCTURE ONLY (see aco_oracle.h).  The prior has no counterpart
 * in the reference: it replaces the eta^beta heuristic of the candidate
 * edges (reference heuristic: src/AntSystem.cu:56-77), so these functions
 * restate the definitions the device code implements, in fp64 (features and
 * Adam in the exact fp32 operation order of the device).
 *
 * Pinning status (tests/test_oracle_prior.py):
 *   orc_prior_features   brute-force per-edge recomputation through Python
 *                        dictionaries and numpy float32 arithmetic
 *   orc_prior_labels     full candidate lists: exactly 2n labels, symmetric
 *   orc_gelu             scipy normal CDF: within the Abramowitz-Stegun bound
 *                        7.5e-8 |x| of x Phi(x); gradient within 7.5e-8 of the
 *                        exact derivative Phi(x) + x phi(x) and of central
 *                        finite differences of orc_gelu
 *   orc_mlp_forward      zero network (= b3); one-path networks against
 *                        scalar closed forms (catches transposed operands)
 *   orc_mlp_grad         central finite differences of orc_mlp_forward's loss
 *   orc_adam_step        first-step closed form lr g / (|g| + eps)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "aco_oracle.h"

#define H 64
#define F 8
#define P_W1 0
#define P_B1 512
#define P_W2 576
#define P_B2 4672
#define P_W3 4736
#define P_B3 4800

/* Edge features (definition in include/acogpu.h); every ratio is one fp32
 * division of integer-valued floats, f4 one fp64 expression rounded once. */
void orc_prior_features(int32_t n, int32_t k, const int32_t *dist, const int32_t *nn, float *feat) {
    uint64_t sum = 0;
    for (int32_t i = 0; i < n; ++i) sum += (uint64_t)dist[(size_t)i * n + nn[(size_t)i * k]];
    for (int32_t i = 0; i < n; ++i) {
        const int32_t di0 = dist[(size_t)i * n + nn[(size_t)i * k]];
        const int32_t dik = dist[(size_t)i * n + nn[(size_t)i * k + k - 1]];
        for (int32_t r = 0; r < k; ++r) {
            const int32_t j = nn[(size_t)i * k + r];
            const int32_t d = dist[(size_t)i * n + j];
            const int32_t dj0 = dist[(size_t)j * n + nn[(size_t)j * k]];
            const int32_t djk = dist[(size_t)j * n + nn[(size_t)j * k + k - 1]];
            int32_t rank = k;
            for (int32_t q = 0; q < k; ++q)
                if (nn[(size_t)j * k + q] == i) {
                    rank = q;
                    break;
                }
            float *f = feat + ((size_t)i * k + r) * F;
            f[0] = (float)d / (float)(dik + 1);
            f[1] = (float)d / (float)(djk + 1);
            f[2] = (float)r / (float)k;
            f[3] = (float)rank / (float)k;
            f[4] = (float)((double)d * (double)n / (double)(sum + (uint64_t)n));
            f[5] = (float)di0 / (float)(dik + 1);
            f[6] = (float)dj0 / (float)(djk + 1);
            f[7] = (float)(d - di0) / (float)(dik - di0 + 1);
        }
    }
}

void orc_prior_labels(int32_t n, int32_t k, const int32_t *nn, const int32_t *tour, float *y) {
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!pos) return;
    for (int32_t t = 0; t < n; ++t) pos[tour[t]] = t;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t t = pos[i];
        const int32_t nx = tour[(t + 1) % n], pv = tour[(t + n - 1) % n];
        for (int32_t r = 0; r < k; ++r) {
            const int32_t j = nn[(size_t)i * k + r];
            y[(size_t)i * k + r] = (nx == j || pv == j) ? 1.0f : 0.0f;
        }
    }
    free(pos);
}

/* Round to bfloat16 (8 significant bits), ties to even, returned as double. */
double orc_bf16(double x) {
    if (x == 0.0 || !isfinite(x)) return x;
    int e;
    const double m = frexp(x, &e); /* x = m 2^e, 0.5 <= |m| < 1 */
    return ldexp(nearbyint(ldexp(m, 8)), e - 8);
}

/* GELU(x) = x Phi(x), Phi from the erf approximation 7.1.26 of Abramowitz
 * and Stegun (Handbook of Mathematical Functions, 1964; |error of erf| <=
 * 1.5e-7, so |error of Phi| <= 7.5e-8): with u = |x| / sqrt(2),
 * erf(u) ~ 1 - P(t) e^{-u^2}, t = 1 / (1 + p u), P(t) = a1 t + ... + a5 t^5,
 * so Phi(x) = 1 - Q for x >= 0 and Q for x < 0, Q = P(t) e^{-x^2/2} / 2.
 * The gradient is GELU'(x) = Phi(x) + x phi(x) with this Phi and the exact
 * normal density phi(x) = e^{-x^2/2} / sqrt(2 pi).  The device evaluates the
 * same expressions in fp32. */
static const double GELU_A[5] = {0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429};
static const double GELU_P = 0.3275911;

/* Phi(x) and e^{-x^2/2} */
static double gelu_phi(double x, double *e) {
    const double a = fabs(x), c = GELU_P / sqrt(2.0);
    *e = exp(-0.5 * x * x);
    const double t = 1.0 / (1.0 + c * a);
    const double P = t * (GELU_A[0] + t * (GELU_A[1] + t * (GELU_A[2] + t * (GELU_A[3] + t * GELU_A[4]))));
    const double q = 0.5 * P * *e;
    return x >= 0 ? 1.0 - q : q;
}

double orc_gelu(double x) {
    double e;
    return x * gelu_phi(x, &e);
}

double orc_gelu_grad(double x) {
    double e;
    const double phi = gelu_phi(x, &e);
    return phi + x * e / sqrt(2.0 * M_PI);
}

/* log(1 + e^x) without overflow */
static double softplus(double x) { return (x > 0 ? x : 0.0) + log1p(exp(-fabs(x))); }

typedef struct {
    double z1[H], h1[H], z2[H], h2[H], out;
} act_t;

static double rnd(double v, int32_t bf) { return bf ? orc_bf16(v) : v; }

static void forward_one(const float *p, const float *x, int32_t bf, act_t *a) {
    double xr[F];
    for (int c = 0; c < F; ++c) xr[c] = rnd((double)x[c], bf);
    for (int o = 0; o < H; ++o) {
        double s = 0.0;
        for (int c = 0; c < F; ++c) s += rnd((double)p[P_W1 + o * F + c], bf) * xr[c];
        a->z1[o] = s + (double)p[P_B1 + o];
        a->h1[o] = rnd(orc_gelu(a->z1[o]), bf);
    }
    double out = 0.0;
    for (int o = 0; o < H; ++o) {
        double s = 0.0;
        for (int i = 0; i < H; ++i) s += rnd((double)p[P_W2 + o * H + i], bf) * a->h1[i];
        a->z2[o] = s + (double)p[P_B2 + o];
        a->h2[o] = rnd(orc_gelu(a->z2[o]), bf);
        out += (double)p[P_W3 + o] * a->h2[o];
    }
    a->out = out + (double)p[P_B3];
}

/* bf16 != 0 rounds, as the device does, the inputs, W1, W2 and the two
 * hidden activations to bfloat16; bf16 == 0 is exact fp64. */
void orc_mlp_forward(const float *params, const float *x, int64_t count, int32_t bf16, double *out) {
    act_t a;
    for (int64_t e = 0; e < count; ++e) {
        forward_one(params, x + e * F, bf16, &a);
        out[e] = a.out;
    }
}

/* Mean weighted binary cross-entropy with logits over `count` edges and its
 * gradient (ACO_PRIOR_PARAMS doubles, overwritten); returns the loss.  With
 * bf16 != 0 the back-propagated pre-activation gradients dz2 and dz1 are
 * rounded to bfloat16 where they enter the weight-gradient and dh1 products,
 * as the device does; b2's gradient takes dz2 unrounded, b1's comes out of
 * the W1 product (a constant-one input) and takes dz1 rounded. */
double orc_mlp_grad(const float *params, const float *x, const float *y, int64_t count, float pos_weight,
                    int32_t bf16, double *grad) {
    memset(grad, 0, sizeof(double) * 4801);
    const double pw = (double)pos_weight;
    double loss = 0.0;
    act_t a;
    double dz2[H], dz2r[H], dz1[H];
    for (int64_t e = 0; e < count; ++e) {
        const float *xe = x + e * F;
        forward_one(params, xe, bf16, &a);
        const double t = (double)y[e];
        const double s = 1.0 / (1.0 + exp(-a.out));
        loss += (pw * t * softplus(-a.out) + (1.0 - t) * softplus(a.out)) / (double)count;
        const double g = (s * (1.0 - t + pw * t) - pw * t) / (double)count;
        grad[P_B3] += g;
        for (int o = 0; o < H; ++o) {
            grad[P_W3 + o] += g * a.h2[o];
            dz2[o] = g * (double)params[P_W3 + o] * orc_gelu_grad(a.z2[o]);
            grad[P_B2 + o] += dz2[o];
            dz2r[o] = rnd(dz2[o], bf16);
        }
        for (int o = 0; o < H; ++o)
            for (int i = 0; i < H; ++i) grad[P_W2 + o * H + i] += dz2r[o] * a.h1[i];
        for (int i = 0; i < H; ++i) {
            double dh = 0.0;
            for (int o = 0; o < H; ++o) dh += rnd((double)params[P_W2 + o * H + i], bf16) * dz2r[o];
            dz1[i] = dh * orc_gelu_grad(a.z1[i]);
            grad[P_B1 + i] += rnd(dz1[i], bf16);
        }
        for (int o = 0; o < H; ++o) {
            const double d = rnd(dz1[o], bf16);
            for (int c = 0; c < F; ++c) grad[P_W1 + o * F + c] += d * rnd((double)xe[c], bf16);
        }
    }
    return loss;
}

/* One Adam step in fp32, in the device's operation order; step is the
 * 1-based step number (bias corrections computed in fp64, rounded once). */
void orc_adam_step(int32_t nparams, float *params, const float *grad, float *m, float *v, int32_t step, float lr,
                   float beta1, float beta2, float eps) {
    const float bc1 = (float)(1.0 - pow((double)beta1, (double)step));
    const float bc2 = (float)(1.0 - pow((double)beta2, (double)step));
    for (int32_t i = 0; i < nparams; ++i) {
        const float g = grad[i];
        const float mi = beta1 * m[i] + (1.0f - beta1) * g;
        const float vi = beta2 * v[i] + (1.0f - beta2) * (g * g);
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / bc1;
        const float vh = vi / bc2;
        params[i] = params[i] - lr * mh / (sqrtf(vh) + eps);
    }
}
