/*
 * aco_oracle.c -- CPU restatement of the reference Ant-System hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never links it.
 * Built with -O3 -ffp-contract=off and without fast-math so that every fp32
 * operation is one IEEE round-to-nearest operation in program order, the same
 * arithmetic the gfx950 kernels perform.  See aco_oracle.h for what pins each
 * function.
 *
 * Citations are to the reference tree of observit17/acogpu.
 */
#include "aco_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Counter-based RNG.  The reference draws from Numerical-Recipes ran2 on the
 * CPU (src/Utils.cu:19-60, 112-114) and cuRAND XORWOW on the GPU
 * (inc/Colony.cuh:58-72); neither is reproducible (clock seeded) and cuRAND is
 * closed.  Both are replaced by Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11), restated here from
 * the paper: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* One 32-bit word per (seed, iteration, ant, step, stream): key {seed lo,
 * seed hi}, counter {block, ant, iteration, stream} with
 * block = 64 * (step / 256) + step % 64 and output word (step / 64) % 4.
 * (A wave computes 64 blocks at once, lane l holding block 64*(s/256) + l,
 * and reads one lane per step.)  Step 0 is the start city
 * (src/Colony.cu:513-524), steps 1..n-1 the roulette draws
 * (src/Colony.cu:826). */
uint32_t orc_rng_word(uint64_t seed, uint32_t iter, uint32_t ant, uint32_t step, uint32_t stream) {
    uint32_t ctr[4] = {((step >> 8) << 6) | (step & 63u), ant, iter, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[(step >> 6) & 3u];
}

/* u in [0,1) with 24 random bits: exact in fp32. */
float orc_uniform(uint32_t word) { return (float)(word >> 8) * 5.9604644775390625e-08f; }

/* Start city floor(word * n / 2^32): integer, always in [0, n).  Replaces
 * (long)(u * n) of src/Colony.cu:518 and inc/Colony.cuh:156, which can reach
 * n on the GPU because curand_uniform is in (0,1]. */
int32_t orc_start_city(uint32_t word, int32_t n) {
    return (int32_t)(((uint64_t)word * (uint64_t)(uint32_t)n) >> 32);
}

/* ------------------------------------------------------------------------ */
/* EUC_2D.  Reference: Utils::round_distance src/Utils.cu:212-218,
 * (int)(sqrt(xd*xd + yd*yd) + 0.5) with xd = xi - xj computed in fp32. */
int32_t orc_euc2d_ref(float xi, float xj, float yi, float yj) {
    double xd = xi - xj;
    double yd = yi - yj;
    double r = sqrt(xd * xd + yd * yd) + 0.5;
    return (int32_t)r;
}

/* Exact nearest integer of sqrt(dx^2 + dy^2) for integer coordinates:
 * the r with (2r-1)^2 <= 4D < (2r+1)^2.  Equal to orc_euc2d_ref for
 * |dx|,|dy| < 2^24 (no half-way case exists for integer D). */
int32_t orc_euc2d_exact(int64_t dx, int64_t dy) {
    int64_t D = dx * dx + dy * dy;
    int64_t r = (int64_t)(sqrt((double)D) + 0.5);
    while (r > 0 && (2 * r - 1) * (2 * r - 1) > 4 * D) r--;
    while ((2 * r + 1) * (2 * r + 1) <= 4 * D) r++;
    return (int32_t)r;
}

/* Distance matrix with MAX_DISTANCE on the diagonal
 * (Utils::compute_distance src/Utils.cu:162-182). xy is interleaved x,y. */
void orc_dist_from_int_coords(const int32_t *xy, int32_t n, int32_t *dist) {
    for (int32_t i = 0; i < n; i++) {
        dist[(int64_t)i * n + i] = ORC_MAX_DISTANCE;
        for (int32_t j = i + 1; j < n; j++) {
            int32_t d = orc_euc2d_exact((int64_t)xy[2 * i] - xy[2 * j],
                                        (int64_t)xy[2 * i + 1] - xy[2 * j + 1]);
            dist[(int64_t)i * n + j] = d;
            dist[(int64_t)j * n + i] = d;
        }
    }
}

void orc_dist_from_float_coords(const float *x, const float *y, int32_t n, int32_t *dist) {
    for (int32_t i = 0; i < n; i++) {
        dist[(int64_t)i * n + i] = ORC_MAX_DISTANCE;
        for (int32_t j = i + 1; j < n; j++) {
            int32_t d = orc_euc2d_ref(x[i], x[j], y[i], y[j]);
            dist[(int64_t)i * n + j] = d;
            dist[(int64_t)j * n + i] = d;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Nearest-neighbour order.  The reference sorts every full row
 * (Colony::computeNNlist src/Colony.cu:276-311 on the CPU, quicksort;
 * dev_compute_nnlist inc/Colony.cuh:101-123 on the GPU, rank-by-count, i.e.
 * stable (d, j) order).  The oracle uses the stable (d, j) order.  For
 * k < n it returns the k nearest cities other than i; for k == n the full
 * row, in which i itself (d = MAX_DISTANCE) sorts last. */
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

void orc_nn_lists(const int32_t *dist, int32_t n, int32_t k, int32_t *nn) {
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        int32_t cnt = 0;
        for (int32_t j = 0; j < n; j++) {
            if (k < n && j == i) continue;
            keys[cnt++] = ((uint64_t)(uint32_t)dist[(int64_t)i * n + j] << 32) | (uint32_t)j;
        }
        qsort(keys, (size_t)cnt, sizeof(uint64_t), cmp_u64);
        for (int32_t r = 0; r < k; r++) nn[(int64_t)i * k + r] = (int32_t)(keys[r] & 0xffffffffu);
    }
    free(keys);
}

/* ------------------------------------------------------------------------ */
/* Heuristic eta^beta.  Reference HURESTIC(d, b) = powf(1.0/(float)d, b),
 * 0 for d == MAX_DISTANCE (inc/Colony.h:17-21, src/Colony.cu:232-246).  Here
 * it is tabulated by integer distance, in fp64 and rounded to fp32 once, so
 * CPU and GPU read identical bits.  For beta = 1, 2 or 3 with dmax^beta <
 * 2^53 the value is 1 / d^beta (d^beta exact, one correctly rounded fp64
 * division), which the GPU sweep also evaluates on the fly; otherwise
 * pow(1 / d, beta).  d == 0 (duplicate cities, release/input/a280.tsp:178)
 * would give +inf; it is defined as d = 1/2. */
int32_t orc_eta_table(double beta, int32_t dmax, float *table) {
    const int32_t b = (int32_t)beta;
    double top = 1.0;
    for (int32_t q = 0; q < b && q < 3; q++) top *= (double)dmax;
    if ((double)b == beta && b >= 1 && b <= 3 && top < 9007199254740992.0) {
        table[0] = (float)(1 << b);
        for (int32_t d = 1; d <= dmax; d++) {
            const double dd = (double)d;
            const double p = b == 1 ? dd : (b == 2 ? dd * dd : dd * dd * dd);
            table[d] = (float)(1.0 / p);
        }
    } else {
        table[0] = (float)pow(2.0, beta);
        for (int32_t d = 1; d <= dmax; d++) table[d] = (float)pow(1.0 / (double)d, beta);
    }
    return dmax + 1;
}

float orc_eta(const float *table, int32_t table_len, int32_t d) {
    if (d < 0 || d >= table_len) return 0.0f; /* includes MAX_DISTANCE */
    return table[d];
}

/* tau^alpha for alpha != 1 (the reference's pow, src/Colony.cu:264), here
 * defined as 2^(alpha log2 tau) in fp64 by a fixed sequence of IEEE
 * operations so that the GPU's pow_alpha (acogpu_amd/csrc/aco_kernels.hip)
 * gives the same bits: log2 = exponent + 2 atanh((m-1)/(m+1)) / ln 2 with the
 * mantissa m in [sqrt(1/2), sqrt(2)) (odd series to t^23), 2^z = 2^floor(z)
 * e^(frac(z) ln 2) (Taylor series to the 18th power). */
float orc_pow_alpha(float p, float alpha) {
    if (!(p > 0.0f)) return p == 0.0f ? 0.0f : NAN;
    union { double d; uint64_t u; } v = {(double)p}, w;
    int32_t e = (int32_t)((v.u >> 52) & 0x7ff) - 1023;
    w.u = (v.u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m = w.d;
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double t = (m - 1.0) / (m + 1.0);
    const double t2 = t * t;
    static const double odd[12] = {1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0,
                                   1.0 / 11.0, 1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0,  1.0};
    double s = odd[0];
    for (int i = 1; i < 12; i++) s = s * t2 + odd[i];
    const double lg = (double)e + ((2.0 * t) * s) * 1.4426950408889634;
    const double z = (double)alpha * lg;
    const double k = floor(z);
    if (k < -1022.0) return 0.0f;
    if (k > 1023.0) return INFINITY;
    const double g = (z - k) * 0.6931471805599453;
    static const double inv_fact[19] = {
        1.0 / 6402373705728000.0, 1.0 / 355687428096000.0, 1.0 / 20922789888000.0, 1.0 / 1307674368000.0,
        1.0 / 87178291200.0,      1.0 / 6227020800.0,      1.0 / 479001600.0,      1.0 / 39916800.0,
        1.0 / 3628800.0,          1.0 / 362880.0,          1.0 / 40320.0,          1.0 / 5040.0,
        1.0 / 720.0,              1.0 / 120.0,             1.0 / 24.0,             1.0 / 6.0,
        0.5,                      1.0,                     1.0};
    double q = inv_fact[0];
    for (int i = 1; i < 19; i++) q = q * g + inv_fact[i];
    union { double d; uint64_t u; } sc;
    sc.u = (uint64_t)((int64_t)k + 1023) << 52;
    return (float)(q * sc.d);
}

static inline float apply_alpha(float p, float alpha) {
    /* tau^alpha; the reference drops alpha in construction
     * (inc/Colony.cuh:29-30), which equals this for alpha == 1. */
    return alpha == 1.0f ? p : orc_pow_alpha(p, alpha);
}

/* ------------------------------------------------------------------------ */
/* Wave-64 inclusive prefix sum in the exact fp32 order of the gfx950 DPP
 * sequence row_shr:1,2,4,8 then row_bcast:15 (rows 1,3) and row_bcast:31
 * (rows 2,3).  This fixes the summation order that the reference leaves to a
 * buggy integer reduction (src/Colony.cu:643-685). */
void orc_wave_scan(float v[64]) {
    float t[64];
    static const int shifts[4] = {1, 2, 4, 8};
    for (int s = 0; s < 4; s++) {
        int sh = shifts[s];
        for (int i = 0; i < 64; i++) t[i] = ((i & 15) >= sh) ? v[i - sh] : 0.0f;
        for (int i = 0; i < 64; i++) v[i] = v[i] + t[i];
    }
    for (int i = 0; i < 64; i++) {
        int row = i >> 4;
        t[i] = (row == 1 || row == 3) ? v[16 * row - 1] : 0.0f;
    }
    for (int i = 0; i < 64; i++) v[i] = v[i] + t[i];
    for (int i = 0; i < 64; i++) t[i] = (i >= 32) ? v[31] : 0.0f;
    for (int i = 0; i < 64; i++) v[i] = v[i] + t[i];
}

/* Roulette over one wave of candidate weights (selectCities
 * src/Colony.cu:821-851): r = u * total; the first lane whose inclusive prefix
 * exceeds r and whose own weight is positive; if rounding leaves no such lane,
 * the last positive lane.  Returns -1 when total is not positive (the
 * reference then falls back, src/Colony.cu:776-778). */
int32_t orc_select_candidate(const float w[64], float u, float *total_out) {
    float p[64];
    memcpy(p, w, sizeof(p));
    orc_wave_scan(p);
    float total = p[63];
    if (total_out) *total_out = total;
    if (!(total > 0.0f)) return -1;
    float rr = u * total;
    for (int r = 0; r < 64; r++)
        if (p[r] > rr && w[r] > 0.0f) return r;
    for (int r = 63; r >= 0; r--)
        if (w[r] > 0.0f) return r;
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Fused row sweep (one iteration's dense work):
 *   choice   = tau^alpha * eta(d)        computeProbabilities src/Colony.cu:252-273,
 *                                        dev_compute_selection inc/Colony.cuh:91-99
 *   tau_evap = max((1-rho) tau, 1e-18)   AntSystem::evaporatePheromones
 *                                        src/AntSystem.cu:9-11,206-225
 *   mean     = row mean of tau           computePheromoneMean src/Colony.cu:427-443
 *   choice_nn[i][r] = choice[i][nn[i][r]]
 * The row mean uses the wave order of the GPU sweep: lane l accumulates
 * columns 256s+4l+q (q = 0..3, s ascending) sequentially, then the 64 lane
 * partials are combined by orc_wave_scan. */
void orc_sweep(int32_t n, int32_t k, const float *tau, const int32_t *dist,
               const float *eta, int32_t eta_len, const int32_t *nn, const float *heur_nn,
               float alpha, float rho, int32_t want_mean,
               float *choice, float *choice_nn, float *tau_evap, float *mean) {
    const float one_minus_rho = 1.0f - rho;
    for (int32_t i = 0; i < n; i++) {
        const float *trow = tau + (int64_t)i * n;
        const int32_t *drow = dist + (int64_t)i * n;
        for (int32_t j = 0; j < n; j++) {
            float t = trow[j];
            float e = orc_eta(eta, eta_len, drow[j]);
            if (choice) choice[(int64_t)i * n + j] = apply_alpha(t, alpha) * e;
            if (tau_evap) {
                float p = one_minus_rho * t;
                tau_evap[(int64_t)i * n + j] = ((double)p > 1e-18) ? p : (float)1e-18;
            }
        }
        if (want_mean && mean) {
            float acc[64];
            for (int l = 0; l < 64; l++) {
                float a = 0.0f;
                for (int32_t s = 0; 256 * s < n; s++)
                    for (int q = 0; q < 4; q++) {
                        int32_t j = 256 * s + 4 * l + q;
                        if (j < n) a = a + trow[j];
                    }
                acc[l] = a;
            }
            orc_wave_scan(acc);
            mean[i] = acc[63] / (float)n;
        }
        if (choice_nn && choice && nn)
            for (int32_t r = 0; r < k; r++) {
                int64_t e = (int64_t)i * k + r;
                choice_nn[e] = heur_nn ? apply_alpha(trow[nn[e]], alpha) * heur_nn[e]
                                       : choice[(int64_t)i * n + nn[e]];
            }
    }
}

/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t n;
    const int32_t *dist;
    const float *choice, *tau, *mean, *eta;
    int32_t eta_len;
    float alpha;
    int young;
    float ratio;
} weight_ctx;

/* Selection weight of edge (c, j) for one ant: compute_fittness
 * inc/Colony.cuh:14-31.  Aging: for ants younger than min_age after the
 * colony has converged, tau is pulled toward the row mean,
 * ((tau - mean) * (age/min_age)) + mean. */
static inline float edge_weight(const weight_ctx *w, int32_t c, int32_t j) {
    int64_t xy = (int64_t)c * w->n + j;
    if (!w->young) return w->choice[xy];
    float m = w->mean[c];
    float p = ((w->tau[xy] - m) * w->ratio) + m;
    return apply_alpha(p, w->alpha) * orc_eta(w->eta, w->eta_len, w->dist[xy]);
}

/* select_local_best inc/Colony.cuh:33-43: the last unvisited city in the
 * row's NN order (the farthest). */
static int32_t farthest_unvisited(const int32_t *nnrow, const uint8_t *vis, int32_t n) {
    for (int32_t r = n - 1; r >= 0; r--)
        if (!vis[nnrow[r]]) return nnrow[r];
    return -1;
}

void orc_construct(int32_t mode, int32_t n, int32_t m, int32_t k,
                   const int32_t *dist, const int32_t *nn, const int32_t *nnfull,
                   const float *choice, const float *choice_nn, const float *heur_nn,
                   const float *tau, const float *mean,
                   const float *eta, int32_t eta_len, float alpha,
                   const int32_t *age, const uint8_t *valid, int32_t converged,
                   int32_t min_age, uint64_t seed, uint32_t iter, uint32_t stream,
                   int32_t *tours, int64_t *lengths, int64_t *fallback_out) {
    uint8_t *vis = (uint8_t *)malloc((size_t)n);
    float *wrow = (float *)malloc(sizeof(float) * (size_t)(n + 64));
    float *prow = (float *)malloc(sizeof(float) * (size_t)(n + 64));
    float *tprev = (float *)malloc(sizeof(float) * (size_t)(n / 64 + 2));
    int64_t nfallback = 0, nscanned = 0;
    weight_ctx wc = {n, dist, choice, tau, mean, eta, eta_len, alpha, 0, 0.0f};

    for (int32_t a = 0; a < m; a++) {
        if (valid && !valid[a]) continue;
        wc.young = (age && converged && age[a] < min_age);
        wc.ratio = wc.young ? (float)age[a] / (float)min_age : 0.0f;
        int32_t *tour = tours + (int64_t)a * n;
        memset(vis, 0, (size_t)n);
        int32_t start = orc_start_city(orc_rng_word(seed, iter, (uint32_t)a, 0, stream), n);
        tour[0] = start;
        vis[start] = 1;
        int64_t len = 0;
        int32_t c = start;
        for (int32_t s = 1; s < n; s++) {
            float u = orc_uniform(orc_rng_word(seed, iter, (uint32_t)a, (uint32_t)s, stream));
            int32_t next = -1;
            if (mode == ORC_MODE_CANDIDATE) {
                float w[64];
                for (int r = 0; r < 64; r++) {
                    w[r] = 0.0f;
                    if (r < k) {
                        int32_t j = nn[(int64_t)c * k + r];
                        if (!vis[j]) {
                            if (!wc.young) {
                                w[r] = choice_nn[(int64_t)c * k + r];
                            } else if (heur_nn) {  /* young weight with the candidate-edge heuristic */
                                int64_t xy = (int64_t)c * n + j;
                                float mc = mean[c];
                                float p = ((tau[xy] - mc) * wc.ratio) + mc;
                                w[r] = apply_alpha(p, alpha) * heur_nn[(int64_t)c * k + r];
                            } else {
                                w[r] = edge_weight(&wc, c, j);
                            }
                        }
                    }
                }
                int32_t sel = orc_select_candidate(w, u, NULL);
                if (sel >= 0) {
                    next = nn[(int64_t)c * k + sel];
                } else {
                    /* choose_best_next: argmax of the weight over every
                     * unvisited city, ties to the lowest index. */
                    float best = -1.0f;
                    int32_t bj = -1;
                    nfallback++;
                    for (int32_t j = 0; j < n; j++) {
                        if (vis[j]) continue;
                        nscanned++;
                        float x = edge_weight(&wc, c, j);
                        if (x > best) {
                            best = x;
                            bj = j;
                        }
                    }
                    if (bj < 0)
                        for (int32_t j = 0; j < n && bj < 0; j++)
                            if (!vis[j]) bj = j;
                    next = bj;
                }
            } else if (mode == ORC_MODE_FULLROW) {
                const int32_t *row = nnfull + (int64_t)c * n;
                int32_t nchunk = (n + 63) / 64;
                float T = 0.0f;
                for (int32_t q = 0; q < nchunk; q++) {
                    float w[64];
                    for (int l = 0; l < 64; l++) {
                        int32_t r = 64 * q + l;
                        w[l] = 0.0f;
                        if (r < n && !vis[row[r]]) w[l] = edge_weight(&wc, c, row[r]);
                        wrow[64 * q + l] = w[l];
                    }
                    orc_wave_scan(w);
                    memcpy(prow + 64 * q, w, sizeof(w));
                    tprev[q] = T;
                    T = T + w[63];
                }
                if (T > 0.0f) {
                    float rr = u * T;
                    for (int32_t q = 0; q < nchunk && next < 0; q++)
                        for (int l = 0; l < 64; l++) {
                            int32_t r = 64 * q + l;
                            float P = tprev[q] + prow[r];
                            if (P > rr && wrow[r] > 0.0f) {
                                next = row[r];
                                break;
                            }
                        }
                }
                if (next < 0) {
                    nfallback++;
                    next = farthest_unvisited(row, vis, n);
                }
            } else { /* ORC_MODE_REFSEQ: calculateProbabilityOfSelection +
                        selectCities / chooseBestNext, src/Colony.cu:795-864 */
                const int32_t *row = nnfull + (int64_t)c * n;
                float sum = 0.0f;
                for (int32_t i = 0; i < n; i++) {
                    if (vis[row[i]]) {
                        wrow[i] = 0.0f;
                    } else {
                        int64_t xy = (int64_t)c * n + row[i];
                        float x;
                        if (wc.young) {
                            x = edge_weight(&wc, c, row[i]);
                        } else {
                            x = apply_alpha(tau[xy], alpha) * orc_eta(eta, eta_len, dist[xy]);
                        }
                        wrow[i] = x;
                        sum += x;
                    }
                }
                if (sum <= 0.0f) {
                    nfallback++;
                    next = farthest_unvisited(row, vis, n);
                } else {
                    float rnd = u * sum;
                    int32_t i = 0;
                    float partial = wrow[0];
                    while (partial <= rnd) {
                        i++;
                        if (i >= n) {
                            nfallback++;
                            i = -1;
                            break;
                        }
                        partial += wrow[i];
                    }
                    next = (i < 0) ? farthest_unvisited(row, vis, n) : row[i];
                }
            }
            len += dist[(int64_t)c * n + next];
            vis[next] = 1;
            tour[s] = next;
            c = next;
        }
        len += dist[(int64_t)c * n + start];
        lengths[a] = len;
    }
    if (fallback_out) {
        fallback_out[0] = nfallback;
        fallback_out[1] = nscanned;
    }
    free(vis);
    free(wrow);
    free(prow);
    free(tprev);
}

/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t len;
    int32_t ant;
} rank_t;

static int cmp_rank(const void *a, const void *b) {
    const rank_t *x = (const rank_t *)a, *y = (const rank_t *)b;
    if (x->len != y->len) return x->len < y->len ? -1 : 1;
    return x->ant - y->ant;
}

/* computeAntDistances src/Colony.cu:996-1043 (CPU branch): per-ant length
 * bookkeeping (aging antperformance), iteration best = first strict minimum,
 * best-so-far update and copy.  Then the deposit list:
 *   AS   every valid ant with weight 1 (src/AntSystem.cu:296-316);
 *   RBAS the w-1 best ranked ants (ties to the lower index) with weights
 *        w-r, plus the best-so-far tour with weight w -- textbook AS_rank
 *        (Bullnheimer, Hartl, Strauss 1999), since RankBasedAntSystem's
 *        updatePheromones is empty in the reference
 *        (src/RankBasedAntSystem.cu:48-56). */
int32_t orc_stats(int32_t variant, int32_t n, int32_t m, int32_t w,
                  const int32_t *tours, const int64_t *lengths, const uint8_t *valid,
                  int32_t *perf, int64_t *prev_len, int64_t *state64,
                  int32_t *best_tour, int32_t *dep_ant, int32_t *dep_w) {
    int64_t iter_best = ORC_MAX_DISTANCE - 1;
    int32_t ib_ant = -1;
    for (int32_t a = 0; a < m; a++) {
        if (valid && !valid[a]) continue;
        int64_t L = lengths[a];
        if (perf && prev_len) {
            perf[a] = (L == prev_len[a]) ? perf[a] + 1 : 0;
            prev_len[a] = L;
        }
        if (L < iter_best) {
            iter_best = L;
            ib_ant = a;
        }
    }
    state64[0] = iter_best;
    state64[2] = ib_ant;
    if (ib_ant > -1 && iter_best < state64[1]) {
        state64[1] = iter_best;
        memcpy(best_tour, tours + (int64_t)ib_ant * n, sizeof(int32_t) * (size_t)n);
        state64[3] = 0;
    } else {
        state64[3] += 1;
    }
    int32_t nd = 0;
    if (variant == ORC_AS) {
        for (int32_t a = 0; a < m; a++) {
            if (valid && !valid[a]) continue;
            dep_ant[nd] = a;
            dep_w[nd] = 1;
            nd++;
        }
    } else {
        rank_t *rk = (rank_t *)malloc(sizeof(rank_t) * (size_t)(m > 0 ? m : 1));
        int32_t nv = 0;
        for (int32_t a = 0; a < m; a++) {
            if (valid && !valid[a]) continue;
            rk[nv].len = lengths[a];
            rk[nv].ant = a;
            nv++;
        }
        qsort(rk, (size_t)nv, sizeof(rank_t), cmp_rank);
        for (int32_t r = 1; r <= w - 1 && r <= nv; r++) {
            dep_ant[nd] = rk[r - 1].ant;
            dep_w[nd] = w - r;
            nd++;
        }
        free(rk);
        if (state64[1] < ORC_MAX_DISTANCE) {
            dep_ant[nd] = -1;
            dep_w[nd] = w;
            nd++;
        }
    }
    return nd;
}

/* FP64 pheromone reference: evaporation (src/AntSystem.cu:218-223, with the
 * fp32 factor 1 - rho of the reference) and symmetric deposit of
 * weight / L on both (i,j) and (j,i) (src/AntSystem.cu:296-316). */
void orc_evaporate_f64(int32_t n, double *tau, float rho) {
    double f = (double)(1.0f - rho);
    for (int64_t x = 0; x < (int64_t)n * n; x++) {
        double p = f * tau[x];
        tau[x] = p > 1e-18 ? p : 1e-18;
    }
}

void orc_deposit_f64(int32_t n, double *tau, const int32_t *tours, const int64_t *lengths,
                     const int32_t *best_tour, int64_t best_len,
                     const int32_t *dep_ant, const int32_t *dep_w, int32_t ndep) {
    for (int32_t e = 0; e < ndep; e++) {
        const int32_t *t = dep_ant[e] >= 0 ? tours + (int64_t)dep_ant[e] * n : best_tour;
        int64_t L = dep_ant[e] >= 0 ? lengths[dep_ant[e]] : best_len;
        double d = (double)dep_w[e] / (double)L;
        for (int32_t i = 0; i < n; i++) {
            int32_t a = t[i], b = t[(i + 1) % n];
            tau[(int64_t)a * n + b] += d;
            tau[(int64_t)b * n + a] += d;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Aging, CPU semantics (the GPU kernels of the reference differ, Q15 of the
 * analysis; the CPU branch is canonical here).
 *   updateAntAges  src/Colony.cu:353-364: every ant ages by one.
 *   retireAnts     src/Colony.cu:378-393: the first valid ant among the first
 *                  m/2 with age >= max_age retires; sets converged.
 *   recruiteAnts   src/Colony.cu:399-423: if the colony has not improved for
 *                  global_perf iterations or fewer than min_pop ants are valid,
 *                  the first invalid ant is recruited with age 0. */
void orc_aging_start(int32_t m, int32_t *age) {
    for (int32_t a = 0; a < m; a++) age[a] = age[a] + 1;
}

void orc_aging_control(int32_t m, int32_t *age, uint8_t *valid, int32_t *num_valid,
                       int32_t *converged, int32_t gperf, int32_t max_age,
                       int32_t global_perf, int32_t min_pop) {
    for (int32_t i = 0; i < m / 2; i++) {
        if (valid[i] && age[i] >= max_age) {
            valid[i] = 0;
            (*num_valid)--;
            *converged = 1;
            break;
        }
    }
    if (gperf >= global_perf || *num_valid < min_pop) {
        for (int32_t i = 0; i < m; i++) {
            if (!valid[i]) {
                valid[i] = 1;
                age[i] = 0;
                (*num_valid)++;
                break;
            }
        }
    }
}

/* Nearest-neighbour tour length from `start` (ties to the lowest index):
 * C_nn of the initial-pheromone formulas that the reference leaves as a stub
 * (Colony::greedyDistance src/Colony.cu:1053-1055, comment inc/Colony.h:51). */
int32_t orc_greedy_length(const int32_t *dist, int32_t n, int32_t start) {
    uint8_t *vis = (uint8_t *)calloc((size_t)n, 1);
    int64_t len = 0;
    int32_t c = start;
    vis[c] = 1;
    for (int32_t s = 1; s < n; s++) {
        int32_t bj = -1;
        int32_t bd = 0;
        for (int32_t j = 0; j < n; j++) {
            if (vis[j]) continue;
            int32_t d = dist[(int64_t)c * n + j];
            if (bj < 0 || d < bd) {
                bd = d;
                bj = j;
            }
        }
        len += bd;
        vis[bj] = 1;
        c = bj;
    }
    len += dist[(int64_t)c * n + start];
    free(vis);
    return (int32_t)(len > 0x7fffffff ? 0x7fffffff : len);
}

/* ------------------------------------------------------------------------ */
static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Whole run in fp32 with the reference's CPU operation order
 * (Colony::constructSolution src/Colony.cu:537-594, TSP::execute
 * src/TSP.cpp:60-71): [mean] -> construct -> lengths/best -> evaporate ->
 * deposit (ant by ant, tau[j][h] += d; tau[h][j] = tau[j][h]) -> aging.
 * In the no-aging build the reference also recomputes the full probability
 * matrix every iteration (src/Colony.cu:557-562); that sweep is kept. */
double orc_run(int32_t n, int32_t m, const int32_t *dist, const int32_t *params,
               const float *fparams, uint64_t seed, int64_t *out, int32_t *best_tour) {
    int32_t variant = params[0], mode = params[1], k = params[2], w = params[3];
    int32_t aging = params[4], min_age = params[5], max_age = params[6];
    int32_t global_perf = params[7], min_pop = params[8], iters = params[9];
    uint32_t stream = (uint32_t)params[10];
    float alpha = fparams[0], beta = fparams[1], rho = fparams[2], tau0 = fparams[3];
    int64_t nn2 = (int64_t)n * n;

    int32_t dmax = 0;
    for (int64_t x = 0; x < nn2; x++)
        if (dist[x] != ORC_MAX_DISTANCE && dist[x] > dmax) dmax = dist[x];
    float *eta = (float *)malloc(sizeof(float) * (size_t)(dmax + 1));
    float *tau = (float *)malloc(sizeof(float) * (size_t)nn2);
    float *choice = (float *)malloc(sizeof(float) * (size_t)nn2);
    float *mean = (float *)malloc(sizeof(float) * (size_t)n);
    int32_t kk = (mode == ORC_MODE_CANDIDATE) ? k : n;
    int32_t *nn = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * (size_t)kk);
    float *choice_nn = (float *)malloc(sizeof(float) * (size_t)n * (size_t)(k > 0 ? k : 1));
    int32_t *tours = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * (size_t)n);
    int64_t *lengths = (int64_t *)calloc((size_t)m, sizeof(int64_t));
    int64_t *prev_len = (int64_t *)calloc((size_t)m, sizeof(int64_t));
    int32_t *perf = (int32_t *)calloc((size_t)m, sizeof(int32_t));
    int32_t *age = (int32_t *)calloc((size_t)m, sizeof(int32_t));
    uint8_t *valid = (uint8_t *)malloc((size_t)m);
    int32_t *dep_ant = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + w + 1));
    int32_t *dep_w = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + w + 1));
    if (!eta || !tau || !choice || !mean || !nn || !choice_nn || !tours || !lengths ||
        !prev_len || !perf || !age || !valid || !dep_ant || !dep_w)
        return -1.0;

    int32_t eta_len = orc_eta_table((double)beta, dmax, eta);
    orc_nn_lists(dist, n, kk, nn);
    for (int64_t x = 0; x < nn2; x++) tau[x] = tau0;
    memset(valid, 1, (size_t)m);
    int32_t num_valid = m, converged = 0;
    int64_t st[4] = {ORC_MAX_DISTANCE - 1, ORC_MAX_DISTANCE, -1, 0};

    double t0 = now_seconds();
    for (int32_t it = 0; it < iters; it++) {
        if (aging) {
            orc_sweep(n, 0, tau, dist, eta, eta_len, NULL, NULL, alpha, rho, 1, NULL, NULL, NULL, mean);
            orc_aging_start(m, age);
        }
        if (mode == ORC_MODE_REFSEQ) {
            if (!aging)
                orc_sweep(n, 0, tau, dist, eta, eta_len, NULL, NULL, alpha, rho, 0, choice, NULL, NULL, NULL);
            orc_construct(mode, n, m, k, dist, NULL, nn, choice, NULL, NULL, tau, mean, eta, eta_len,
                          alpha, aging ? age : NULL, aging ? valid : NULL, converged, min_age,
                          seed, (uint32_t)it, stream, tours, lengths, NULL);
        } else {
            orc_sweep(n, k, tau, dist, eta, eta_len, mode == ORC_MODE_CANDIDATE ? nn : NULL, NULL,
                      alpha, rho, 0, choice, mode == ORC_MODE_CANDIDATE ? choice_nn : NULL, NULL,
                      NULL);
            orc_construct(mode, n, m, k, dist, mode == ORC_MODE_CANDIDATE ? nn : NULL,
                          mode == ORC_MODE_FULLROW ? nn : NULL, choice, choice_nn, NULL, tau, mean, eta,
                          eta_len, alpha, aging ? age : NULL, aging ? valid : NULL, converged,
                          min_age, seed, (uint32_t)it, stream, tours, lengths, NULL);
        }
        int32_t nd = orc_stats(variant, n, m, w, tours, lengths, aging ? valid : NULL,
                               aging ? perf : NULL, aging ? prev_len : NULL, st, best_tour,
                               dep_ant, dep_w);
        /* evaporate, lower triangle mirrored (src/AntSystem.cu:218-223) */
        for (int32_t i = 0; i < n; i++)
            for (int32_t j = 0; j <= i; j++) {
                float p = (1.0f - rho) * tau[(int64_t)i * n + j];
                float v = ((double)p > 1e-18) ? p : (float)1e-18;
                tau[(int64_t)i * n + j] = v;
                tau[(int64_t)j * n + i] = v;
            }
        for (int32_t e = 0; e < nd; e++) {
            const int32_t *t = dep_ant[e] >= 0 ? tours + (int64_t)dep_ant[e] * n : best_tour;
            int64_t L = dep_ant[e] >= 0 ? lengths[dep_ant[e]] : st[1];
            float d = (float)((double)dep_w[e] / (double)L);
            for (int32_t i = 0; i < n; i++) {
                int32_t a = t[i], b = t[(i + 1) % n];
                tau[(int64_t)a * n + b] += d;
                tau[(int64_t)b * n + a] = tau[(int64_t)a * n + b];
            }
        }
        if (aging)
            orc_aging_control(m, age, valid, &num_valid, &converged, (int32_t)st[3], max_age,
                              global_perf, min_pop);
    }
    double elapsed = now_seconds() - t0;
    out[0] = st[1];
    out[1] = st[0];
    free(eta); free(tau); free(choice); free(mean); free(nn); free(choice_nn);
    free(tours); free(lengths); free(prev_len); free(perf); free(age); free(valid);
    free(dep_ant); free(dep_w);
    return elapsed;
}
