/*
 * aco_oracle.h -- CPU restatement of the reference Ant-System hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (acogpu_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference tree of observit17/acogpu).  Where the reference is
 * non-deterministic or buggy (GPU summation of float bit patterns, cuRAND,
 * clock-seeded ran2) the oracle follows the *algorithm* and fixes the
 * arithmetic order explicitly so that the MI355X kernels can be replayed
 * bit-exactly; each such choice is named at the function.
 *
 * Pinning status (what each function is checked against in tests/):
 *   orc_philox4x32_10      published Random123 known-answer vectors
 *   orc_euc2d_*            TSPLIB closed form, brute force, a280 duplicate
 *   orc_nn_lists           brute-force stable argsort
 *   orc_eta_table          closed form (pow) spot checks
 *   orc_pow_alpha          fp64 numpy power rounded to fp32 on a grid (<= 1 ulp,
 *                          equal almost everywhere); exact powers of two
 *   orc_wave_scan          exact integer prefix sums; fp64 error bound
 *   orc_select_candidate   chi-square frequencies; exact-arithmetic boundaries
 *   orc_sweep              elementwise numpy closed forms of choice, choice_nn
 *                          and the evaporated tau (floor included), alpha 1
 *                          and 1.5, heuristic override, d = 0, diagonal
 *   orc_construct          permutation + length invariants; chi-square of the
 *                          first move against numpy weights in all three modes
 *                          (young ants too); brute-force farthest-unvisited
 *                          walks on zero and overflowing sums; argmax fallback
 *   (tests/oracle_pins.py; tests/test_oracle_mutants.py plants one mistake per
 *    copy of this file and requires these pins to fail on it)
 *   orc_stats              brute-force argmin / ranking
 *   orc_deposit_f64        numpy np.add.at restatement over tour edges
 *   orc_aging_*            closed-form retirement/recruitment traces
 *   orc_run_refseq         TSPLIB optimum lower bound, long-run convergence
 */
#ifndef ACO_ORACLE_H
#define ACO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_DISTANCE 0x7fffffff

/* construction modes */
#define ORC_MODE_CANDIDATE 0   /* k nearest neighbours, argmax fallback      */
#define ORC_MODE_FULLROW 1     /* all n cities in NN order, wave-chunk sums  */
#define ORC_MODE_REFSEQ 2      /* all n cities, sequential sums (reference)  */

/* variants */
#define ORC_AS 0
#define ORC_RBAS 1

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t orc_rng_word(uint64_t seed, uint32_t iter, uint32_t ant, uint32_t step, uint32_t stream);
float orc_uniform(uint32_t word);
int32_t orc_start_city(uint32_t word, int32_t n);

int32_t orc_euc2d_exact(int64_t dx, int64_t dy);
int32_t orc_euc2d_ref(float xi, float xj, float yi, float yj);
void orc_dist_from_int_coords(const int32_t *xy, int32_t n, int32_t *dist);
void orc_dist_from_float_coords(const float *x, const float *y, int32_t n, int32_t *dist);

void orc_nn_lists(const int32_t *dist, int32_t n, int32_t k, int32_t *nn);
int32_t orc_eta_table(double beta, int32_t dmax, float *table);
float orc_pow_alpha(float p, float alpha);
float orc_eta(const float *table, int32_t table_len, int32_t d);

void orc_wave_scan(float v[64]);
int32_t orc_select_candidate(const float w[64], float u, float *total_out);

/* heur_nn (n*k, may be NULL): heuristic on candidate edges replacing eta
 * for choice_nn (a learned prior); the full choice matrix keeps eta. */
void orc_sweep(int32_t n, int32_t k, const float *tau, const int32_t *dist,
               const float *eta, int32_t eta_len, const int32_t *nn, const float *heur_nn,
               float alpha, float rho, int32_t want_mean,
               float *choice, float *choice_nn, float *tau_evap, float *mean);

/* Per-iteration construction of every valid ant.  Arrays are row-major:
 * tours[m*n] (int32), lengths[m] (int64).  age/valid may be NULL (no aging).
 * fallback_out (may be NULL) receives [number of fallback steps, number of
 * unvisited cities scanned by fallback steps]. */
void orc_construct(int32_t mode, int32_t n, int32_t m, int32_t k,
                   const int32_t *dist, const int32_t *nn, const int32_t *nnfull,
                   const float *choice, const float *choice_nn, const float *heur_nn,
                   const float *tau, const float *mean,
                   const float *eta, int32_t eta_len, float alpha,
                   const int32_t *age, const uint8_t *valid, int32_t converged,
                   int32_t min_age, uint64_t seed, uint32_t iter, uint32_t stream,
                   int32_t *tours, int64_t *lengths, int64_t *fallback_out);

/* Iteration bookkeeping after construction (computeAntDistances).
 * state64[0]=iter_best, [1]=glob_best, [2]=iter_best_ant, [3]=gperf.
 * perf/prev_len may be NULL when aging is off.
 * dep_ant[] (size >= max(m, w)) receives the depositing ants in order
 * (-1 = best-so-far tour), dep_w[] their integer weights; returns the count. */
int32_t orc_stats(int32_t variant, int32_t n, int32_t m, int32_t w,
                  const int32_t *tours, const int64_t *lengths, const uint8_t *valid,
                  int32_t *perf, int64_t *prev_len, int64_t *state64,
                  int32_t *best_tour, int32_t *dep_ant, int32_t *dep_w);

void orc_deposit_f64(int32_t n, double *tau, const int32_t *tours, const int64_t *lengths,
                     const int32_t *best_tour, int64_t best_len,
                     const int32_t *dep_ant, const int32_t *dep_w, int32_t ndep);
void orc_evaporate_f64(int32_t n, double *tau, float rho);

void orc_aging_start(int32_t m, int32_t *age);
void orc_aging_control(int32_t m, int32_t *age, uint8_t *valid, int32_t *num_valid,
                       int32_t *converged, int32_t gperf, int32_t max_age,
                       int32_t global_perf, int32_t min_pop);

int32_t orc_greedy_length(const int32_t *dist, int32_t n, int32_t start);

/* Whole-run driver with the reference CPU semantics (timed CPU baseline).
 * params: [0]=variant [1]=mode [2]=k [3]=w [4]=aging [5]=min_age [6]=max_age
 *         [7]=global_perf [8]=min_pop [9]=iterations [10]=Philox stream
 * fparams: [0]=alpha [1]=beta [2]=rho [3]=tau0
 * out: [0]=glob_best [1]=last iter_best ; returns elapsed seconds of the
 * iteration loop (steady clock), or -1 on allocation failure. */
double orc_run(int32_t n, int32_t m, const int32_t *dist, const int32_t *params,
               const float *fparams, uint64_t seed, int64_t *out, int32_t *best_tour);

/* Learned edge prior (prior_oracle.c). */
void orc_prior_features(int32_t n, int32_t k, const int32_t *dist, const int32_t *nn, float *feat);
void orc_prior_labels(int32_t n, int32_t k, const int32_t *nn, const int32_t *tour, float *y);
double orc_bf16(double x);
double orc_gelu(double x);
double orc_gelu_grad(double x);
void orc_mlp_forward(const float *params, const float *x, int64_t count, int32_t bf16, double *out);
double orc_mlp_grad(const float *params, const float *x, const float *y, int64_t count, float pos_weight,
                    int32_t bf16, double *grad);
void orc_adam_step(int32_t nparams, float *params, const float *grad, float *m, float *v, int32_t step, float lr,
                   float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif
