/*
 * acogpu.h -- C ABI of the MI355X-native Ant Colony Optimization solver.
 *
 * The reference (observit17/acogpu) has no C or FFI interface: its boundary
 * is the C++ class Colony (inc/Colony.h:24-155) with subclasses AntSystem
 * (inc/AntSystem.h:7-23) and RankBasedAntSystem (inc/RankBasedAntSystem.h:
 * 12-28), called by main (src/Setup.cpp:73-129) and TSP::execute
 * (src/TSP.cpp:23-79).  Each entry point below names the member function it
 * replaces.  The C++ classes of acogpu_amd/csrc/colony/ are thin wrappers over
 * this ABI; any other host language binds it directly (ctypes, cgo, JNI).
 *
 * Conventions: every call returns 0 on success or a negative ACO_E* code and
 * never exits the process; aco_last_error() holds the message.  No process
 * globals: one aco_ctx per colony, calls on one ctx are not thread-safe,
 * calls on different ctxs are independent.  Host arrays are plain row-major
 * C arrays; device memory never crosses the ABI except the optional device
 * distance matrix accepted by aco_create (the reference's GPU convention,
 * src/Setup.cpp:60-66).
 */
#ifndef ACOGPU_H
#define ACOGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACO_OK 0
#define ACO_E_ARG (-1)      /* invalid argument / parameter combination */
#define ACO_E_HIP (-2)      /* HIP runtime error (message has the HIP string) */
#define ACO_E_NOMEM (-3)    /* device allocation failed */
#define ACO_E_STATE (-4)    /* call out of order (e.g. iterate before initialize) */
#define ACO_E_RANGE (-5)    /* input outside the supported range */

#define ACO_VARIANT_AS 0    /* AntSystem (src/AntSystem.cu)                   */
#define ACO_VARIANT_RBAS 1  /* RankBasedAntSystem (src/RankBasedAntSystem.cu) */

#define ACO_MODE_CANDIDATE 0 /* k-nearest candidate lists + argmax fallback   */
#define ACO_MODE_FULLROW 1   /* every city in NN order, as the reference does */

#define ACO_TAU0_CONST 0    /* tau0 field (reference: 1e-6, inc/default.h:16) */
#define ACO_TAU0_AS 1       /* m / C_nn                                       */
#define ACO_TAU0_RBAS 2     /* 0.5 w (w-1) / (rho C_nn) (src/RankBasedAntSystem.cu:44) */

/* timers[] slots, same order as enum StatsTimer (inc/Statistics.h:7-15) */
#define ACO_TIME_INIT_TOUR 0
#define ACO_TIME_TOUR 1
#define ACO_TIME_COMPUTE_ANT_DISTANCE 2
#define ACO_TIME_EVAPORATE_PHEROMONE 3
#define ACO_TIME_UPDATE_PHEROMONE 4
#define ACO_TIME_COMPUTE_PROBABILITIES 5
#define ACO_TOTAL_TIMERS 6

typedef struct aco_ctx aco_ctx;

/* Colony + AntSystem + aging constructor arguments (src/Colony.cu:25-137,
 * src/AntSystem.cu:160-181, inc/Colony.h:118-124) and CLI flags
 * (src/Commandline.cpp:98-231). */
typedef struct aco_params {
    int32_t variant;        /* ACO_VARIANT_*                                  */
    int32_t mode;           /* ACO_MODE_*                                     */
    int32_t n;              /* cities                                         */
    int32_t m;              /* ants (-m)                                      */
    int32_t k;              /* candidate list length, 1..64 (candidate mode)  */
    int32_t w;              /* RBAS ranked ants (RankBasedAntSystem::setW)    */
    float alpha;            /* -a                                             */
    float beta;             /* -b                                             */
    float rho;              /* -r                                             */
    int32_t tau0_policy;    /* ACO_TAU0_*                                     */
    float tau0;             /* used by ACO_TAU0_CONST                         */
    int32_t aging;          /* 0/1: the "aging" build variant                 */
    int32_t min_age;        /* -j minRetirementAge                            */
    int32_t max_age;        /* -k maxRetirementAge                            */
    int32_t local_perf;     /* -l localPeformance (unused by the reference)   */
    int32_t global_perf;    /* -g globalPerformance                           */
    int32_t min_pop;        /* -x minPopulationLimit                          */
    int32_t max_pop;        /* -y maxPopulationLimit                          */
    int32_t broker;         /* "broker" build variant: accepted, no behaviour */
    int32_t ants_per_block; /* construction launch shape 1..8 (batch 1..16; 1..8 if n > 2048), 0 = automatic */
    uint64_t seed;          /* Philox key (reference: RANDOM_SEED, inc/Colony.h:15) */
    uint32_t stream_id;     /* Philox counter word 3 (instance id)            */
    int32_t construct_path; /* candidate mode: 0 auto, 1 global memory, 2 LDS-staged */
    int32_t deposit_path;   /* Ant System: 0 auto, 1 fp32 atomics, 2 adjacency gather (n <= 32768;
                               batch n <= 10240) */
    int32_t ants_per_wave;  /* batch construction: ants sharing one wave, 1, 2 (k <= 32, n <= 1024)
                               or 4 (k <= 16, n <= 512); 0 = the most that fit */
    int32_t reserved[4];
} aco_params;

/* Reference defaults (inc/default.h:8-39, src/Colony.cu:32-36). */
int aco_params_init(aco_params *p);

/* Colony::Colony + AntSystem::AntSystem / RankBasedAntSystem::
 * RankBasedAntSystem.  Exactly one of dist (n*n int32, host or device
 * pointer; diagonal ignored) or xy (2n int32 host coordinates, interleaved,
 * |coordinate| < 2^29) is non-NULL.  Allocates every device buffer once. */
int aco_create(const aco_params *p, const int32_t *dist, const int32_t *xy, aco_ctx **out);

/* Colony::initialize (src/Colony.cu:185-214): pheromone fill, heuristic
 * table, NN lists, aging arrays, best-so-far reset. */
int aco_initialize(aco_ctx *c);

/* Colony::constructSolution (src/Colony.cu:537-594): one iteration, phase
 * times in milliseconds (timers may be NULL), each from the start of the
 * phase's first kernel to the start of the next phase's (the last phase: to
 * the end of the iteration), from wall-clock stamps the kernels store
 * themselves; one wait. */
int aco_iterate(aco_ctx *c, float *timers);

/* aco_iterate plus what Colony::constructSolution's caller reads next:
 * getIterBestDist / getGlobBestDist / getNumValidAnts after the iteration
 * (src/Colony.cu:1096-1104, 447-467; each output may be NULL), written by the
 * statistics kernel to host memory, so the wait is the only host round trip. */
int aco_iterate_stats(aco_ctx *c, float *timers, int64_t *iter_best, int64_t *glob_best, uint32_t *valid_ants);

/* One phase of Colony::constructSolution (src/Colony.cu:537-594), for hosts
 * that drive the virtual steps themselves.  Order per iteration:
 * PROBABILITIES (computeProbabilities: choice_info, row means, and the
 * evaporated pheromone of AntSystem::evaporatePheromones written to the next
 * buffer), TOUR (initTour + constructPath), ANT_DISTANCE
 * (computeAntDistances), EVAPORATE (no kernel: folded into PROBABILITIES;
 * may be skipped), UPDATE_PHEROMONE (updatePheromones + aging
 * retire/recruit).  A non-NULL ms waits for the phase and gets its HIP-event
 * time; ms == NULL returns at once (aco_phase_end completes the iteration). */
#define ACO_PHASE_PROBABILITIES 0
#define ACO_PHASE_TOUR 1
#define ACO_PHASE_ANT_DISTANCE 2
#define ACO_PHASE_EVAPORATE 3
#define ACO_PHASE_UPDATE_PHEROMONE 4
int aco_phase(aco_ctx *c, int32_t phase, float *ms);

/* With ms == NULL aco_phase enqueues its phase and returns without waiting;
 * after the UPDATE_PHEROMONE phase of such an iteration, aco_phase_end waits
 * once and returns what Colony::constructSolution's caller reads next: the
 * six timers (StatsTimer slots, ms) and getIterBestDist / getGlobBestDist /
 * getNumValidAnts after the iteration (src/Colony.cu:537-594, 1096-1104;
 * each output may be NULL).  One host synchronisation per iteration. */
int aco_phase_end(aco_ctx *c, float *timers, int64_t *iter_best, int64_t *glob_best, uint32_t *valid_ants);

/* TSP::execute's loop body for `iterations` iterations with no host
 * synchronisation inside (hipGraph replay).  elapsed_ms (may be NULL) gets the
 * HIP-event time of the whole loop. */
int aco_run(aco_ctx *c, int32_t iterations, float *elapsed_ms);

/* TSP::execute's loop (src/TSP.cpp:23-79) with the optimum test of
 * TSP::terminationCondition (src/TSP.cpp:168-171) evaluated on the device:
 * up to max_iterations graph-replayed iterations, halting after the first
 * iteration whose iteration-best length is <= stop_at (stop_at < 0: never);
 * later iterations are skipped on the device and the host reads the state
 * once per `chunk` iterations.  iterations_run / halted (may be NULL) report
 * what happened; the colony continues normally afterwards. */
int aco_run_until(aco_ctx *c, int32_t max_iterations, int64_t stop_at, int32_t chunk, int32_t *iterations_run,
                  int32_t *halted);

/* aco_run_until with the time test of TSP::terminationCondition as well
 * (elapsed >= maxTime, src/TSP.cpp:168-171): before every chunk the run stops
 * once the HIP-event time of the iterations run reaches time_limit_ms (< 0:
 * no limit; 0 runs none).  halted = 1 (optimum), 2 (time limit) or 0;
 * elapsed_ms (may be NULL) gets that time. */
int aco_run_until_time(aco_ctx *c, int32_t max_iterations, int64_t stop_at, float time_limit_ms, int32_t chunk,
                       int32_t *iterations_run, int32_t *halted, float *elapsed_ms);

/* getIterBestDist / getGlobBestDist / getNumValidAnts (src/Colony.cu:
 * 1096-1104, 447-467). */
int aco_get_stats(aco_ctx *c, int64_t *iter_best, int64_t *glob_best, uint32_t *valid_ants);

/* Colony::getTour (src/Colony.cu:1173-1180) as 0-based city indices (n). */
int aco_get_best_tour(aco_ctx *c, int32_t *tour);

/* Parity taps. */
int aco_get_tours(aco_ctx *c, int32_t *tours);            /* m*n          */
int aco_get_lengths(aco_ctx *c, int64_t *lengths);        /* m            */
int aco_get_pheromone(aco_ctx *c, float *tau);            /* n*n, current */
int aco_set_pheromone(aco_ctx *c, const float *tau);      /* n*n          */
int aco_get_choice(aco_ctx *c, float *choice);            /* n*n, last sweep */
int aco_get_choice_nn(aco_ctx *c, float *choice_nn);      /* n*k          */
int aco_get_nn(aco_ctx *c, int32_t *nn);                  /* n*k (or n*n full-row) */
int aco_get_mean(aco_ctx *c, float *mean);                /* n            */
int aco_get_dist(aco_ctx *c, int32_t *dist);              /* n*n          */
int aco_get_eta(aco_ctx *c, float *eta, int32_t *eta_len);/* eta may be NULL to query */
/* aging state: age[m], valid[m], perf[m], prev_len[m];
 * scalars = {num_valid, converged, gperf, iteration} */
int aco_get_aging(aco_ctx *c, int32_t *age, uint8_t *valid, int32_t *perf, int64_t *prev_len,
                  int32_t *scalars);
int aco_set_aging(aco_ctx *c, const int32_t *age, const uint8_t *valid, const int32_t *perf,
                  const int64_t *prev_len, const int32_t *scalars);
/* Heuristic on the candidate edges, n*k values in candidate-list order:
 * replaces eta^beta on NN-list edges only (fallback steps keep plain
 * eta^beta); NULL restores eta^beta.  No counterpart in the reference: the
 * entry point for a learned edge prior (aco_prior_apply writes it on the
 * device). */
int aco_set_heuristic_nn(aco_ctx *c, const float *h);
int aco_get_heuristic_nn(aco_ctx *c, float *h);
/* best-so-far length/tour overwrite (parity replays) */
int aco_set_best(aco_ctx *c, int64_t glob_best, const int32_t *tour);
/* counters[4] = {fallback steps, cities whose weight fallback steps compared
 * (every unvisited city of a row scan, one for a top-list answer), tours
 * constructed (the valid ants of every iteration run), fallback steps
 * answered by the row's top list (n >= 4096, alpha = 1)} since
 * aco_initialize */
int aco_get_counters(aco_ctx *c, uint64_t *counters);

/* Launch configuration chosen at create: config = {construction path
 * (0 global memory, 1 full row, 2 LDS-staged), ants per block, blocks
 * reserved for young aging ants, LDS bytes per block available}. */
int aco_get_config(aco_ctx *c, int32_t *config);

/* Runs `iterations` iterations on the ctx stream, each phase's first kernel
 * storing its start time (device wall clock) and a last one-thread kernel the
 * end; kernel_ms[0..3] = average ms of {sweep, construct, stats, deposit},
 * each from its kernel's start to the next one's. */
int aco_profile(aco_ctx *c, int32_t iterations, float *kernel_ms);

/* Latency floor of one construction step at this colony's k: a single wave
 * repeats the LDS-staged step's dependent chain (candidate entry and weight
 * from LDS, visited test by ds_bpermute, roulette scan, ballot and pick)
 * `iters` times; out[0] = core clock cycles per step, out[1] = ns per step.
 * No counterpart in the reference (measurement only). */
int aco_probe_step(aco_ctx *c, int32_t iters, float *out);

/* Memory-bandwidth calibration: copies `bytes` (multiple of 16) device to
 * device with a plain streaming kernel `reps` times; returns ms per copy. */
int aco_calibrate_copy(aco_ctx *c, int64_t bytes, int32_t reps, float *ms);

/* ------------------------------------------------------------------------
 * Batched solver: `count` independent instances solved together, one
 * persistent construction launch per iteration.  No counterpart in the
 * reference (one instance per process, src/Setup.cpp:73-129); every instance
 * runs the same iteration as an aco_ctx.  Shared parameters come from p (its n
 * and m are ignored; modes: candidate lists staged in LDS, no aging); n[i]
 * cities and m[i] ants (m == NULL: m[i] = n[i]) per instance, xy the
 * concatenated integer coordinates (2 n[i] per instance).  Instance i draws
 * from Philox stream i, so its tours equal an aco_ctx's with stream_id = i. */
typedef struct aco_batch aco_batch;
int aco_batch_create(const aco_params *p, int32_t count, const int32_t *n, const int32_t *m, const int32_t *xy,
                     aco_batch **out);
int aco_batch_initialize(aco_batch *b);
int aco_batch_run(aco_batch *b, int32_t iterations, float *elapsed_ms);
/* glob_best[count]; tours (may be NULL): sum of n[i], instance after instance */
int aco_batch_get_best(aco_batch *b, int64_t *glob_best, int32_t *tours);
int aco_batch_get_tours(aco_batch *b, int32_t instance, int32_t *tours);     /* m[i] x n[i] */
int aco_batch_get_lengths(aco_batch *b, int32_t instance, int64_t *lengths); /* m[i]       */
int aco_batch_get_pheromone(aco_batch *b, int32_t instance, float *tau);      /* n[i] x n[i] */
int aco_batch_set_pheromone(aco_batch *b, int32_t instance, const float *tau);
/* config[5] = {ants per block, persistent blocks, work items, LDS bytes per
 * block, ants per wave} */
int aco_batch_get_config(aco_batch *b, int32_t *config);
/* Per-kernel HIP-event times of `iterations` eagerly launched iterations:
 * kernel_ms[4] = averages of {sweep, construction, statistics, deposit}, as
 * aco_profile. */
int aco_batch_profile(aco_batch *b, int32_t iterations, float *kernel_ms);
/* counters[2 * count] = per instance {fallback steps, unvisited cities
 * scanned by fallback steps} since create, as aco_get_counters */
int aco_batch_get_counters(aco_batch *b, uint64_t *counters);
/* candidate-edge heuristic of one instance (n[i] * k values), as aco_set_heuristic_nn */
int aco_batch_set_heuristic_nn(aco_batch *b, int32_t instance, const float *h);
/* Times construction launches for each ants-per-block choice, with one ant
 * per wave and with the most ants per wave the batch allows (unless
 * ants_per_wave fixes it): launch parameters only, tours do not change.
 * Keeps the fastest; reps launches per choice; chosen (may be NULL) gets the
 * ants per block (aco_batch_get_config reports both). */
int aco_batch_autotune(aco_batch *b, int32_t reps, int32_t *chosen);
/* Candidate-list width tuner: builds a batch of the same instances for each
 * k in ks[nks], runs trial_iterations iterations (after aco_batch_autotune)
 * and picks the fastest k whose mean best length is within `tolerance`
 * (relative) of the best mean; report (may be NULL) gets nks x {ms per
 * iteration, mean best length, ants per block}.  Changes the candidate set,
 * hence the tours: the caller creates the production batch with chosen_k. */
int aco_batch_tune_k(const aco_params *p, int32_t count, const int32_t *n, const int32_t *m, const int32_t *xy,
                     const int32_t *ks, int32_t nks, int32_t trial_iterations, float tolerance, int32_t *chosen_k,
                     float *report);
/* Island model on a batch whose instances are copies of one TSP (island i
 * draws from Philox stream i).  The reference's broker build variant is an
 * empty extension point (inc/Colony.h:142-149); this is its multi-colony
 * reading: aco_batch_exchange_best gives every island the shortest
 * best-so-far tour over all islands (ties: lowest island); run_islands runs
 * `iterations` iterations exchanging after every `exchange_every`.  The
 * exchange matters for the rank-based variant, whose best-so-far tour
 * deposits. */
int aco_batch_exchange_best(aco_batch *b);
int aco_batch_run_islands(aco_batch *b, int32_t iterations, int32_t exchange_every, float *elapsed_ms);
int aco_batch_destroy(aco_batch *b);
const char *aco_batch_last_error(const aco_batch *b);

/* ------------------------------------------------------------------------
 * Learned edge prior.  No counterpart in the reference (its heuristic is
 * eta^beta only, src/AntSystem.cu:56-77).  An MLP 8 -> 64 -> 64 -> 1 with
 * GELU, bf16 weights/activations and fp32 accumulation scores every
 * candidate edge (i, nn[i][r]) from 8 features of the distance matrix; the
 * heuristic sigmoid(score)^gamma (transform 0) or exp(gamma score)
 * (transform 1) then replaces eta^beta on the candidate edges
 * (aco_set_heuristic_nn); fallback steps keep eta^beta.  Training data
 * are candidate edges labelled 1 when they lie on a solver's best-so-far
 * tour; training runs on the device (weighted binary cross-entropy, Adam).
 *
 * Edge features, d = dist(i,j), D_i^q = distance to the q-th neighbour of i:
 *   f0 d/(D_i^{k-1}+1)  f1 d/(D_j^{k-1}+1)  f2 r/k  f3 (rank of i in NN(j),
 *   k if absent)/k  f4 d n/(sum_i D_i^0 + n)  f5 D_i^0/(D_i^{k-1}+1)
 *   f6 D_j^0/(D_j^{k-1}+1)  f7 (d-D_i^0)/(D_i^{k-1}-D_i^0+1).
 * Parameters: ACO_PRIOR_PARAMS floats, W1 (64x8, row = output) at 0, b1 at
 * 512, W2 (64x64) at 576, b2 at 4672, w3 (64) at 4736, b3 at 4800. */
#define ACO_PRIOR_PARAMS 4801
#define ACO_PRIOR_FEATURES 8

typedef struct aco_prior aco_prior;
typedef struct aco_prior_config {
    uint64_t seed;      /* weight initialisation (uniform Glorot, zero biases) */
    float lr;           /* Adam step size */
    float beta1, beta2, eps;
    float pos_weight;   /* loss weight of positive (on-tour) edges */
    float gamma;        /* heuristic exponent */
    int32_t minibatch;  /* edges per Adam step */
    int32_t transform;  /* 0: sigmoid(score)^gamma, 1: exp(gamma min(score, 80/gamma)) */
    int32_t reserved[7];
} aco_prior_config;

void aco_prior_config_init(aco_prior_config *cfg);
int aco_prior_create(const aco_prior_config *cfg, aco_prior **out);
int aco_prior_get_params(aco_prior *pr, float *params);
int aco_prior_set_params(aco_prior *pr, const float *params); /* resets Adam moments */
/* Appends the candidate edges of a colony / every batch instance, labelled
 * by its best-so-far tour, to the training set (device resident). */
int aco_prior_collect(aco_prior *pr, aco_ctx *c);
int aco_prior_collect_batch(aco_prior *pr, aco_batch *b);
int aco_prior_clear(aco_prior *pr);
int64_t aco_prior_size(const aco_prior *pr);
/* `epochs` passes of minibatch Adam over the training set in collection
 * order; loss (may be NULL) gets the mean loss of each epoch. */
int aco_prior_train(aco_prior *pr, int32_t epochs, float *loss);
/* Writes the prior heuristic into the candidate edges of an initialized
 * colony / every batch instance (after aco_initialize, which resets them). */
int aco_prior_apply(aco_prior *pr, aco_ctx *c);
int aco_prior_apply_batch(aco_prior *pr, aco_batch *b);
/* aco_prior_apply without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code in
 * tight serving loops, harnesses that forbid hipMalloc).  The features and
 * scores live in a fixed library scratch array of 9.0 MiB per GPU, held
 * until the process ends; colonies with n*k > 262144 candidate edges are
 * refused with ACO_E_RANGE.  Calls that overlap on one GPU share that array
 * and race.  The heuristic written equals aco_prior_apply's bit for bit. */
int aco_prior_apply_noalloc(aco_prior *pr, aco_ctx *c);
/* Host-array taps: features (n*k*8) and labels (n*k) of a colony's edges;
 * scores of `count` feature rows; loss gradient (no update); one Adam step
 * from a given gradient. */
int aco_prior_features(aco_prior *pr, aco_ctx *c, float *features, float *labels);
int aco_prior_forward(aco_prior *pr, const float *x, int64_t count, float *scores);
int aco_prior_gradient(aco_prior *pr, const float *x, const float *y, int64_t count, float *grad, float *loss);
int aco_prior_adam_step(aco_prior *pr, const float *grad);
/* ms = {forward ms, backward ms, edges}: average of `reps` launches of each
 * kernel over the first minibatch of the training set (HIP events). */
int aco_prior_profile(aco_prior *pr, int32_t reps, float *ms);
int aco_prior_destroy(aco_prior *pr);
const char *aco_prior_last_error(const aco_prior *pr);

/* Colony::~Colony */
int aco_destroy(aco_ctx *c);

/* Message of the last failing call on c (NULL: last aco_create failure on
 * this thread). Never NULL. */
const char *aco_last_error(const aco_ctx *c);

const char *aco_version(void);

#ifdef __cplusplus
}
#endif
#endif
