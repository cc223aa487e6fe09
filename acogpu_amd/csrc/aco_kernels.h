// Internal launch API of the gfx950 kernels (aco_kernels.hip).  Not part of
// the public ABI; aco_capi.hip is the only caller.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aco {

// Device-resident iteration state: everything the host used to read back
// every iteration in the reference (src/Colony.cu:918-925,985, 447-467) stays
// here so an iteration is a fixed sequence of kernels (graph-capturable).
struct DevState {
    int64_t iter_best;       // Colony::iterBestDist
    int64_t glob_best;       // Colony::globBestDist
    int32_t iter_best_ant;
    int32_t gperf;           // Colony::globalperformance (aging)
    uint32_t iter;           // iteration counter (Philox counter word 2)
    int32_t num_valid;       // numValidAnts (inc/Colony.cuh:10)
    int32_t converged;       // isconverged (inc/Colony.cuh:11)
    int32_t ndep;            // entries in the deposit list
    unsigned long long fb_steps;    // fallback steps taken
    unsigned long long fb_scanned;  // unvisited cities read by fallback steps
    int64_t stop_at;         // halt after the first iteration whose best <= stop_at (-1: never)
    int32_t halted;          // set by the statistics kernel; cleared by the host
    int32_t skip;            // the current iteration is skipped (the sweep copies halted)
    unsigned long long tours;       // tours constructed (valid ants of every iteration run)
    unsigned long long fb_listed;   // fallback steps answered by the row's top list
    uint32_t fuse_count;     // blocks of the fused statistics + deposit done reading (0 between launches)
};

// Top list of a pheromone row (large instances, alpha = 1): the TOP_K cities
// with the largest selection weight, in the order (weight descending, index
// ascending) that the fallback's argmax uses.  Rebuilt by every sweep; a
// fallback takes the first unvisited entry and scans the row only when all
// TOP_K are visited.  TOP_INVALID in entry 0 marks a row without a list.
constexpr int TOP_K = 256;
constexpr int TOP_BUF = 512;               // per-wave LDS key buffer of the selection
constexpr int32_t TOP_MIN_N = 4096;        // smallest n with top lists
constexpr uint16_t TOP_INVALID = 0xffff;

struct SweepArgs {
    const float* tau;        // current pheromone (pitched n x P)
    const int32_t* dist;     // pitched n x P
    const float* eta;        // heuristic table by distance
    int32_t eta_len;
    const int32_t* nn;       // n x k candidate lists (may be null)
    int32_t k;
    float* choice;           // pitched n x P
    float* choice_nn;        // n x k
    const float* eta_nn;     // n x k heuristic on candidate edges (eta^beta or a learned prior)
    float* tau_nn;           // n x k pheromone on candidate edges (aging)
    float* tau_next;         // evaporated pheromone (pitched)
    float* mean;             // n row means (aging)
    int32_t* age;            // m ant ages: incremented here (aging)
    int32_t m;
    int32_t n, P;
    float alpha, one_minus_rho;
    int32_t want_mean;
    int32_t eta_pow;         // 1..3: eta computed as 1 / d^eta_pow (equal to the table); 0: table
    DevState* state;         // single colony: halting flags (null in a batch)
    uint16_t* top;           // n x TOP_K row top lists (null: none)
    float* top_thr;          // n: the TOP_K-th weight of the previous sweep
    const float* eta_full;   // pitched heuristic of every edge (read instead of dist; may be null)
    const int32_t* xy;       // 2n interleaved coordinates, |x| < XY_SWEEP_LIMIT: distances computed, not read (may be null)
    uint64_t* stamp;         // set: block 0 writes the kernel's start (wall_clock64) here
};

// Largest coordinate magnitude (exclusive) for which the sweep computes its
// rows' distances from the coordinates (see euc2d_fast).
constexpr int32_t XY_SWEEP_LIMIT = 1 << 20;

struct ConstructArgs {
    const int32_t* nn;       // n x k (candidate) or n x n (full row)
    const uint16_t* nn16;    // n x k, 16-bit copy for LDS staging
    const float* choice_nn;  // n x k
    const float* tau_nn;     // n x k (aging young ants)
    const float* eta_nn;     // n x k heuristic on candidate edges
    const float* choice;     // pitched
    const int32_t* dist;     // pitched
    const float* tau;        // pitched, current pheromone (aging young ants)
    const float* mean;       // n
    const float* eta;
    int32_t eta_len;
    int32_t* tours;          // m x n
    int64_t* lengths;        // m
    int32_t* age;            // m (aging)
    const uint8_t* valid;    // m (aging)
    DevState* state;
    int32_t n, P, m, k;
    int32_t aging, min_age;
    float alpha;
    uint32_t seed_lo, seed_hi, stream;
    int32_t words;           // visited-bitmap words per ant
    int32_t young_blocks;    // trailing blocks reserved for young aging ants (staged path)
    const uint16_t* top;     // n x TOP_K row top lists (null: none)
    const float* eta_full;   // pitched heuristic of every edge (aging: young ants' fallback rows; null: gathered)
    uint64_t* stamp;         // set: block 0 writes the kernel's start (wall_clock64) here
};

struct StatsArgs {
    const int32_t* tours;
    const int64_t* lengths;
    uint8_t* valid;
    int32_t* age;
    int32_t* perf;
    int64_t* prev_len;
    int32_t* best_tour;
    int32_t* dep_ant;
    float* dep_val;
    DevState* state;
    int32_t n, m, variant, w;
    int32_t aging, max_age, global_perf, min_pop;
    // nullable: {iteration best, best so far, valid ants} written to host
    // (coherent pinned) memory after the statistics, so the host reads them
    // after the iteration with no copy
    int64_t* host_stats;
    uint64_t* stamp;         // set: block 0 writes the kernel's start (wall_clock64) here
};

struct DepositArgs {
    const int32_t* tours;
    const int32_t* best_tour;
    const int32_t* dep_ant;
    const float* dep_val;
    const DevState* state;
    float* tau_next;
    int32_t n, P, dense;
    // Ant System adjacency deposit (dense list): row x of adj (adj_pitch(m)
    // 16-bit slots) holds, for ant a, its successor (slot 2a) and
    // predecessor (2a+1) of x.
    uint16_t* adj;
    int32_t m;
    int32_t list_cap;        // slots of dep_ant / dep_val (m + w + 2)
    uint64_t* stamp;         // set: block 0 writes the kernel's start (wall_clock64) here
};

// Heuristic eta^beta of an integer distance d.  beta = 1, 2 or 3 with
// dmax^beta < 2^53 (d^beta exact in fp64): 1 / d^beta, one fp64 division
// rounded once to fp32, which a kernel evaluates on the fly instead of
// gathering it from the table (eta_pow = beta); any other beta: fp64
// pow(1 / d, beta) rounded once (table only, eta_pow = 0).  d = 0 (duplicate
// cities) is taken as d = 1/2.
__host__ __device__ inline int32_t eta_integer_power(double beta, int32_t dmax) {
    const int32_t b = (int32_t)beta;
    if ((double)b != beta || b < 1 || b > 3) return 0;
    double p = 1.0;
    for (int32_t q = 0; q < b; ++q) p *= (double)dmax;
    return p < 9007199254740992.0 ? b : 0;
}
__host__ __device__ inline float eta_power_value(int32_t pw, int32_t d) {
    if (d == 0) return (float)(1 << pw);
    const double dd = (double)d;
    double p = dd;
    if (pw >= 2) p *= dd;
    if (pw >= 3) p *= dd;
    return (float)(1.0 / p);
}

// Largest n whose deposit row (n fp32) is accumulated in LDS.
constexpr int32_t ADJ_MAX_N = 32768;

// Slots per adjacency row: 2m rounded up to 16, so rows are 32-byte aligned
// and the gather's 16-slot loads never leave their row.
__host__ __device__ inline int32_t adj_pitch(int32_t m) { return (2 * m + 15) & ~15; }

// Largest n of a batch with the adjacency deposit (its scatter needs
// adj_stage_ants(n) > 0).
constexpr int32_t ADJ_BATCH_MAX_N = 8192;
// Rows per wave of the batched gather: four while 4 waves x 4 rows of nmax
// floats fit in 64 KB of LDS, else one.
__host__ __device__ inline int gather_rows(int32_t nmax) { return nmax <= 1024 ? 4 : 1; }

// Staged scatter: an LDS row holds a block's 2A slots of one city plus one
// padding dword (2A + 2 16-bit entries, an odd number of dwords), so the
// slots that one store instruction writes into rows of random cities fall
// into different LDS banks.
__host__ __device__ inline int32_t adj_stage_stride(int32_t A) { return 2 * A + 2; }
__host__ __device__ inline size_t adj_stage_lds(int32_t rows, int32_t A) {
    return (size_t)rows * adj_stage_stride(A) * sizeof(uint16_t);
}
// Ants per staged scatter block for rows of n cities: the largest A whose
// staging of every row fits in 160 KB of LDS; 0 beyond (row bands instead).
__host__ __device__ inline int32_t adj_stage_ants(int32_t n) {
    for (int32_t A = 32; A >= 4; A >>= 1)
        if (adj_stage_lds(n, A) <= 163840) return A;
    return 0;
}
// Larger n: blocks of ADJ_BAND_ANTS ants and ADJ_BAND_ROWS rows (136 KB of LDS).
constexpr int32_t ADJ_BAND_ANTS = 16;
constexpr int32_t ADJ_BAND_ROWS = 2048;

// ---------------------------------------------------------------------------
// Batched instances: all matrices of all instances live in shared buffers;
// InstMeta gives each instance's sizes and offsets.
struct InstMeta {
    int32_t n, m, P, words;
    int64_t mat_off;   // element offset of row 0 in the pitched n x P matrices
    int64_t nk_off;    // offset in the n x k arrays
    int64_t tour_off;  // offset in tours (m x n)
    int32_t ant_off;   // offset in per-ant arrays
    int32_t city_off;  // offset in per-city arrays (global row of city 0)
    int32_t dep_off;   // offset in the deposit lists (m + w + 2 per instance)
    int32_t pad;
    int64_t adj_off;   // offset in the adjacency slots (n x adj_pitch(m) per instance)
};

struct BatchBuffers {
    const InstMeta* meta;
    int32_t count, k, total_rows, total_ants;
    const int32_t* row_inst;  // global row -> instance
    const int32_t* ant_inst;  // global ant -> instance
    const int32_t* xy;        // concatenated coordinates, 2 per city
    int32_t* dist;
    float* tau_cur;
    float* tau_next;
    float* choice;
    int32_t* nn;
    uint16_t* nn16;
    float* choice_nn;
    float* eta_nn;
    const float* eta;
    int32_t eta_len;
    int32_t* tours;
    int64_t* lengths;
    int32_t* best_tours;     // per-city layout (city_off)
    int32_t* dep_ant;
    float* dep_val;
    uint16_t* adj;           // Ant System adjacency slots (null: atomic deposit)
    const int2* adj_items;   // staged scatter blocks (instance, first ant), adj_ants ants each
    int32_t nadj_items, adj_ants;
    DevState* state;         // one per instance
    const int2* items;       // construction work items (instance, first ant), longest first
    int32_t nitems;
    uint32_t* queue;         // QUEUE_BINS work-queue counters, reset by the statistics kernel
    const int32_t* bin_off;  // items of bin h: [bin_off[h], bin_off[h + 1])
    int64_t* cnn;            // per-instance nearest-neighbour tour length
    float* tau0;             // per-instance initial pheromone
    float alpha, one_minus_rho;
    int32_t eta_pow;         // as SweepArgs::eta_pow
    int32_t variant, w, ants_per_block, tau0_policy;
    int32_t ants_per_wave;   // construction: ants sharing one wave (1, 2 or 4)
    float tau0_const, rho;
    uint32_t seed_lo, seed_hi;
};

// Construction work-queue bins, one per XCD of the MI355X.
constexpr int32_t QUEUE_BINS = 8;

hipError_t launch_batch_init(const BatchBuffers& b, int32_t pmax, int32_t nmax, hipStream_t s);
hipError_t launch_batch_iteration_kernel(const BatchBuffers& b, int32_t which, int32_t nmax, int32_t grid,
                                         hipStream_t s);
size_t batch_construct_lds_bytes(int32_t nmax, int32_t k, int32_t ants_per_block, int32_t ants_per_wave);
int32_t batch_max_ants_per_wave(int32_t nmax, int32_t k);
hipError_t launch_batch_exchange(const BatchBuffers& b, hipStream_t s);
int32_t batch_construct_blocks_per_cu(int32_t nmax, int32_t ants_per_block, int32_t ants_per_wave,
                                      size_t lds_bytes);

hipError_t launch_dist_from_xy(const int32_t* xy, int32_t n, int32_t P, int32_t* dist, hipStream_t s);
// End stamp of a timed iteration: stamp[4] = now, then stamp[0..4] copied to
// host (coherent host memory, system fence).
hipError_t launch_stamp_end(uint64_t* stamp, int64_t* host, hipStream_t s);
// *flag |= 1 unless the sweep's computed heuristic 1 / d^eta_pow equals the
// table eta[0 .. eta_len) bit for bit.
hipError_t launch_eta_check(int32_t eta_pow, const float* eta, int32_t eta_len, int32_t* flag, hipStream_t s);
// out (pitched n x P) = the heuristic table's value of every edge's distance
hipError_t launch_eta_full(const int32_t* dist, const float* eta, int32_t eta_len, int32_t n, int32_t P, float* out,
                           hipStream_t s);
// Largest n x P for which aging colonies keep the heuristic of every edge.
constexpr size_t ETA_FULL_MAX = (size_t)1 << 26;
hipError_t launch_fix_diagonal(int32_t* dist, int32_t n, int32_t P, hipStream_t s);
hipError_t launch_dist_max(const int32_t* dist, int32_t n, int32_t P, int32_t* out, hipStream_t s);
hipError_t launch_nn_lists(const int32_t* dist, int32_t n, int32_t P, int32_t k, int32_t* nn, hipStream_t s);
hipError_t launch_fill_rows(float* a, int32_t n, int32_t P, float v, hipStream_t s);
// nn: n x k candidate lists (k < n) to shortcut the steps, or null.
hipError_t launch_greedy(const int32_t* dist, int32_t n, int32_t P, int32_t start, int64_t* out, const int32_t* nn,
                         int32_t k, hipStream_t s);
// Timing of one phase's launches: its first kernel records `start` and its
// last kernel `stop` from their own dispatches (hipExtLaunchKernel), which
// adds no marker packets to the stream; both null: plain launches.
struct LaunchTiming {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
    uint64_t* stamp = nullptr;  // the phase's first kernel stores its start time (wall_clock64) here
};

// Start stamp of a phase's first kernel: block 0's first thread, before any
// early exit (a plain store; the next kernel on the stream sees it).
__device__ __forceinline__ void stamp_start(uint64_t* stamp) {
    if (stamp != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *stamp = wall_clock64();
}

hipError_t launch_sweep(const SweepArgs& a, hipStream_t s, const LaunchTiming& t = LaunchTiming());
// Full-row construction: chunks of 64 candidates whose list entries and
// weights are loaded together; LDS words per ant (visited bitmap, the weight
// row in city order, then per candidate its weight, prefix and city, then
// the running total before each chunk and after the last).
constexpr int FULL_GROUP = 16;
// word offset of the weight row (16-byte aligned: per-ant areas are multiples of 4 words)
__host__ __device__ inline size_t full_row_offset(int32_t words) { return ((size_t)words + 3) & ~(size_t)3; }
__host__ __device__ inline size_t full_lds_words(int32_t n, int32_t words) {
    const size_t nchunk = ((size_t)n + 63) / 64;
    return ((full_row_offset(words) + 4 * nchunk * 64 + nchunk + 1) + 3) & ~(size_t)3;
}
// Construction paths: 0 = candidate lists with global reads, 1 = full row,
// 2 = candidate lists staged in LDS (n*k*6 bytes + per-ant state per block).
size_t staged_lds_bytes(int32_t n, int32_t k, int32_t words, int32_t ants_per_block);
hipError_t launch_construct(const ConstructArgs& a, int32_t path, int32_t ants_per_block, hipStream_t s,
                            const LaunchTiming& t = LaunchTiming());
hipError_t launch_nn_aux(const int32_t* nn, const int32_t* dist, const float* eta, int32_t eta_len, int32_t n,
                         int32_t P, int32_t k, uint16_t* nn16, float* eta_nn, hipStream_t s);
hipError_t launch_stats(const StatsArgs& a, hipStream_t s, const LaunchTiming& t = LaunchTiming());
hipError_t launch_deposit(const DepositArgs& a, int32_t grid, hipStream_t s,
                          const LaunchTiming& t = LaunchTiming());
// Rank-based statistics and deposit as one launch (stats_deposit_fusable):
// the results of launch_stats followed by launch_deposit_rows.  s.stamp, when
// set, receives the start time in s.stamp[0] and, in s.stamp[1], the time
// block 0 has its deposit list (the statistics phase's end).
bool stats_deposit_fusable(int32_t variant, int32_t n, int32_t m, int32_t w);
hipError_t launch_stats_deposit(const StatsArgs& s, const DepositArgs& d, hipStream_t st,
                                const LaunchTiming& t = LaunchTiming());
// Rank-based deposit (list entries in order, one writer per cell, no
// atomics): the same fp32 sums on every launch.
hipError_t launch_deposit_rows(const DepositArgs& a, hipStream_t s, const LaunchTiming& t = LaunchTiming());
// Ant System deposit without global atomics: tours scattered into adj, then
// every pheromone row gathers its deposits in LDS (needs a.adj, a.m, n <= ADJ_MAX_N).
hipError_t launch_deposit_adj(const DepositArgs& a, hipStream_t s, const LaunchTiming& t = LaunchTiming());
hipError_t launch_copy(const float4* src, float4* dst, int64_t count, hipStream_t s);
hipError_t launch_probe_step(int32_t iters, int32_t k, unsigned long long* out, int32_t* sink, hipStream_t s);

}  // namespace aco
