// C ABI of the MI355X ACO solver (include/acogpu.h): context lifetime, device
// buffers, the per-iteration kernel sequence and its hipGraph replay, parity
// taps.  All device memory is allocated in aco_create; an iteration allocates
// nothing and performs no host synchronisation.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <vector>

#include "../../include/acogpu.h"
#include "aco_kernels.h"
#include "aco_prior.h"

using aco::DevState;

struct aco_ctx {
    aco_params p;
    int32_t n = 0, m = 0, k = 0, P = 0, words = 0, apb = 1;
    int32_t path = 0;          // construction kernel: 0 global, 1 full row, 2 LDS-staged
    int32_t young_blocks = 0;  // staged path: blocks reserved for young aging ants
    bool fused = false;        // whole iterations launch statistics + deposit as one kernel
    int32_t max_lds = 0;       // LDS bytes per block on this device
    int32_t nn_cols = 0;  // k (candidate) or n (full row)
    hipStream_t stream = nullptr;
    int32_t* d_dist = nullptr;
    float* d_tau[2] = {nullptr, nullptr};
    float* d_choice = nullptr;
    int32_t* d_xy = nullptr;  // coordinates, padded to a multiple of four cities (sweep distances)
    float* d_choice_nn = nullptr;
    float* d_tau_nn = nullptr;
    float* d_eta_nn = nullptr;
    int32_t* d_nn = nullptr;
    uint16_t* d_nn16 = nullptr;
    float* d_eta = nullptr;
    float* d_eta_full = nullptr;  // aging: heuristic of every edge, pitched (young ants' fallback rows)
    int32_t eta_len = 0;
    int32_t eta_pow = 0;  // aco::eta_integer_power of (beta, largest distance)
    int32_t sweep_eta_pow = 0;  // eta_pow when the sweep's own reciprocal matches the table, else 0
    float* d_mean = nullptr;
    int32_t* d_tours = nullptr;
    int64_t* d_len = nullptr;
    int64_t* d_prev_len = nullptr;
    int32_t* d_age = nullptr;
    int32_t* d_perf = nullptr;
    uint8_t* d_valid = nullptr;
    int32_t* d_best_tour = nullptr;
    int32_t* d_dep_ant = nullptr;
    uint16_t* d_adj = nullptr;  // Ant System adjacency slots, n x adj_pitch(m) (null: atomic deposit)
    uint16_t* d_top = nullptr;  // row top lists, n x TOP_K (null: none)
    float* d_top_thr = nullptr; // per row: TOP_K-th weight of the previous sweep
    float* d_dep_val = nullptr;
    DevState* d_state = nullptr;
    int64_t* d_scratch64 = nullptr;
    int32_t cur = 0;           // index of the current pheromone buffer
    int32_t next_phase = 0;    // aco_phase ordering
    bool initialized = false;
    hipGraphExec_t graph = nullptr;       // two iterations starting at cur == 0
    hipGraphExec_t graph_long = nullptr;  // kGraphLong iterations starting at cur == 0
    hipEvent_t ev[6] = {};
    hipEvent_t pev[6] = {};     // asynchronous phases: iteration start, then the end of phase p at 1 + p
    uint32_t phases_async = 0;  // bit p: phase p enqueued without waiting since the iteration began
    hipEvent_t tev_end = nullptr;  // end of a timed iteration (phase times come from in-kernel stamps)
    int64_t* h_stats = nullptr; // coherent pinned host memory: {iteration best, best so far, valid ants}
    int64_t* h_stamp = nullptr; // coherent pinned host memory: a timed iteration's five wall-clock stamps
    uint64_t* d_stamp = nullptr;  // the phases' start stamps and the end stamp (device)
    double wall_khz = 0.0;      // wall_clock64 ticks per millisecond
    char err[512] = {0};
};

static thread_local char g_create_err[512] = "no error";

static int fail(aco_ctx* c, int code, const char* fmt, ...) {
    char* buf = c ? c->err : g_create_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(ctx, expr)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail((ctx), ACO_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                     \
    } while (0)

extern "C" const char* aco_version(void) { return "acogpu_amd 0.1 (gfx950)"; }

extern "C" const char* aco_last_error(const aco_ctx* c) { return c ? c->err : g_create_err; }

extern "C" int aco_params_init(aco_params* p) {
    if (!p) return ACO_E_ARG;
    std::memset(p, 0, sizeof(*p));
    p->variant = ACO_VARIANT_AS;
    p->mode = ACO_MODE_CANDIDATE;
    p->m = 10;           // DEFAULT_NUM_ANTS
    p->k = 20;
    p->w = 5;            // DEFAULT_NUM_RANK_ANTS
    p->alpha = 1.0f;     // DEFAULT_ALPHA
    p->beta = 5.0f;      // DEFAULT_BETA
    p->rho = 0.1f;       // DEFAULT_RHO
    p->tau0_policy = ACO_TAU0_CONST;
    p->tau0 = 1e-6f;     // DEFAULT_PHEROMONE
    p->min_age = 30;     // DEFAULT_MIN_RITIREMENT_AGE
    p->max_age = 50;     // DEFAULT_MAX_RITIREMENT_AGE
    p->local_perf = 10;  // DEFAULT_LOCAL_NO_CONVERGENCE
    p->global_perf = 10; // DEFAULT_GLOBAL_NO_CONVERGENCE
    p->min_pop = 40;     // DEFAULT_MIN_POPULATION_LIMIT
    p->max_pop = 100;    // DEFAULT_MAX_POPULATION_LIMIT
    p->seed = 12345678;  // RANDOM_SEED (inc/Colony.h:15)
    return ACO_OK;
}

static void free_all(aco_ctx* c) {
    if (c->graph) hipGraphExecDestroy(c->graph);
    if (c->graph_long) hipGraphExecDestroy(c->graph_long);
    void* ptrs[] = {c->d_dist, c->d_tau[0], c->d_tau[1], c->d_choice, c->d_xy, c->d_choice_nn, c->d_nn, c->d_eta,
                    c->d_tau_nn, c->d_eta_nn, c->d_nn16,
                    c->d_mean, c->d_tours, c->d_len, c->d_prev_len, c->d_age, c->d_perf, c->d_valid,
                    c->d_best_tour, c->d_dep_ant, c->d_dep_val, c->d_state, c->d_scratch64, c->d_adj,
                    c->d_top, c->d_top_thr, c->d_eta_full};
    for (void* q : ptrs)
        if (q) hipFree(q);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    for (auto& e : c->pev)
        if (e) hipEventDestroy(e);
    if (c->tev_end) hipEventDestroy(c->tev_end);
    if (c->h_stats) hipHostFree(c->h_stats);
    if (c->h_stamp) hipHostFree(c->h_stamp);
    if (c->d_stamp) hipFree(c->d_stamp);
    if (c->stream) hipStreamDestroy(c->stream);
}

template <typename T>
static int dalloc(aco_ctx* c, T** out, size_t count) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(out), count * sizeof(T) + 16);
    if (e != hipSuccess) {
        *out = nullptr;
        (void)hipGetLastError();
        return fail(c, ACO_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    return ACO_OK;
}

static int check_params(const aco_params* p) {
    if (p->n < 2) return fail(nullptr, ACO_E_ARG, "n must be >= 2 (got %d)", p->n);
    if (p->m < 1) return fail(nullptr, ACO_E_ARG, "m must be >= 1 (got %d)", p->m);
    if (p->variant != ACO_VARIANT_AS && p->variant != ACO_VARIANT_RBAS)
        return fail(nullptr, ACO_E_ARG, "unknown variant %d", p->variant);
    if (p->mode == ACO_MODE_CANDIDATE) {
        if (p->k < 1 || p->k > 64 || p->k > p->n - 1)
            return fail(nullptr, ACO_E_ARG, "candidate list length k=%d must be in [1, min(64, n-1)]", p->k);
    } else if (p->mode == ACO_MODE_FULLROW) {
        if (p->n > 8192) return fail(nullptr, ACO_E_RANGE, "full-row mode supports n <= 8192 (got %d)", p->n);
    } else {
        return fail(nullptr, ACO_E_ARG, "unknown mode %d", p->mode);
    }
    if (!(p->rho >= 0.0f && p->rho <= 1.0f)) return fail(nullptr, ACO_E_ARG, "rho must be in [0,1]");
    if (!(p->alpha >= 0.0f) || !(p->beta >= 0.0f)) return fail(nullptr, ACO_E_ARG, "alpha, beta must be >= 0");
    if (p->variant == ACO_VARIANT_RBAS && p->w < 1) return fail(nullptr, ACO_E_ARG, "w must be >= 1");
    if (p->aging && (p->min_age < 1 || p->max_age < 0))
        return fail(nullptr, ACO_E_ARG, "aging needs min_age >= 1 and max_age >= 0");
    if (p->ants_per_block < 0 || p->ants_per_block > 8)
        return fail(nullptr, ACO_E_ARG, "ants_per_block must be in [0, 8]");
    if (p->construct_path < 0 || p->construct_path > 2)
        return fail(nullptr, ACO_E_ARG, "construct_path must be 0 (auto), 1 (global) or 2 (LDS-staged)");
    if (p->deposit_path < 0 || p->deposit_path > 2)
        return fail(nullptr, ACO_E_ARG, "deposit_path must be 0 (auto), 1 (atomics) or 2 (adjacency gather)");
    if (p->deposit_path == 2 && (p->variant != ACO_VARIANT_AS || p->n > aco::ADJ_MAX_N))
        return fail(nullptr, ACO_E_ARG, "the adjacency deposit needs the Ant System and n <= %d", aco::ADJ_MAX_N);
    return ACO_OK;
}

extern "C" int aco_create(const aco_params* p, const int32_t* dist, const int32_t* xy, aco_ctx** out) {
    if (!p || !out) return fail(nullptr, ACO_E_ARG, "null argument");
    *out = nullptr;
    if ((dist == nullptr) == (xy == nullptr))
        return fail(nullptr, ACO_E_ARG, "exactly one of dist and xy must be given");
    int rc = check_params(p);
    if (rc) return rc;

    aco_ctx* c = new aco_ctx();
    c->p = *p;
    // RankBasedAntSystem::computeParameters clamps w to n (src/RankBasedAntSystem.cu:35-40).
    if (c->p.variant == ACO_VARIANT_RBAS && c->p.w > c->p.n) c->p.w = c->p.n;
    c->n = p->n;
    c->m = p->m;
    c->fused = c->p.variant == ACO_VARIANT_RBAS && aco::stats_deposit_fusable(c->p.variant, c->n, c->m, c->p.w);
    c->k = p->mode == ACO_MODE_CANDIDATE ? p->k : 0;
    c->nn_cols = p->mode == ACO_MODE_CANDIDATE ? p->k : p->n;
    c->P = (c->n + 63) & ~63;
    c->words = (c->n + 31) / 32;
    {
        int dev = 0, max_lds = 65536;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            max_lds = 65536;
        (void)hipGetLastError();
        c->max_lds = max_lds;
        max_lds -= 256;  // static LDS of the kernels
        if (p->mode == ACO_MODE_FULLROW) {
            c->path = 1;
            c->apb = p->ants_per_block ? p->ants_per_block : 1;
            const size_t per_ant = aco::full_lds_words(c->n, c->words) * 4;
            while (c->apb > 1 && per_ant * c->apb > (size_t)max_lds) --c->apb;
            if (per_ant > (size_t)max_lds) {
                int rc2 = fail(nullptr, ACO_E_RANGE, "full-row mode needs %zu bytes of LDS per ant (limit %d)", per_ant,
                               max_lds);
                delete c;
                return rc2;
            }
        } else {
            // Staged path: one block per CU (the instance fills most of the
            // LDS); choose the ants per block so that the normal blocks plus
            // the blocks reserved for young aging ants fit the CUs in one pass.
            int cus = 256;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
            (void)hipGetLastError();
            const int32_t young = p->aging ? std::min(c->m, std::max(p->min_age, 1)) : 0;
            // Measured on MI355X (n=1000, m=1000, k=20): 8 ants per block is
            // fastest; fewer only when the colony is too small to fill the CUs.
            int32_t want = p->ants_per_block;
            if (!want) {
                want = 8;
                while (want > 1 && (c->m + want - 1) / want + young < cus / 2) --want;
            }
            int32_t W = want;
            while (W > 1 && aco::staged_lds_bytes(c->n, c->k, c->words, W) > (size_t)max_lds) --W;
            const bool fits = c->n <= 65535 && aco::staged_lds_bytes(c->n, c->k, c->words, W) <= (size_t)max_lds;
            if (p->construct_path == 2 && !fits) {
                int rc2 = fail(nullptr, ACO_E_RANGE, "LDS-staged construction needs %zu bytes of LDS (limit %d)",
                               aco::staged_lds_bytes(c->n, c->k, c->words, 1), max_lds);
                delete c;
                return rc2;
            }
            if (fits && p->construct_path != 1) {
                c->path = 2;
                c->apb = W;
                c->young_blocks = p->aging ? std::min(c->m, std::max(p->min_age, 1)) : 0;
            } else {
                // global-memory candidate rows: ants per block up to 4 as the
                // colony fills the CUs, one block per CU and one ant per SIMD
                // at m = 4 x CUs (n = 20000, m = 1024: construction 7.8 ->
                // 7.65 ms against one ant per block)
                c->path = 0;
                c->apb = p->ants_per_block ? p->ants_per_block : std::min(4, std::max(1, c->m / cus));
            }
        }
    }

    auto bail = [&](int code) {
        free_all(c);
        if (code != ACO_OK && c->err[0]) std::snprintf(g_create_err, sizeof(g_create_err), "%s", c->err);
        delete c;
        return code;
    };
#define TRYB(expr)                     \
    do {                               \
        int r_ = (expr);               \
        if (r_ != ACO_OK) return bail(r_); \
    } while (0)
#define HIPB(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return bail(fail(c, ACO_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)));          \
    } while (0)

    HIPB(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) HIPB(hipEventCreate(&e));
    for (auto& e : c->pev) HIPB(hipEventCreate(&e));
    HIPB(hipEventCreate(&c->tev_end));
    HIPB(hipHostMalloc(reinterpret_cast<void**>(&c->h_stats), 4 * sizeof(int64_t), hipHostMallocCoherent));
    HIPB(hipHostMalloc(reinterpret_cast<void**>(&c->h_stamp), 8 * sizeof(int64_t), hipHostMallocCoherent));
    std::memset(c->h_stamp, 0, 8 * sizeof(int64_t));
    {
        int dev = 0, khz = 0;
        HIPB(hipGetDevice(&dev));
        HIPB(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) return bail(fail(c, ACO_E_HIP, "device reports no wall-clock rate"));
        c->wall_khz = (double)khz;
    }
    std::memset(c->h_stats, 0, 4 * sizeof(int64_t));
    const size_t nP = (size_t)c->n * c->P;
    TRYB(dalloc(c, &c->d_dist, nP));
    TRYB(dalloc(c, &c->d_tau[0], nP));
    TRYB(dalloc(c, &c->d_tau[1], nP));
    TRYB(dalloc(c, &c->d_choice, nP));
    TRYB(dalloc(c, &c->d_nn, (size_t)c->n * c->nn_cols));
    TRYB(dalloc(c, &c->d_choice_nn, (size_t)c->n * (c->k > 0 ? c->k : 1)));
    TRYB(dalloc(c, &c->d_tau_nn, (size_t)c->n * (c->k > 0 ? c->k : 1)));
    TRYB(dalloc(c, &c->d_eta_nn, (size_t)c->n * (c->k > 0 ? c->k : 1)));
    TRYB(dalloc(c, &c->d_nn16, (size_t)c->n * (c->k > 0 ? c->k : 1)));
    TRYB(dalloc(c, &c->d_mean, (size_t)c->n));
    TRYB(dalloc(c, &c->d_tours, (size_t)c->m * c->n));
    TRYB(dalloc(c, &c->d_len, (size_t)c->m));
    TRYB(dalloc(c, &c->d_prev_len, (size_t)c->m));
    TRYB(dalloc(c, &c->d_age, (size_t)c->m));
    TRYB(dalloc(c, &c->d_perf, (size_t)c->m));
    TRYB(dalloc(c, &c->d_valid, (size_t)c->m));
    TRYB(dalloc(c, &c->d_best_tour, (size_t)c->n));
    TRYB(dalloc(c, &c->d_dep_ant, (size_t)c->m + c->p.w + 2));
    TRYB(dalloc(c, &c->d_dep_val, (size_t)c->m + c->p.w + 2));
    if (c->p.variant == ACO_VARIANT_AS && c->n <= aco::ADJ_MAX_N && c->p.deposit_path != 1)
        TRYB(dalloc(c, &c->d_adj, (size_t)c->n * aco::adj_pitch(c->m)));
    // Row top lists for large candidate-mode instances (alpha = 1: the
    // sweep's lower bound of the TOP_K-th weight holds for tau times eta)
    if (c->p.mode == ACO_MODE_CANDIDATE && c->n >= aco::TOP_MIN_N && c->n <= 65535 && c->p.alpha == 1.0f) {
        TRYB(dalloc(c, &c->d_top, (size_t)c->n * aco::TOP_K));
        TRYB(dalloc(c, &c->d_top_thr, (size_t)c->n));
    }
    TRYB(dalloc(c, &c->d_state, 1));
    TRYB(dalloc(c, &c->d_scratch64, 4));
    TRYB(dalloc(c, &c->d_stamp, 8));
    HIPB(hipMemsetAsync(c->d_dist, 0, nP * sizeof(int32_t), c->stream));
    HIPB(hipMemsetAsync(c->d_tau[0], 0, nP * sizeof(float), c->stream));
    HIPB(hipMemsetAsync(c->d_tau[1], 0, nP * sizeof(float), c->stream));
    HIPB(hipMemsetAsync(c->d_choice, 0, nP * sizeof(float), c->stream));
    HIPB(hipMemsetAsync(c->d_tours, 0, (size_t)c->m * c->n * sizeof(int32_t), c->stream));
    HIPB(hipMemsetAsync(c->d_len, 0, (size_t)c->m * sizeof(int64_t), c->stream));

    // Distance matrix.
    if (xy) {
        for (int32_t i = 0; i < 2 * c->n; ++i)
            if (xy[i] >= (1 << 29) || xy[i] <= -(1 << 29))
                return bail(fail(c, ACO_E_RANGE, "coordinate %d out of range (|x| < 2^29)", xy[i]));
        // Coordinates below XY_SWEEP_LIMIT in magnitude stay on the device:
        // the sweep computes its rows' distances from them (exactly the
        // matrix's values) instead of reading the n^2 matrix every iteration.
        bool keep = true;
        for (int32_t i = 0; i < 2 * c->n; ++i)
            keep = keep && xy[i] < aco::XY_SWEEP_LIMIT && xy[i] > -aco::XY_SWEEP_LIMIT;
        const size_t padded = (size_t)2 * (((size_t)c->n + 3) & ~(size_t)3);
        int32_t* d_xy = nullptr;
        TRYB(dalloc(c, &d_xy, padded));
        hipError_t e = hipMemsetAsync(d_xy, 0, sizeof(int32_t) * padded, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_xy, xy, sizeof(int32_t) * 2 * c->n, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = aco::launch_dist_from_xy(d_xy, c->n, c->P, c->d_dist, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (keep && e == hipSuccess)
            c->d_xy = d_xy;
        else
            hipFree(d_xy);
        HIPB(e);
    } else {
        // The reference passes a device pointer in its GPU build and a host
        // pointer in its CPU build (src/Setup.cpp:54-70); accept both.
        hipPointerAttribute_t attr;
        hipMemcpyKind kind = hipMemcpyHostToDevice;
        if (hipPointerGetAttributes(&attr, dist) == hipSuccess && attr.type == hipMemoryTypeDevice)
            kind = hipMemcpyDeviceToDevice;
        (void)hipGetLastError();
        HIPB(hipMemcpy2DAsync(c->d_dist, (size_t)c->P * 4, dist, (size_t)c->n * 4, (size_t)c->n * 4, c->n, kind,
                              c->stream));
        HIPB(aco::launch_fix_diagonal(c->d_dist, c->n, c->P, c->stream));
    }

    // Heuristic table eta^beta by integer distance (aco::eta_integer_power).
    int32_t* d_max = reinterpret_cast<int32_t*>(c->d_scratch64);
    HIPB(hipMemsetAsync(d_max, 0, sizeof(int32_t), c->stream));
    HIPB(aco::launch_dist_max(c->d_dist, c->n, c->P, d_max, c->stream));
    int32_t dmax = 0;
    HIPB(hipMemcpyAsync(&dmax, d_max, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPB(hipStreamSynchronize(c->stream));
    if (dmax < 0 || dmax > (1 << 26))
        return bail(fail(c, ACO_E_RANGE, "largest distance %d exceeds the heuristic table limit 2^26", dmax));
    c->eta_len = dmax + 1;
    std::vector<float> eta((size_t)c->eta_len);
    c->eta_pow = aco::eta_integer_power((double)c->p.beta, dmax);
    if (c->eta_pow) {
        for (int32_t d = 0; d <= dmax; ++d) eta[d] = aco::eta_power_value(c->eta_pow, d);
    } else {
        eta[0] = (float)std::pow(2.0, (double)c->p.beta);  // d == 0: duplicate cities, taken as d = 1/2
        for (int32_t d = 1; d <= dmax; ++d) eta[d] = (float)std::pow(1.0 / (double)d, (double)c->p.beta);
    }
    TRYB(dalloc(c, &c->d_eta, (size_t)c->eta_len));
    HIPB(hipMemcpy(c->d_eta, eta.data(), sizeof(float) * eta.size(), hipMemcpyHostToDevice));
    // The sweep computes 1 / d^beta with its own reciprocal sequence; it is
    // used only if it reproduces every table entry.
    if (c->eta_pow) {
        int32_t* d_flag = reinterpret_cast<int32_t*>(c->d_scratch64);
        int32_t flag = 1;
        HIPB(hipMemsetAsync(d_flag, 0, sizeof(int32_t), c->stream));
        HIPB(aco::launch_eta_check(c->eta_pow, c->d_eta, c->eta_len, d_flag, c->stream));
        HIPB(hipMemcpyAsync(&flag, d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPB(hipStreamSynchronize(c->stream));
        c->sweep_eta_pow = flag == 0 ? c->eta_pow : 0;
    }
    // Aging colonies: young ants weigh a fallback row's every city with the
    // heuristic, a gather per city from the table; a pitched copy of the
    // table's values turns that into row loads (one memory round trip)
    if (c->p.aging && (size_t)c->n * c->P <= aco::ETA_FULL_MAX) {
        TRYB(dalloc(c, &c->d_eta_full, (size_t)c->n * c->P));
        HIPB(aco::launch_eta_full(c->d_dist, c->d_eta, c->eta_len, c->n, c->P, c->d_eta_full, c->stream));
    }
    HIPB(hipDeviceSynchronize());
    *out = c;
    return ACO_OK;
#undef TRYB
#undef HIPB
}

extern "C" int aco_destroy(aco_ctx* c) {
    if (!c) return ACO_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return ACO_OK;
}

static int reset_state(aco_ctx* c) {
    DevState st{};
    st.iter_best = (int64_t)0x7fffffff - 1;  // iterBestDist = MAX_DISTANCE - 1 (src/Colony.cu:38)
    st.glob_best = (int64_t)0x7fffffff;      // globBestDist = MAX_DISTANCE (src/Colony.cu:39)
    st.iter_best_ant = -1;
    st.gperf = 0;  // globalperformance, uninitialised in the reference (inc/Colony.h:114)
    st.iter = 0;
    st.num_valid = c->m;  // intialpopulation = numAnts (src/Colony.cu:132)
    st.converged = 0;
    st.ndep = 0;
    st.stop_at = -1;
    st.halted = 0;
    st.skip = 0;
    HIP_TRY(c, hipMemcpyAsync(c->d_state, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
    std::vector<uint8_t> ones((size_t)c->m, 1);
    HIP_TRY(c, hipMemcpyAsync(c->d_valid, ones.data(), (size_t)c->m, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_age, 0, sizeof(int32_t) * c->m, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_perf, 0, sizeof(int32_t) * c->m, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_prev_len, 0, sizeof(int64_t) * c->m, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_len, 0, sizeof(int64_t) * c->m, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_best_tour, 0, sizeof(int32_t) * c->n, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ACO_OK;
}

extern "C" int aco_initialize(aco_ctx* c) {
    if (!c) return ACO_E_ARG;
    // Colony::computeParameters -> computeInitialPheromone (src/AntSystem.cu:194-203,
    // src/RankBasedAntSystem.cu:353-363).
    float tau0 = c->p.tau0;
    HIP_TRY(c, aco::launch_nn_lists(c->d_dist, c->n, c->P, c->nn_cols, c->d_nn, c->stream));
    if (c->p.tau0_policy != ACO_TAU0_CONST) {
        const bool lists = c->nn_cols < c->n;
        HIP_TRY(c, aco::launch_greedy(c->d_dist, c->n, c->P, 0, c->d_scratch64, lists ? c->d_nn : nullptr,
                                      lists ? c->nn_cols : 0, c->stream));
        int64_t cnn = 0;
        HIP_TRY(c, hipMemcpyAsync(&cnn, c->d_scratch64, sizeof(cnn), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (cnn <= 0) return fail(c, ACO_E_STATE, "nearest-neighbour tour length is %lld", (long long)cnn);
        if (c->p.tau0_policy == ACO_TAU0_AS)
            tau0 = (float)((double)c->m / (double)cnn);
        else
            tau0 = (float)(0.5 * c->p.w * (c->p.w - 1) / ((double)c->p.rho * (double)cnn));
        c->p.tau0 = tau0;
    }
    c->cur = 0;
    c->next_phase = ACO_PHASE_PROBABILITIES;
    HIP_TRY(c, aco::launch_fill_rows(c->d_tau[0], c->n, c->P, tau0, c->stream));
    HIP_TRY(c, aco::launch_fill_rows(c->d_tau[1], c->n, c->P, 0.0f, c->stream));
    if (c->d_top_thr) HIP_TRY(c, hipMemsetAsync(c->d_top_thr, 0, sizeof(float) * c->n, c->stream));
    if (c->k > 0)
        HIP_TRY(c, aco::launch_nn_aux(c->d_nn, c->d_dist, c->d_eta, c->eta_len, c->n, c->P, c->k, c->d_nn16,
                                      c->d_eta_nn, c->stream));
    int rc = reset_state(c);
    if (rc) return rc;
    // the iteration graphs capture buffer addresses and create-time
    // parameters only: they stay valid across initialisations
    c->initialized = true;
    return ACO_OK;
}

// The kernels of one iteration, one function per phase, on the ctx stream;
// they read pheromone buffer `cur` and produce buffer 1-cur.
static hipError_t enqueue_sweep(aco_ctx* c, int cur, const aco::LaunchTiming& t) {
    hipError_t e;
    float* tau = c->d_tau[cur];
    float* tau_next = c->d_tau[1 - cur];
    const bool cand = c->p.mode == ACO_MODE_CANDIDATE;
    (void)tau;
    (void)tau_next;
    (void)cand;
    aco::SweepArgs sa{};
    sa.tau = tau;
    sa.dist = c->d_dist;
    sa.eta = c->d_eta;
    sa.eta_full = c->d_eta_full;
    sa.xy = c->d_xy;
    sa.eta_len = c->eta_len;
    sa.nn = cand ? c->d_nn : nullptr;
    sa.k = c->k;
    sa.choice = c->d_choice;
    sa.choice_nn = c->d_choice_nn;
    sa.eta_nn = c->d_eta_nn;
    sa.tau_nn = (cand && c->p.aging) ? c->d_tau_nn : nullptr;
    sa.tau_next = tau_next;
    sa.mean = c->d_mean;
    sa.age = c->p.aging ? c->d_age : nullptr;
    sa.m = c->m;
    sa.n = c->n;
    sa.P = c->P;
    sa.alpha = c->p.alpha;
    sa.one_minus_rho = 1.0f - c->p.rho;
    sa.want_mean = c->p.aging;
    sa.eta_pow = c->sweep_eta_pow;
    sa.state = c->d_state;
    sa.stamp = t.stamp;
    sa.top = c->d_top;
    sa.top_thr = c->d_top_thr;
    if ((e = aco::launch_sweep(sa, c->stream, t)) != hipSuccess) return e;
    return hipSuccess;
}

static hipError_t enqueue_construct(aco_ctx* c, int cur, const aco::LaunchTiming& t) {
    hipError_t e;
    float* tau = c->d_tau[cur];
    float* tau_next = c->d_tau[1 - cur];
    const bool cand = c->p.mode == ACO_MODE_CANDIDATE;
    (void)tau;
    (void)tau_next;
    (void)cand;
    aco::ConstructArgs ca{};
    ca.nn = c->d_nn;
    ca.nn16 = c->n <= 65535 ? c->d_nn16 : nullptr;
    ca.choice_nn = c->d_choice_nn;
    ca.tau_nn = c->d_tau_nn;
    ca.eta_nn = c->d_eta_nn;
    ca.choice = c->d_choice;
    ca.dist = c->d_dist;
    ca.tau = tau;
    ca.mean = c->d_mean;
    ca.eta = c->d_eta;
    ca.eta_full = c->d_eta_full;
    ca.eta_len = c->eta_len;
    ca.tours = c->d_tours;
    ca.lengths = c->d_len;
    ca.age = c->d_age;
    ca.valid = c->d_valid;
    ca.state = c->d_state;
    ca.stamp = t.stamp;
    ca.n = c->n;
    ca.P = c->P;
    ca.m = c->m;
    ca.k = c->k;
    ca.aging = c->p.aging;
    ca.min_age = c->p.min_age;
    ca.alpha = c->p.alpha;
    ca.seed_lo = (uint32_t)c->p.seed;
    ca.seed_hi = (uint32_t)(c->p.seed >> 32);
    ca.stream = c->p.stream_id;
    ca.words = c->words;
    ca.young_blocks = c->young_blocks;
    ca.top = c->d_top;
    if ((e = aco::launch_construct(ca, c->path, c->apb, c->stream, t)) != hipSuccess) return e;
    return hipSuccess;
}

static aco::StatsArgs stats_args(aco_ctx* c, bool mirror, const aco::LaunchTiming& t) {
    aco::StatsArgs st{};
    st.tours = c->d_tours;
    st.lengths = c->d_len;
    st.valid = c->d_valid;
    st.age = c->d_age;
    st.perf = c->d_perf;
    st.prev_len = c->d_prev_len;
    st.best_tour = c->d_best_tour;
    st.dep_ant = c->d_dep_ant;
    st.dep_val = c->d_dep_val;
    st.state = c->d_state;
    st.stamp = t.stamp;
    st.n = c->n;
    st.m = c->m;
    st.variant = c->p.variant;
    st.w = c->p.w;
    st.aging = c->p.aging;
    st.max_age = c->p.max_age;
    st.global_perf = c->p.global_perf;
    st.min_pop = c->p.min_pop;
    st.host_stats = mirror ? c->h_stats : nullptr;
    return st;
}

static aco::DepositArgs deposit_args(aco_ctx* c, int cur, const aco::LaunchTiming& t) {
    aco::DepositArgs da{};
    da.tours = c->d_tours;
    da.best_tour = c->d_best_tour;
    da.dep_ant = c->d_dep_ant;
    da.dep_val = c->d_dep_val;
    da.state = c->d_state;
    da.stamp = t.stamp;
    da.tau_next = c->d_tau[1 - cur];
    da.n = c->n;
    da.P = c->P;
    da.dense = c->p.variant == ACO_VARIANT_AS;
    da.adj = c->d_adj;
    da.m = c->m;
    da.list_cap = c->m + c->p.w + 2;
    return da;
}

static hipError_t enqueue_stats(aco_ctx* c, bool mirror, const aco::LaunchTiming& t) {
    return aco::launch_stats(stats_args(c, mirror, t), c->stream, t);
}

// Ant System: the adjacency deposit where its buffer exists, else fp32
// atomics; rank-based: the ordered row deposit.
static hipError_t enqueue_deposit(aco_ctx* c, int cur, const aco::LaunchTiming& t) {
    const aco::DepositArgs da = deposit_args(c, cur, t);
    if (da.dense && c->d_adj) return aco::launch_deposit_adj(da, c->stream, t);
    if (!da.dense) return aco::launch_deposit_rows(da, c->stream, t);
    return aco::launch_deposit(da, c->m, c->stream, t);
}

// Rank-based colonies within aco::stats_deposit_fusable: statistics and
// deposit as one launch (stamping both phases' slots: the statistics from
// the kernel's start to block 0's deposit list, the deposit from there).
static hipError_t enqueue_stats_deposit(aco_ctx* c, int cur, bool mirror, const aco::LaunchTiming& t) {
    return aco::launch_stats_deposit(stats_args(c, mirror, t), deposit_args(c, cur, aco::LaunchTiming()),
                                     c->stream, t);
}

// mirror: the statistics kernel also writes its results to h_stats; t: the
// phase's first kernel records t.start and its last t.stop (may be null)
static hipError_t enqueue_phase(aco_ctx* c, int cur, int phase, bool mirror = false,
                                const aco::LaunchTiming& t = aco::LaunchTiming()) {
    switch (phase) {
        case ACO_PHASE_PROBABILITIES: return enqueue_sweep(c, cur, t);
        case ACO_PHASE_TOUR: return enqueue_construct(c, cur, t);
        case ACO_PHASE_ANT_DISTANCE: return enqueue_stats(c, mirror, t);
        case ACO_PHASE_UPDATE_PHEROMONE: return enqueue_deposit(c, cur, t);
        default: return hipSuccess;  // evaporation is part of the probability sweep
    }
}

// A whole iteration; ev (may be null) brackets the four kernels.
static hipError_t enqueue_iteration(aco_ctx* c, int cur, hipEvent_t* ev) {
    static const int order[4] = {ACO_PHASE_PROBABILITIES, ACO_PHASE_TOUR, ACO_PHASE_ANT_DISTANCE,
                                 ACO_PHASE_UPDATE_PHEROMONE};
    hipError_t e;
    if (ev && (e = hipEventRecord(ev[0], c->stream)) != hipSuccess) return e;
    for (int i = 0; i < 4; ++i) {
        if (c->fused && order[i] == ACO_PHASE_ANT_DISTANCE)
            e = enqueue_stats_deposit(c, cur, false, aco::LaunchTiming());
        else if (!(c->fused && order[i] == ACO_PHASE_UPDATE_PHEROMONE))
            e = enqueue_phase(c, cur, order[i]);
        if (e != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev[i + 1], c->stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// One timed iteration: the four phases launched in order with plain
// launches, each phase's first kernel storing its start time (wall_clock64)
// in d_stamp[phase], a one-thread kernel the end time and all five stamps to
// h_stamp; the statistics kernel writes the iteration's statistics to
// h_stats; one wait.  (Launches that carry timestamp events cost about 12 us
// per iteration more at n = 1000.)
static hipError_t enqueue_timed_iteration(aco_ctx* c, int cur) {
    static const int order[4] = {ACO_PHASE_PROBABILITIES, ACO_PHASE_TOUR, ACO_PHASE_ANT_DISTANCE,
                                 ACO_PHASE_UPDATE_PHEROMONE};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 4 && e == hipSuccess; ++i) {
        aco::LaunchTiming t;
        t.stamp = c->d_stamp + i;
        if (c->fused && order[i] == ACO_PHASE_ANT_DISTANCE)
            e = enqueue_stats_deposit(c, cur, true, t);  // also stamps the deposit's slot
        else if (!(c->fused && order[i] == ACO_PHASE_UPDATE_PHEROMONE))
            e = enqueue_phase(c, cur, order[i], true, t);
    }
    if (e == hipSuccess) e = aco::launch_stamp_end(c->d_stamp, c->h_stamp, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->tev_end, c->stream);
    return e;
}

static void fill_timers(float* timers, const float t[4]) {
    // Placement is fused into construction and evaporation into the
    // probability sweep: their slots report 0, the sweep reports under
    // TIME_COMPUTE_PROBABILITIES.
    timers[ACO_TIME_INIT_TOUR] = 0.0f;
    timers[ACO_TIME_TOUR] = t[1];
    timers[ACO_TIME_COMPUTE_ANT_DISTANCE] = t[2];
    timers[ACO_TIME_EVAPORATE_PHEROMONE] = 0.0f;
    timers[ACO_TIME_UPDATE_PHEROMONE] = t[3];
    timers[ACO_TIME_COMPUTE_PROBABILITIES] = t[0];
}

static void read_host_stats(const aco_ctx* c, int64_t* iter_best, int64_t* glob_best, uint32_t* valid_ants) {
    const volatile int64_t* h = c->h_stats;
    if (iter_best) *iter_best = h[0];
    if (glob_best) *glob_best = h[1];
    if (valid_ants) *valid_ants = (uint32_t)h[2];
}

extern "C" int aco_iterate_stats(aco_ctx* c, float* timers, int64_t* iter_best, int64_t* glob_best,
                                 uint32_t* valid_ants) {
    if (!c) return ACO_E_ARG;
    if (!c->initialized) return fail(c, ACO_E_STATE, "aco_iterate before aco_initialize");
    if (c->next_phase != ACO_PHASE_PROBABILITIES) return fail(c, ACO_E_STATE, "aco_iterate inside a phased iteration");
    HIP_TRY(c, enqueue_timed_iteration(c, c->cur));
    c->cur ^= 1;
    HIP_TRY(c, hipEventSynchronize(c->tev_end));
    if (timers) {
        // phase i: from its first kernel's start to the next phase's (the
        // last one to the end stamp)
        const volatile int64_t* h = c->h_stamp;
        float t[4];
        for (int i = 0; i < 4; ++i) t[i] = (float)((double)(h[i + 1] - h[i]) / c->wall_khz);
        fill_timers(timers, t);
    }
    read_host_stats(c, iter_best, glob_best, valid_ants);
    return ACO_OK;
}

extern "C" int aco_iterate(aco_ctx* c, float* timers) { return aco_iterate_stats(c, timers, nullptr, nullptr, nullptr); }

// Graphs of an even number of iterations, starting with pheromone buffer 0.
// Each graph launch leaves a short gap before its first kernel, so long runs
// replay a graph of kGraphLong iterations (n = 1000: 0.3063 -> 0.3043 ms per
// iteration) and the rest of a run the two-iteration graph.
constexpr int32_t kGraphLong = 16;

static int build_graph(aco_ctx* c, int32_t iterations, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int32_t it = 0; it < iterations && e == hipSuccess; ++it) e = enqueue_iteration(c, it & 1, nullptr);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return fail(c, ACO_E_HIP, "kernel launch during capture failed: %s", hipGetErrorString(e));
    }
    HIP_TRY(c, e2);
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIP_TRY(c, e);
    return ACO_OK;
}

extern "C" int aco_run(aco_ctx* c, int32_t iterations, float* elapsed_ms) {
    if (!c) return ACO_E_ARG;
    if (!c->initialized) return fail(c, ACO_E_STATE, "aco_run before aco_initialize");
    if (iterations < 0) return fail(c, ACO_E_ARG, "negative iteration count");
    if (c->next_phase != ACO_PHASE_PROBABILITIES) return fail(c, ACO_E_STATE, "aco_run inside a phased iteration");
    if (!c->graph) {
        int rc = build_graph(c, 2, &c->graph);
        if (rc) return rc;
    }
    if (!c->graph_long && iterations > kGraphLong) {
        int rc = build_graph(c, kGraphLong, &c->graph_long);
        if (rc) return rc;
    }
    HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
    int32_t left = iterations;
    if (left > 0 && c->cur == 1) {
        HIP_TRY(c, enqueue_iteration(c, 1, nullptr));
        c->cur = 0;
        --left;
    }
    if (c->graph_long)
        for (; left >= kGraphLong; left -= kGraphLong) HIP_TRY(c, hipGraphLaunch(c->graph_long, c->stream));
    for (; left >= 2; left -= 2) HIP_TRY(c, hipGraphLaunch(c->graph, c->stream));
    if (left == 1) {
        HIP_TRY(c, enqueue_iteration(c, 0, nullptr));
        c->cur = 1;
    }
    HIP_TRY(c, hipEventRecord(c->ev[5], c->stream));
    HIP_TRY(c, hipEventSynchronize(c->ev[5]));
    if (elapsed_ms) HIP_TRY(c, hipEventElapsedTime(elapsed_ms, c->ev[0], c->ev[5]));
    return ACO_OK;
}

static int copy_in(aco_ctx* c, void* dst, const void* src, size_t bytes);
static int copy_out(aco_ctx* c, void* dst, const void* src, size_t bytes);
static int read_state(aco_ctx* c, DevState* st);

// TSP::execute's loop (src/TSP.cpp:23-79, termination src/TSP.cpp:168-171)
// without a host round trip per iteration: the statistics kernel halts the
// run after the first iteration whose best <= stop_at and the remaining
// graph-replayed iterations are skipped on the device; the host looks at the
// state once per `chunk` iterations.
extern "C" int aco_run_until(aco_ctx* c, int32_t max_iterations, int64_t stop_at, int32_t chunk,
                             int32_t* iterations_run, int32_t* halted) {
    return aco_run_until_time(c, max_iterations, stop_at, -1.0f, chunk, iterations_run, halted, nullptr);
}

// The same loop with the time test of TSP::terminationCondition
// (elapsed >= maxTime, src/TSP.cpp:168-171): elapsed is the HIP-event time
// of the iterations run, checked before every chunk.
extern "C" int aco_run_until_time(aco_ctx* c, int32_t max_iterations, int64_t stop_at, float time_limit_ms,
                                  int32_t chunk, int32_t* iterations_run, int32_t* halted, float* elapsed_ms) {
    if (!c) return ACO_E_ARG;
    if (!c->initialized) return fail(c, ACO_E_STATE, "aco_run_until before aco_initialize");
    if (max_iterations < 0 || chunk < 1) return fail(c, ACO_E_ARG, "negative iteration count or chunk < 1");
    if (c->next_phase != ACO_PHASE_PROBABILITIES) return fail(c, ACO_E_STATE, "aco_run_until inside a phased iteration");
    DevState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const uint32_t iter0 = st.iter;
    const int cur0 = c->cur;
    st.stop_at = stop_at < 0 ? -1 : stop_at;
    st.halted = 0;
    st.skip = 0;
    if ((rc = copy_in(c, c->d_state, &st, sizeof(st)))) return rc;
    int32_t left = max_iterations;
    double elapsed = 0.0;
    bool timed_out = false;
    while (left > 0) {
        if (time_limit_ms >= 0.0f && elapsed >= (double)time_limit_ms) {
            timed_out = true;
            break;
        }
        const int32_t step = left < chunk ? left : chunk;
        float ms = 0.0f;
        if ((rc = aco_run(c, step, &ms))) return rc;
        elapsed += ms;
        left -= step;
        if ((rc = read_state(c, &st))) return rc;
        if (st.halted) break;
    }
    const int32_t ran = (int32_t)(st.iter - iter0);
    const int32_t was_halted = st.halted ? 1 : (timed_out ? 2 : 0);
    // skipped iterations wrote nothing: the current pheromone buffer follows
    // the iterations that ran
    c->cur = (cur0 + ran) & 1;
    st.stop_at = -1;
    st.halted = 0;
    st.skip = 0;
    if ((rc = copy_in(c, c->d_state, &st, sizeof(st)))) return rc;
    if (iterations_run) *iterations_run = ran;
    if (halted) *halted = was_halted;
    if (elapsed_ms) *elapsed_ms = (float)elapsed;
    return ACO_OK;
}

extern "C" int aco_profile(aco_ctx* c, int32_t iterations, float* kernel_ms) {
    if (!c || !kernel_ms || iterations < 1) return ACO_E_ARG;
    if (!c->initialized) return fail(c, ACO_E_STATE, "aco_profile before aco_initialize");
    if (c->next_phase != ACO_PHASE_PROBABILITIES) return fail(c, ACO_E_STATE, "aco_profile inside a phased iteration");
    // each phase from its first kernel's start to the next phase's first
    // kernel's start (the last one to the end stamp), from the wall-clock
    // stamps the kernels themselves store: events around the launches add
    // their own dispatch overhead to every interval
    double acc[4] = {0, 0, 0, 0};
    for (int32_t it = 0; it < iterations; ++it) {
        HIP_TRY(c, enqueue_timed_iteration(c, c->cur));
        c->cur ^= 1;
        HIP_TRY(c, hipEventSynchronize(c->tev_end));
        const volatile int64_t* h = c->h_stamp;
        for (int i = 0; i < 4; ++i) acc[i] += (double)(h[i + 1] - h[i]) / c->wall_khz;
    }
    for (int i = 0; i < 4; ++i) kernel_ms[i] = (float)(acc[i] / iterations);
    return ACO_OK;
}

extern "C" int aco_calibrate_copy(aco_ctx* c, int64_t bytes, int32_t reps, float* ms) {
    if (!c || !ms || bytes <= 0 || (bytes % 16) != 0 || reps < 1) return ACO_E_ARG;
    float4 *a = nullptr, *b = nullptr;
    HIP_TRY(c, hipMalloc(reinterpret_cast<void**>(&a), (size_t)bytes));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b), (size_t)bytes);
    if (e != hipSuccess) {
        hipFree(a);
        return fail(c, ACO_E_NOMEM, "calibration buffer: %s", hipGetErrorString(e));
    }
    e = hipMemsetAsync(a, 0, (size_t)bytes, c->stream);
    if (e == hipSuccess) e = aco::launch_copy(a, b, bytes / 16, c->stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(c->ev[0], c->stream);
    for (int32_t r = 0; r < reps && e == hipSuccess; ++r) e = aco::launch_copy(a, b, bytes / 16, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, c->ev[0], c->ev[1]);
    hipFree(a);
    hipFree(b);
    HIP_TRY(c, e);
    *ms = t / reps;
    return ACO_OK;
}

// Latency floor of one construction step at this colony's k: a single wave
// repeating the step's dependent chain (k_probe_step).  out = {core clock
// cycles per step, ns per step}.
extern "C" int aco_probe_step(aco_ctx* c, int32_t iters, float* out) {
    if (!c || !out || iters < 1) return ACO_E_ARG;
    if (c->k < 1) return fail(c, ACO_E_STATE, "the step probe needs candidate lists");
    int dev = 0, wall_khz = 100000;
    HIP_TRY(c, hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || wall_khz <= 0)
        wall_khz = 100000;
    (void)hipGetLastError();
    unsigned long long* d_out = reinterpret_cast<unsigned long long*>(c->d_scratch64);  // 4 x 8 bytes
    int32_t* d_sink = reinterpret_cast<int32_t*>(c->d_scratch64 + 2);
    HIP_TRY(c, aco::launch_probe_step(iters < 1000 ? iters : 1000, c->k, d_out, d_sink, c->stream));  // warm-up
    HIP_TRY(c, aco::launch_probe_step(iters, c->k, d_out, d_sink, c->stream));
    unsigned long long h[2] = {0, 0};
    int rc = copy_out(c, h, d_out, sizeof(h));
    if (rc) return rc;
    out[0] = (float)((double)h[0] / iters);
    out[1] = (float)((double)h[1] * 1e6 / wall_khz / iters);
    return ACO_OK;
}

// ---------------------------------------------------------------------------
// Taps and state access.
static int copy_pitched_out(aco_ctx* c, void* dst, const void* src, size_t elem) {
    HIP_TRY(c, hipMemcpy2DAsync(dst, (size_t)c->n * elem, src, (size_t)c->P * elem, (size_t)c->n * elem, c->n,
                                hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ACO_OK;
}

static int copy_out(aco_ctx* c, void* dst, const void* src, size_t bytes) {
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ACO_OK;
}

static int copy_in(aco_ctx* c, void* dst, const void* src, size_t bytes) {
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ACO_OK;
}

static int read_state(aco_ctx* c, DevState* st) { return copy_out(c, st, c->d_state, sizeof(*st)); }

extern "C" int aco_get_stats(aco_ctx* c, int64_t* iter_best, int64_t* glob_best, uint32_t* valid_ants) {
    if (!c) return ACO_E_ARG;
    DevState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    if (iter_best) *iter_best = st.iter_best;
    if (glob_best) *glob_best = st.glob_best;
    if (valid_ants) *valid_ants = c->p.aging ? (uint32_t)st.num_valid : (uint32_t)c->m;
    return ACO_OK;
}

extern "C" int aco_get_best_tour(aco_ctx* c, int32_t* tour) {
    if (!c || !tour) return ACO_E_ARG;
    return copy_out(c, tour, c->d_best_tour, sizeof(int32_t) * c->n);
}

extern "C" int aco_set_best(aco_ctx* c, int64_t glob_best, const int32_t* tour) {
    if (!c) return ACO_E_ARG;
    DevState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    st.glob_best = glob_best;
    rc = copy_in(c, c->d_state, &st, sizeof(st));
    if (rc) return rc;
    if (tour) return copy_in(c, c->d_best_tour, tour, sizeof(int32_t) * c->n);
    return ACO_OK;
}

extern "C" int aco_get_tours(aco_ctx* c, int32_t* tours) {
    if (!c || !tours) return ACO_E_ARG;
    return copy_out(c, tours, c->d_tours, sizeof(int32_t) * (size_t)c->m * c->n);
}

extern "C" int aco_get_lengths(aco_ctx* c, int64_t* lengths) {
    if (!c || !lengths) return ACO_E_ARG;
    return copy_out(c, lengths, c->d_len, sizeof(int64_t) * c->m);
}

extern "C" int aco_get_pheromone(aco_ctx* c, float* tau) {
    if (!c || !tau) return ACO_E_ARG;
    return copy_pitched_out(c, tau, c->d_tau[c->cur], sizeof(float));
}

extern "C" int aco_set_pheromone(aco_ctx* c, const float* tau) {
    if (!c || !tau) return ACO_E_ARG;
    HIP_TRY(c, hipMemcpy2DAsync(c->d_tau[c->cur], (size_t)c->P * 4, tau, (size_t)c->n * 4, (size_t)c->n * 4, c->n,
                                hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ACO_OK;
}

extern "C" int aco_get_choice(aco_ctx* c, float* choice) {
    if (!c || !choice) return ACO_E_ARG;
    return copy_pitched_out(c, choice, c->d_choice, sizeof(float));
}

extern "C" int aco_get_choice_nn(aco_ctx* c, float* choice_nn) {
    if (!c || !choice_nn) return ACO_E_ARG;
    if (c->k == 0) return fail(c, ACO_E_STATE, "no candidate lists in full-row mode");
    return copy_out(c, choice_nn, c->d_choice_nn, sizeof(float) * (size_t)c->n * c->k);
}

extern "C" int aco_get_nn(aco_ctx* c, int32_t* nn) {
    if (!c || !nn) return ACO_E_ARG;
    return copy_out(c, nn, c->d_nn, sizeof(int32_t) * (size_t)c->n * c->nn_cols);
}

extern "C" int aco_get_mean(aco_ctx* c, float* mean) {
    if (!c || !mean) return ACO_E_ARG;
    return copy_out(c, mean, c->d_mean, sizeof(float) * c->n);
}

extern "C" int aco_get_dist(aco_ctx* c, int32_t* dist) {
    if (!c || !dist) return ACO_E_ARG;
    return copy_pitched_out(c, dist, c->d_dist, sizeof(int32_t));
}

extern "C" int aco_get_eta(aco_ctx* c, float* eta, int32_t* eta_len) {
    if (!c || !eta_len) return ACO_E_ARG;
    *eta_len = c->eta_len;
    if (!eta) return ACO_OK;
    return copy_out(c, eta, c->d_eta, sizeof(float) * c->eta_len);
}

extern "C" int aco_get_aging(aco_ctx* c, int32_t* age, uint8_t* valid, int32_t* perf, int64_t* prev_len,
                             int32_t* scalars) {
    if (!c) return ACO_E_ARG;
    int rc;
    if (age && (rc = copy_out(c, age, c->d_age, sizeof(int32_t) * c->m))) return rc;
    if (valid && (rc = copy_out(c, valid, c->d_valid, (size_t)c->m))) return rc;
    if (perf && (rc = copy_out(c, perf, c->d_perf, sizeof(int32_t) * c->m))) return rc;
    if (prev_len && (rc = copy_out(c, prev_len, c->d_prev_len, sizeof(int64_t) * c->m))) return rc;
    if (scalars) {
        DevState st;
        if ((rc = read_state(c, &st))) return rc;
        scalars[0] = st.num_valid;
        scalars[1] = st.converged;
        scalars[2] = st.gperf;
        scalars[3] = (int32_t)st.iter;
    }
    return ACO_OK;
}

extern "C" int aco_set_aging(aco_ctx* c, const int32_t* age, const uint8_t* valid, const int32_t* perf,
                             const int64_t* prev_len, const int32_t* scalars) {
    if (!c) return ACO_E_ARG;
    int rc;
    if (age && (rc = copy_in(c, c->d_age, age, sizeof(int32_t) * c->m))) return rc;
    if (valid && (rc = copy_in(c, c->d_valid, valid, (size_t)c->m))) return rc;
    if (perf && (rc = copy_in(c, c->d_perf, perf, sizeof(int32_t) * c->m))) return rc;
    if (prev_len && (rc = copy_in(c, c->d_prev_len, prev_len, sizeof(int64_t) * c->m))) return rc;
    if (scalars) {
        DevState st;
        if ((rc = read_state(c, &st))) return rc;
        st.num_valid = scalars[0];
        st.converged = scalars[1];
        st.gperf = scalars[2];
        st.iter = (uint32_t)scalars[3];
        if ((rc = copy_in(c, c->d_state, &st, sizeof(st)))) return rc;
    }
    return ACO_OK;
}

extern "C" int aco_get_counters(aco_ctx* c, uint64_t* counters) {
    if (!c || !counters) return ACO_E_ARG;
    DevState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    counters[0] = st.fb_steps;
    counters[1] = st.fb_scanned;
    counters[2] = st.tours;
    counters[3] = st.fb_listed;
    return ACO_OK;
}

extern "C" int aco_get_config(aco_ctx* c, int32_t* config) {
    if (!c || !config) return ACO_E_ARG;
    config[0] = c->path;
    config[1] = c->apb;
    config[2] = c->young_blocks;
    config[3] = c->max_lds;
    return ACO_OK;
}

// Phase-by-phase iteration for hosts that drive Colony::constructSolution's
// virtual steps themselves (src/Colony.cu:537-594).  Phases must come in the
// order probabilities, tour, ant distances, [evaporate], update; `update`
// completes the iteration (pheromone buffers swap).
extern "C" int aco_phase(aco_ctx* c, int32_t phase, float* ms) {
    if (!c) return ACO_E_ARG;
    if (!c->initialized) return fail(c, ACO_E_STATE, "aco_phase before aco_initialize");
    static const int next_of[5] = {ACO_PHASE_TOUR, ACO_PHASE_ANT_DISTANCE, ACO_PHASE_EVAPORATE,
                                   ACO_PHASE_UPDATE_PHEROMONE, ACO_PHASE_PROBABILITIES};
    if (phase < 0 || phase > 4) return fail(c, ACO_E_ARG, "unknown phase %d", phase);
    if (phase != c->next_phase) {
        // evaporation may be skipped: it is already folded into the sweep
        if (!(c->next_phase == ACO_PHASE_EVAPORATE && phase == ACO_PHASE_UPDATE_PHEROMONE))
            return fail(c, ACO_E_STATE, "phase %d out of order (expected %d)", phase, c->next_phase);
    }
    if (ms) {
        if (phase == ACO_PHASE_EVAPORATE) {
            *ms = 0.0f;  // folded into the probability sweep: no kernel
        } else {
            aco::LaunchTiming t;
            t.start = c->ev[0];
            t.stop = c->ev[1];
            HIP_TRY(c, enqueue_phase(c, c->cur, phase, false, t));
            HIP_TRY(c, hipEventSynchronize(c->ev[1]));
            HIP_TRY(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
        }
        c->phases_async = 0;
    } else {
        // no wait: one event at the iteration's start and one after each
        // phase, read by aco_phase_end; the statistics go to h_stats
        if (phase == ACO_PHASE_PROBABILITIES) c->phases_async = 0;
        if (phase == ACO_PHASE_EVAPORATE) {
            HIP_TRY(c, hipEventRecord(c->pev[1 + phase], c->stream));  // no kernel: a marker
        } else {
            aco::LaunchTiming t;
            t.start = phase == ACO_PHASE_PROBABILITIES ? c->pev[0] : nullptr;
            t.stop = c->pev[1 + phase];
            HIP_TRY(c, enqueue_phase(c, c->cur, phase, phase == ACO_PHASE_ANT_DISTANCE, t));
        }
        c->phases_async |= 1u << phase;
    }
    c->next_phase = next_of[phase];
    if (phase == ACO_PHASE_UPDATE_PHEROMONE) c->cur ^= 1;
    return ACO_OK;
}

// Completes an iteration whose phases were enqueued without waiting: one
// wait for the stream, then the phase times from their events and the
// statistics from the state copied after the update phase.
extern "C" int aco_phase_end(aco_ctx* c, float* timers, int64_t* iter_best, int64_t* glob_best,
                             uint32_t* valid_ants) {
    if (!c) return ACO_E_ARG;
    if (!(c->phases_async & (1u << ACO_PHASE_UPDATE_PHEROMONE)))
        return fail(c, ACO_E_STATE, "aco_phase_end without an iteration of asynchronous phases (ms == NULL)");
    HIP_TRY(c, hipEventSynchronize(c->pev[1 + ACO_PHASE_UPDATE_PHEROMONE]));
    float t[5] = {0, 0, 0, 0, 0};
    hipEvent_t prev = c->pev[0];
    for (int ph = 0; ph < 5; ++ph)
        if (c->phases_async & (1u << ph)) {
            HIP_TRY(c, hipEventElapsedTime(&t[ph], prev, c->pev[1 + ph]));
            prev = c->pev[1 + ph];
        }
    const uint32_t done = c->phases_async;
    c->phases_async = 0;
    if (timers) {
        timers[ACO_TIME_INIT_TOUR] = 0.0f;  // placement runs inside the tour kernel
        timers[ACO_TIME_TOUR] = t[ACO_PHASE_TOUR];
        timers[ACO_TIME_COMPUTE_ANT_DISTANCE] = t[ACO_PHASE_ANT_DISTANCE];
        timers[ACO_TIME_EVAPORATE_PHEROMONE] = t[ACO_PHASE_EVAPORATE];
        timers[ACO_TIME_UPDATE_PHEROMONE] = t[ACO_PHASE_UPDATE_PHEROMONE];
        timers[ACO_TIME_COMPUTE_PROBABILITIES] = t[ACO_PHASE_PROBABILITIES];
    }
    if ((done & (1u << ACO_PHASE_ANT_DISTANCE)) == 0) {  // its statistics were not mirrored
        int64_t ib = 0, gb = 0;
        uint32_t nv = 0;
        int rc = aco_get_stats(c, &ib, &gb, &nv);
        if (rc) return rc;
        if (iter_best) *iter_best = ib;
        if (glob_best) *glob_best = gb;
        if (valid_ants) *valid_ants = nv;
        return ACO_OK;
    }
    read_host_stats(c, iter_best, glob_best, valid_ants);
    return ACO_OK;
}

// Heuristic on candidate edges (n x k): replaces eta^beta on the NN-list
// edges only (the fallback keeps plain eta^beta); NULL restores eta^beta.
extern "C" int aco_set_heuristic_nn(aco_ctx* c, const float* h) {
    if (!c) return ACO_E_ARG;
    if (c->k == 0) return fail(c, ACO_E_STATE, "no candidate lists in full-row mode");
    if (!c->initialized) return fail(c, ACO_E_STATE, "aco_set_heuristic_nn before aco_initialize");
    if (h == nullptr) {
        HIP_TRY(c, aco::launch_nn_aux(c->d_nn, c->d_dist, c->d_eta, c->eta_len, c->n, c->P, c->k, c->d_nn16,
                                      c->d_eta_nn, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return ACO_OK;
    }
    return copy_in(c, c->d_eta_nn, h, sizeof(float) * (size_t)c->n * c->k);
}

extern "C" int aco_get_heuristic_nn(aco_ctx* c, float* h) {
    if (!c || !h) return ACO_E_ARG;
    if (c->k == 0) return fail(c, ACO_E_STATE, "no candidate lists in full-row mode");
    return copy_out(c, h, c->d_eta_nn, sizeof(float) * (size_t)c->n * c->k);
}

// Device arrays of a colony for the learned edge prior (aco_prior.hip).
int aco::colony_view(aco_ctx* c, aco::InstanceView* v, char* err, size_t errlen) {
    if (!c) return ACO_E_ARG;
    if (c->k == 0) {
        snprintf(err, errlen, "the prior needs candidate lists (full-row mode has none)");
        return ACO_E_STATE;
    }
    if (!c->initialized) {
        snprintf(err, errlen, "colony not initialized");
        return ACO_E_STATE;
    }
    v->in.dist = c->d_dist;
    v->in.nn = c->d_nn;
    v->in.n = c->n;
    v->in.P = c->P;
    v->in.k = c->k;
    v->best_tour = c->d_best_tour;
    v->eta_nn = c->d_eta_nn;
    v->stream = c->stream;
    return ACO_OK;
}
