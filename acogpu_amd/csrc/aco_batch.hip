// Batched solver (include/acogpu.h, aco_batch_*): many independent instances
// in shared device buffers, one persistent construction launch per iteration,
// two iterations per hipGraph.  All device memory is allocated at create.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/acogpu.h"
#include "aco_kernels.h"
#include "aco_prior.h"

using aco::BatchBuffers;
using aco::DevState;
using aco::InstMeta;

struct aco_batch {
    aco_params p;
    int32_t count = 0, k = 0, nmax = 0, mmax = 0, pmax = 0, apb = 8, apw = 1, grid = 0;
    int64_t total_rows = 0, total_ants = 0, mat_elems = 0, nk_elems = 0, tour_elems = 0, dep_elems = 0;
    std::vector<InstMeta> meta;
    std::vector<int32_t> n, m;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {};
    hipEvent_t kev[5] = {};  // aco_batch_profile: around the four kernels of an iteration
    BatchBuffers buf{};
    float* tau[2] = {nullptr, nullptr};
    int2* items = nullptr;
    int32_t cur = 0;
    bool initialized = false;
    hipGraphExec_t graph = nullptr;
    std::vector<void*> allocs;
    char err[512] = {0};
};

static thread_local char g_batch_err[512] = "no error";

static int bfail(aco_batch* b, int code, const char* fmt, ...) {
    char* out = b ? b->err : g_batch_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(out, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define BHIP(b, expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return bfail((b), ACO_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static int balloc(aco_batch* b, T** out, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return bfail(b, ACO_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    b->allocs.push_back(q);
    *out = static_cast<T*>(q);
    return ACO_OK;
}

static void bfree(aco_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (void* q : b->allocs) (void)hipFree(q);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : b->kev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

extern "C" const char* aco_batch_last_error(const aco_batch* b) { return b ? b->err : g_batch_err; }

// Work items (instance, first ant) for W ants per block, longest instances first.
// Work items of W waves x G ants (one instance each) in QUEUE_BINS bins,
// one per XCD: instances by decreasing work (ant slots x steps) go to the
// least loaded bin, and each instance's items stay together in its bin.
// Also sizes the persistent grid that fills the device with blocks.
static int build_items(aco_batch* b, int32_t W, int32_t G) {
    if (G == 1 && b->nmax > 2048 && W > 8)
        return bfail(b, ACO_E_ARG, "more than 8 ants per block need n <= 2048 (largest n=%d)", b->nmax);
    const int32_t per = W * G;
    std::vector<int64_t> work(b->count);
    for (int32_t i = 0; i < b->count; ++i) work[i] = (int64_t)((b->m[i] + per - 1) / per) * per * b->n[i];
    std::vector<int32_t> order(b->count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return work[x] > work[y]; });
    std::vector<std::vector<int32_t>> bins(aco::QUEUE_BINS);
    std::vector<int64_t> load(aco::QUEUE_BINS, 0);
    for (int32_t i : order) {
        const int32_t h = (int32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        bins[h].push_back(i);
        load[h] += work[i];
    }
    std::vector<int2> it;
    std::vector<int32_t> off(aco::QUEUE_BINS + 1, 0);
    for (int32_t h = 0; h < aco::QUEUE_BINS; ++h) {
        off[h] = (int32_t)it.size();
        for (int32_t i : bins[h])
            for (int32_t a0 = 0; a0 < b->m[i]; a0 += per) it.push_back(make_int2(i, a0));
    }
    off[aco::QUEUE_BINS] = (int32_t)it.size();
    BHIP(b, hipMemcpy(b->items, it.data(), sizeof(int2) * it.size(), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(const_cast<int32_t*>(b->buf.bin_off), off.data(), sizeof(int32_t) * off.size(),
                      hipMemcpyHostToDevice));
    b->buf.nitems = (int32_t)it.size();
    b->buf.ants_per_block = W;
    b->buf.ants_per_wave = G;
    b->apb = W;
    b->apw = G;
    int cus = 256, dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    (void)hipGetLastError();
    const size_t lds = aco::batch_construct_lds_bytes(b->nmax, b->k, W, G);
    const int32_t per_cu = aco::batch_construct_blocks_per_cu(b->nmax, W, G, lds);
    b->grid = std::min<int32_t>(cus * per_cu, b->buf.nitems);
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return ACO_OK;
}

extern "C" int aco_batch_create(const aco_params* p, int32_t count, const int32_t* n, const int32_t* m,
                                const int32_t* xy, aco_batch** out) {
    if (!p || !n || !xy || !out || count < 1) return bfail(nullptr, ACO_E_ARG, "null argument or count < 1");
    *out = nullptr;
    if (p->mode != ACO_MODE_CANDIDATE || p->k < 1 || p->k > 64)
        return bfail(nullptr, ACO_E_ARG, "batched instances need candidate lists with 1 <= k <= 64");
    if (p->aging) return bfail(nullptr, ACO_E_ARG, "batched instances do not support aging");
    if (p->variant != ACO_VARIANT_AS && p->variant != ACO_VARIANT_RBAS)
        return bfail(nullptr, ACO_E_ARG, "unknown variant %d", p->variant);
    if (!(p->rho >= 0.0f && p->rho <= 1.0f) || !(p->alpha >= 0.0f) || !(p->beta >= 0.0f))
        return bfail(nullptr, ACO_E_ARG, "rho must be in [0,1], alpha and beta >= 0");
    if (p->ants_per_block < 0 || p->ants_per_block > 16)
        return bfail(nullptr, ACO_E_ARG, "ants_per_block must be in [0, 16]");
    if (p->ants_per_wave != 0 && p->ants_per_wave != 1 && p->ants_per_wave != 2 && p->ants_per_wave != 4)
        return bfail(nullptr, ACO_E_ARG, "ants_per_wave must be 0, 1, 2 or 4");
    if (p->deposit_path < 0 || p->deposit_path > 2 || (p->deposit_path == 2 && p->variant != ACO_VARIANT_AS))
        return bfail(nullptr, ACO_E_ARG, "deposit_path must be 0, 1 or 2 (2: Ant System only)");

    aco_batch* b = new aco_batch();
    b->p = *p;
    b->count = count;
    b->k = p->k;
    b->n.assign(n, n + count);
    b->m.resize(count);
    int64_t off_xy = 0;
    int32_t dmax = 0;
    for (int32_t i = 0; i < count; ++i) {
        b->m[i] = m ? m[i] : n[i];
        if (n[i] < 2 || n[i] - 1 < p->k || n[i] > 65535 || b->m[i] < 1) {
            int rc = bfail(nullptr, ACO_E_ARG, "instance %d: n=%d m=%d invalid for k=%d", i, n[i], b->m[i], p->k);
            delete b;
            return rc;
        }
        int64_t xlo = INT64_MAX, xhi = INT64_MIN, ylo = INT64_MAX, yhi = INT64_MIN;
        for (int32_t c = 0; c < n[i]; ++c) {
            const int64_t x = xy[off_xy + 2 * c], y = xy[off_xy + 2 * c + 1];
            if (x >= (1 << 29) || x <= -(1 << 29) || y >= (1 << 29) || y <= -(1 << 29)) {
                int rc = bfail(nullptr, ACO_E_RANGE, "instance %d: coordinate out of range", i);
                delete b;
                return rc;
            }
            xlo = std::min(xlo, x);
            xhi = std::max(xhi, x);
            ylo = std::min(ylo, y);
            yhi = std::max(yhi, y);
        }
        off_xy += 2 * (int64_t)n[i];
        // the bounding-box diagonal bounds every distance of the instance
        const double diag = std::sqrt((double)(xhi - xlo) * (xhi - xlo) + (double)(yhi - ylo) * (yhi - ylo));
        dmax = std::max<int32_t>(dmax, (int32_t)std::ceil(diag) + 1);
    }
    if (dmax > (1 << 26)) {
        int rc = bfail(nullptr, ACO_E_RANGE, "largest distance %d exceeds the heuristic table limit", dmax);
        delete b;
        return rc;
    }
    const int32_t wdep = (p->variant == ACO_VARIANT_RBAS ? p->w : 1) + 2;
    b->meta.resize(count);
    int64_t mat = 0, nk = 0, tours = 0, ants = 0, rows = 0, dep = 0, adj = 0;
    for (int32_t i = 0; i < count; ++i) {
        InstMeta& im = b->meta[i];
        im.n = n[i];
        im.m = b->m[i];
        im.P = (n[i] + 63) & ~63;
        im.words = (n[i] + 31) / 32;
        im.mat_off = mat;
        im.nk_off = nk;
        im.tour_off = tours;
        im.ant_off = (int32_t)ants;
        im.city_off = (int32_t)rows;
        im.dep_off = (int32_t)dep;
        im.pad = 0;
        im.adj_off = adj;
        adj += (int64_t)im.n * aco::adj_pitch(im.m);
        mat += (int64_t)im.n * im.P;
        nk += (int64_t)im.n * p->k;
        tours += (int64_t)im.m * im.n;
        ants += im.m;
        rows += im.n;
        dep += im.m + wdep;
        b->nmax = std::max(b->nmax, im.n);
        b->mmax = std::max(b->mmax, im.m);
        b->pmax = std::max(b->pmax, im.P);
    }
    if (ants > INT32_MAX / 2 || rows > INT32_MAX / 2) {
        delete b;
        return bfail(nullptr, ACO_E_RANGE, "too many ants or cities in one batch");
    }
    if (p->deposit_path == 2 && b->nmax > aco::ADJ_BATCH_MAX_N) {
        const int32_t nm = b->nmax;
        delete b;
        return bfail(nullptr, ACO_E_ARG, "the batched adjacency deposit needs n <= %d (largest n=%d)",
                     aco::ADJ_BATCH_MAX_N, nm);
    }
    const int32_t gmax = aco::batch_max_ants_per_wave(b->nmax, p->k);
    if (p->ants_per_wave > gmax) {
        const int32_t nm = b->nmax;
        delete b;
        return bfail(nullptr, ACO_E_ARG, "ants_per_wave %d needs k <= %d and n <= %d (k=%d, largest n=%d)",
                     p->ants_per_wave, 64 / p->ants_per_wave, 2048 / p->ants_per_wave, p->k, nm);
    }
    {
        // the construction kernel stages the largest instance's candidate
        // table in LDS: refuse what cannot launch on this device
        int dev = 0, max_lds = 65536;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            max_lds = 65536;
        (void)hipGetLastError();
        const size_t need = aco::batch_construct_lds_bytes(b->nmax, p->k, 1, 1) + 256;
        if (need > (size_t)max_lds) {
            const int32_t nm = b->nmax;
            delete b;
            return bfail(nullptr, ACO_E_RANGE,
                         "the largest instance (n=%d, k=%d) needs %zu bytes of LDS for batched construction "
                         "(limit %d)", nm, p->k, need, max_lds);
        }
    }
    b->total_rows = rows;
    b->total_ants = ants;

    auto bail = [&](int code) {
        if (b->err[0]) std::snprintf(g_batch_err, sizeof(g_batch_err), "%s", b->err);
        bfree(b);
        delete b;
        return code;
    };
#define TRYB(expr)                         \
    do {                                   \
        int r_ = (expr);                   \
        if (r_ != ACO_OK) return bail(r_); \
    } while (0)
#define HIPB(expr)                                                                             \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bail(bfail(b, ACO_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_))); \
    } while (0)

    HIPB(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto& e : b->ev) HIPB(hipEventCreate(&e));
    BatchBuffers& B = b->buf;
    InstMeta* d_meta = nullptr;
    int32_t *d_row_inst = nullptr, *d_ant_inst = nullptr, *d_xy = nullptr;
    TRYB(balloc(b, &d_meta, (size_t)count));
    TRYB(balloc(b, &d_row_inst, (size_t)rows));
    TRYB(balloc(b, &d_ant_inst, (size_t)ants));
    TRYB(balloc(b, &d_xy, (size_t)2 * rows));
    TRYB(balloc(b, &B.dist, (size_t)mat));
    TRYB(balloc(b, &b->tau[0], (size_t)mat));
    TRYB(balloc(b, &b->tau[1], (size_t)mat));
    TRYB(balloc(b, &B.choice, (size_t)mat));
    TRYB(balloc(b, &B.nn, (size_t)nk));
    TRYB(balloc(b, &B.nn16, (size_t)nk));
    TRYB(balloc(b, &B.choice_nn, (size_t)nk));
    TRYB(balloc(b, &B.eta_nn, (size_t)nk));
    TRYB(balloc(b, &B.tours, (size_t)tours));
    TRYB(balloc(b, &B.lengths, (size_t)ants));
    TRYB(balloc(b, &B.best_tours, (size_t)rows));
    TRYB(balloc(b, &B.dep_ant, (size_t)dep));
    TRYB(balloc(b, &B.dep_val, (size_t)dep));
    B.adj = nullptr;
    B.adj_items = nullptr;
    std::vector<int2> adj_items;
    if (p->variant == ACO_VARIANT_AS && p->deposit_path != 1 && b->nmax <= aco::ADJ_BATCH_MAX_N) {
        TRYB(balloc(b, &B.adj, (size_t)adj));
        B.adj_ants = aco::adj_stage_ants(b->nmax);
        for (int32_t i = 0; i < count; ++i)
            for (int32_t a0 = 0; a0 < b->m[i]; a0 += B.adj_ants) adj_items.push_back(make_int2(i, a0));
        int2* d_items = nullptr;
        TRYB(balloc(b, &d_items, adj_items.size()));
        B.adj_items = d_items;
        B.nadj_items = (int32_t)adj_items.size();
    }
    TRYB(balloc(b, &B.state, (size_t)count));
    TRYB(balloc(b, &b->items, (size_t)ants));
    TRYB(balloc(b, &B.queue, (size_t)aco::QUEUE_BINS));
    int32_t* d_bin_off = nullptr;
    TRYB(balloc(b, &d_bin_off, (size_t)aco::QUEUE_BINS + 1));
    B.bin_off = d_bin_off;
    TRYB(balloc(b, &B.cnn, (size_t)count));
    TRYB(balloc(b, &B.tau0, (size_t)count));
    float* d_eta = nullptr;
    TRYB(balloc(b, &d_eta, (size_t)dmax + 1));

    std::vector<int32_t> row_inst((size_t)rows), ant_inst((size_t)ants);
    for (int32_t i = 0; i < count; ++i) {
        std::fill(row_inst.begin() + b->meta[i].city_off, row_inst.begin() + b->meta[i].city_off + n[i], i);
        std::fill(ant_inst.begin() + b->meta[i].ant_off, ant_inst.begin() + b->meta[i].ant_off + b->m[i], i);
    }
    std::vector<float> eta((size_t)dmax + 1);
    B.eta_pow = aco::eta_integer_power((double)p->beta, dmax);
    if (B.eta_pow) {
        for (int32_t d = 0; d <= dmax; ++d) eta[d] = aco::eta_power_value(B.eta_pow, d);
    } else {
        eta[0] = (float)std::pow(2.0, (double)p->beta);  // duplicate cities: d taken as 1/2
        for (int32_t d = 1; d <= dmax; ++d) eta[d] = (float)std::pow(1.0 / (double)d, (double)p->beta);
    }
    HIPB(hipMemcpy(d_meta, b->meta.data(), sizeof(InstMeta) * count, hipMemcpyHostToDevice));
    HIPB(hipMemcpy(d_row_inst, row_inst.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
    HIPB(hipMemcpy(d_ant_inst, ant_inst.data(), sizeof(int32_t) * ants, hipMemcpyHostToDevice));
    HIPB(hipMemcpy(d_xy, xy, sizeof(int32_t) * 2 * rows, hipMemcpyHostToDevice));
    HIPB(hipMemcpy(d_eta, eta.data(), sizeof(float) * eta.size(), hipMemcpyHostToDevice));
    if (!adj_items.empty())
        HIPB(hipMemcpy(const_cast<int2*>(B.adj_items), adj_items.data(), sizeof(int2) * adj_items.size(),
                       hipMemcpyHostToDevice));
    HIPB(hipMemset(B.dist, 0, sizeof(int32_t) * mat));
    HIPB(hipMemset(B.choice, 0, sizeof(float) * mat));
    HIPB(hipMemset(B.tours, 0, sizeof(int32_t) * tours));
    HIPB(hipMemset(B.lengths, 0, sizeof(int64_t) * ants));

    B.meta = d_meta;
    B.count = count;
    B.k = p->k;
    B.total_rows = (int32_t)rows;
    B.total_ants = (int32_t)ants;
    B.row_inst = d_row_inst;
    B.ant_inst = d_ant_inst;
    B.xy = d_xy;
    B.tau_cur = b->tau[0];
    B.tau_next = b->tau[1];
    B.eta = d_eta;
    B.eta_len = dmax + 1;
    B.items = b->items;
    B.alpha = p->alpha;
    B.one_minus_rho = 1.0f - p->rho;
    B.variant = p->variant;
    B.w = p->variant == ACO_VARIANT_RBAS ? std::max(1, p->w) : 1;
    B.tau0_policy = p->tau0_policy;
    B.tau0_const = p->tau0;
    B.rho = p->rho;
    B.seed_lo = (uint32_t)p->seed;
    B.seed_hi = (uint32_t)(p->seed >> 32);
    b->dep_elems = dep;
    // Default block: 8 waves; 16 when several ants share a wave (no per-wave
    // LDS) and the candidate table leaves room for one block per CU, so the
    // CU still runs four waves per SIMD (256 instances of n = 1000, k = 20:
    // 17.1 -> 11.3 ms per construction).
    int32_t W = p->ants_per_block;
    const int32_t G = p->ants_per_wave ? p->ants_per_wave : gmax;
    if (!W) {
        int max_lds = 65536, dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            max_lds = 65536;
        (void)hipGetLastError();
        const size_t table = aco::batch_construct_lds_bytes(b->nmax, b->k, 16, G) + 256;
        W = (G > 1 && 2 * table > (size_t)max_lds) ? 16 : 8;
    }
    TRYB(build_items(b, W, G));
    *out = b;
    return ACO_OK;
#undef TRYB
#undef HIPB
}

extern "C" int aco_batch_destroy(aco_batch* b) {
    if (!b) return ACO_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    bfree(b);
    delete b;
    return ACO_OK;
}

extern "C" int aco_batch_initialize(aco_batch* b) {
    if (!b) return ACO_E_ARG;
    b->cur = 0;
    b->buf.tau_cur = b->tau[0];
    b->buf.tau_next = b->tau[1];
    BHIP(b, aco::launch_batch_init(b->buf, b->pmax, b->nmax, b->stream));
    std::vector<DevState> st(b->count);
    for (auto& s : st) {
        std::memset(&s, 0, sizeof(s));
        s.iter_best = (int64_t)0x7fffffff - 1;
        s.glob_best = (int64_t)0x7fffffff;
        s.iter_best_ant = -1;
        s.stop_at = -1;
    }
    BHIP(b, hipMemcpyAsync(b->buf.state, st.data(), sizeof(DevState) * b->count, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemsetAsync(b->buf.queue, 0, sizeof(uint32_t) * aco::QUEUE_BINS, b->stream));
    BHIP(b, hipMemsetAsync(b->buf.best_tours, 0, sizeof(int32_t) * b->total_rows, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    // the iteration graph captures buffer addresses and the launch shape
    // only (aco_batch_autotune rebuilds it): it stays valid across
    // initialisations
    b->initialized = true;
    return ACO_OK;
}

static hipError_t batch_enqueue(aco_batch* b, int cur) {
    BatchBuffers B = b->buf;
    B.tau_cur = b->tau[cur];
    B.tau_next = b->tau[1 - cur];
    hipError_t e;
    if ((e = aco::launch_batch_iteration_kernel(B, 0, b->nmax, 0, b->stream)) != hipSuccess) return e;
    if ((e = aco::launch_batch_iteration_kernel(B, 1, b->nmax, b->grid, b->stream)) != hipSuccess) return e;
    if ((e = aco::launch_batch_iteration_kernel(B, 2, b->nmax, (b->mmax + 31) / 32, b->stream)) != hipSuccess)
        return e;
    return aco::launch_batch_iteration_kernel(B, 3, b->nmax, 0, b->stream);
}

extern "C" int aco_batch_run(aco_batch* b, int32_t iterations, float* elapsed_ms) {
    if (!b) return ACO_E_ARG;
    if (!b->initialized) return bfail(b, ACO_E_STATE, "aco_batch_run before aco_batch_initialize");
    if (iterations < 0) return bfail(b, ACO_E_ARG, "negative iteration count");
    if (!b->graph) {
        hipGraph_t g = nullptr;
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = batch_enqueue(b, 0);
        if (e == hipSuccess) e = batch_enqueue(b, 1);
        hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return bfail(b, ACO_E_HIP, "graph capture failed: %s",
                         hipGetErrorString(e != hipSuccess ? e : e2));
        }
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BHIP(b, e);
    }
    BHIP(b, hipEventRecord(b->ev[0], b->stream));
    int32_t left = iterations;
    if (left > 0 && b->cur == 1) {
        BHIP(b, batch_enqueue(b, 1));
        b->cur = 0;
        --left;
    }
    for (; left >= 2; left -= 2) BHIP(b, hipGraphLaunch(b->graph, b->stream));
    if (left == 1) {
        BHIP(b, batch_enqueue(b, 0));
        b->cur = 1;
    }
    BHIP(b, hipEventRecord(b->ev[1], b->stream));
    BHIP(b, hipEventSynchronize(b->ev[1]));
    if (elapsed_ms) BHIP(b, hipEventElapsedTime(elapsed_ms, b->ev[0], b->ev[1]));
    return ACO_OK;
}

extern "C" int aco_batch_autotune(aco_batch* b, int32_t reps, int32_t* chosen) {
    if (!b || reps < 1) return ACO_E_ARG;
    if (!b->initialized) return bfail(b, ACO_E_STATE, "aco_batch_autotune before aco_batch_initialize");
    BatchBuffers B = b->buf;
    B.tau_cur = b->tau[b->cur];
    B.tau_next = b->tau[1 - b->cur];
    // The sweep writes choice_nn from the current pheromone and the next
    // buffer's evaporated values, which the next real iteration rewrites.
    BHIP(b, aco::launch_batch_iteration_kernel(B, 0, b->nmax, 0, b->stream));
    int best_w = b->apb, best_g = b->apw;
    float best_ms = 1e30f;
    int max_lds = 65536, dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        max_lds = 65536;
    (void)hipGetLastError();
    // one ant per wave and the most the batch allows (unless fixed by
    // ants_per_wave), each with every block size
    const int32_t gmax = aco::batch_max_ants_per_wave(b->nmax, b->k);
    std::vector<int32_t> gs;
    if (b->p.ants_per_wave)
        gs.push_back(b->p.ants_per_wave);
    else
        gs = gmax > 1 ? std::vector<int32_t>{1, gmax} : std::vector<int32_t>{1};
    for (int32_t G : gs) {
        for (int32_t W : {1, 2, 4, 8, 16}) {
            if (aco::batch_construct_lds_bytes(b->nmax, b->k, W, G) + 256 > (size_t)max_lds) continue;
            if (G == 1 && b->nmax > 2048 && W > 8) continue;
            int rc = build_items(b, W, G);
            if (rc) return rc;
            B.nitems = b->buf.nitems;
            B.ants_per_block = W;
            B.ants_per_wave = G;
            float total = 0.0f;
            for (int32_t r = 0; r < reps; ++r) {
                BHIP(b, hipMemsetAsync(b->buf.queue, 0, sizeof(uint32_t) * aco::QUEUE_BINS, b->stream));
                BHIP(b, hipEventRecord(b->ev[0], b->stream));
                BHIP(b, aco::launch_batch_iteration_kernel(B, 1, b->nmax, b->grid, b->stream));
                BHIP(b, hipEventRecord(b->ev[1], b->stream));
                BHIP(b, hipEventSynchronize(b->ev[1]));
                float ms = 0.0f;
                BHIP(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
                total += ms;
            }
            if (total < best_ms) {
                best_ms = total;
                best_w = W;
                best_g = G;
            }
        }
    }
    int rc = build_items(b, best_w, best_g);
    if (rc) return rc;
    BHIP(b, hipMemsetAsync(b->buf.queue, 0, sizeof(uint32_t) * aco::QUEUE_BINS, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (chosen) *chosen = best_w;
    return ACO_OK;
}

static int check_inst(aco_batch* b, int32_t i) {
    if (!b) return ACO_E_ARG;
    if (i < 0 || i >= b->count) return bfail(b, ACO_E_ARG, "instance %d out of range", i);
    return ACO_OK;
}

extern "C" int aco_batch_get_best(aco_batch* b, int64_t* glob_best, int32_t* tours) {
    if (!b || !glob_best) return ACO_E_ARG;
    std::vector<DevState> st(b->count);
    BHIP(b, hipMemcpy(st.data(), b->buf.state, sizeof(DevState) * b->count, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < b->count; ++i) glob_best[i] = st[i].glob_best;
    if (tours)
        BHIP(b, hipMemcpy(tours, b->buf.best_tours, sizeof(int32_t) * b->total_rows, hipMemcpyDeviceToHost));
    return ACO_OK;
}

extern "C" int aco_batch_get_tours(aco_batch* b, int32_t i, int32_t* tours) {
    int rc = check_inst(b, i);
    if (rc || !tours) return rc ? rc : ACO_E_ARG;
    const InstMeta& im = b->meta[i];
    BHIP(b, hipMemcpy(tours, b->buf.tours + im.tour_off, sizeof(int32_t) * im.m * im.n, hipMemcpyDeviceToHost));
    return ACO_OK;
}

extern "C" int aco_batch_get_lengths(aco_batch* b, int32_t i, int64_t* lengths) {
    int rc = check_inst(b, i);
    if (rc || !lengths) return rc ? rc : ACO_E_ARG;
    const InstMeta& im = b->meta[i];
    BHIP(b, hipMemcpy(lengths, b->buf.lengths + im.ant_off, sizeof(int64_t) * im.m, hipMemcpyDeviceToHost));
    return ACO_OK;
}

extern "C" int aco_batch_get_pheromone(aco_batch* b, int32_t i, float* tau) {
    int rc = check_inst(b, i);
    if (rc || !tau) return rc ? rc : ACO_E_ARG;
    const InstMeta& im = b->meta[i];
    BHIP(b, hipMemcpy2D(tau, sizeof(float) * im.n, b->tau[b->cur] + im.mat_off, sizeof(float) * im.P,
                        sizeof(float) * im.n, im.n, hipMemcpyDeviceToHost));
    return ACO_OK;
}

extern "C" int aco_batch_set_pheromone(aco_batch* b, int32_t i, const float* tau) {
    int rc = check_inst(b, i);
    if (rc || !tau) return rc ? rc : ACO_E_ARG;
    const InstMeta& im = b->meta[i];
    BHIP(b, hipMemcpy2D(b->tau[b->cur] + im.mat_off, sizeof(float) * im.P, tau, sizeof(float) * im.n,
                        sizeof(float) * im.n, im.n, hipMemcpyHostToDevice));
    return ACO_OK;
}

extern "C" int aco_batch_get_config(aco_batch* b, int32_t* config) {
    if (!b || !config) return ACO_E_ARG;
    config[0] = b->apb;
    config[1] = b->grid;
    config[2] = b->buf.nitems;
    config[3] = (int32_t)aco::batch_construct_lds_bytes(b->nmax, b->k, b->apb, b->apw);
    config[4] = b->apw;
    return ACO_OK;
}

// Per-kernel HIP-event times of `iterations` eagerly launched iterations
// (the kernels the graph replays): kernel_ms = {sweep, construction,
// statistics, deposit} averages.
extern "C" int aco_batch_profile(aco_batch* b, int32_t iterations, float* kernel_ms) {
    if (!b || !kernel_ms || iterations < 1) return ACO_E_ARG;
    if (!b->initialized) return bfail(b, ACO_E_STATE, "aco_batch_profile before aco_batch_initialize");
    for (auto& e : b->kev)
        if (!e) BHIP(b, hipEventCreate(&e));
    double acc[4] = {0, 0, 0, 0};
    for (int32_t it = 0; it < iterations; ++it) {
        BatchBuffers B = b->buf;
        B.tau_cur = b->tau[b->cur];
        B.tau_next = b->tau[1 - b->cur];
        const int32_t grids[4] = {0, b->grid, (b->mmax + 31) / 32, 0};
        BHIP(b, hipEventRecord(b->kev[0], b->stream));
        for (int q = 0; q < 4; ++q) {
            BHIP(b, aco::launch_batch_iteration_kernel(B, q, b->nmax, grids[q], b->stream));
            BHIP(b, hipEventRecord(b->kev[q + 1], b->stream));
        }
        b->cur ^= 1;
        BHIP(b, hipEventSynchronize(b->kev[4]));
        for (int q = 0; q < 4; ++q) {
            float t = 0.0f;
            BHIP(b, hipEventElapsedTime(&t, b->kev[q], b->kev[q + 1]));
            acc[q] += t;
        }
    }
    for (int q = 0; q < 4; ++q) kernel_ms[q] = (float)(acc[q] / iterations);
    return ACO_OK;
}

extern "C" int aco_batch_get_counters(aco_batch* b, uint64_t* counters) {
    if (!b || !counters) return ACO_E_ARG;
    std::vector<DevState> st(b->count);
    BHIP(b, hipMemcpy(st.data(), b->buf.state, sizeof(DevState) * b->count, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < b->count; ++i) {
        counters[2 * i] = st[i].fb_steps;
        counters[2 * i + 1] = st[i].fb_scanned;
    }
    return ACO_OK;
}

extern "C" int aco_batch_set_heuristic_nn(aco_batch* b, int32_t i, const float* h) {
    int rc = check_inst(b, i);
    if (rc || !h) return rc ? rc : ACO_E_ARG;
    const InstMeta& im = b->meta[i];
    BHIP(b, hipMemcpy(b->buf.eta_nn + im.nk_off, h, sizeof(float) * im.n * b->k, hipMemcpyHostToDevice));
    return ACO_OK;
}

int aco::batch_count(aco_batch* b) { return b ? b->count : 0; }

// Device arrays of one batch instance for the learned edge prior.
int aco::batch_view(aco_batch* b, int32_t i, aco::InstanceView* v, char* err, size_t errlen) {
    if (!b || i < 0 || i >= b->count) {
        snprintf(err, errlen, "batch instance %d out of range", i);
        return ACO_E_ARG;
    }
    if (!b->initialized) {
        snprintf(err, errlen, "batch not initialized");
        return ACO_E_STATE;
    }
    const InstMeta& im = b->meta[i];
    v->in.dist = b->buf.dist + im.mat_off;
    v->in.nn = b->buf.nn + im.nk_off;
    v->in.n = im.n;
    v->in.P = im.P;
    v->in.k = b->k;
    v->best_tour = b->buf.best_tours + im.city_off;
    v->eta_nn = b->buf.eta_nn + im.nk_off;
    v->stream = b->stream;
    return ACO_OK;
}

// Candidate-list width: for each k in ks, a batch of the same instances runs
// `trial_iterations` iterations (after the ants-per-block autotune); the
// choice is the fastest k whose mean best tour is within `tolerance` of the
// best mean over all k.  report (may be NULL) gets, per k, {ms per iteration,
// mean best length, chosen ants per block}.  Unlike the ants-per-block tuner
// this changes the algorithm (the candidate set), so it is a separate call.
extern "C" int aco_batch_tune_k(const aco_params* p, int32_t count, const int32_t* n, const int32_t* m,
                                const int32_t* xy, const int32_t* ks, int32_t nks, int32_t trial_iterations,
                                float tolerance, int32_t* chosen_k, float* report) {
    if (!p || !ks || nks < 1 || nks > 64 || trial_iterations < 1 || !(tolerance >= 0.0f) || !chosen_k)
        return bfail(nullptr, ACO_E_ARG, "aco_batch_tune_k: bad arguments");
    std::vector<double> ms((size_t)nks), mean((size_t)nks);
    std::vector<int32_t> apb((size_t)nks);
    for (int32_t t = 0; t < nks; ++t) {
        aco_params q = *p;
        q.k = ks[t];
        aco_batch* b = nullptr;
        int rc = aco_batch_create(&q, count, n, m, xy, &b);
        if (rc) return rc;
        int32_t w = 0;
        float el = 0.0f;
        if ((rc = aco_batch_initialize(b)) || (rc = aco_batch_autotune(b, 2, &w)) || (rc = aco_batch_initialize(b)) ||
            (rc = aco_batch_run(b, trial_iterations, &el))) {
            std::snprintf(g_batch_err, sizeof(g_batch_err), "k=%d: %s", ks[t], b->err);
            aco_batch_destroy(b);
            return rc;
        }
        std::vector<int64_t> gb((size_t)count);
        rc = aco_batch_get_best(b, gb.data(), nullptr);
        aco_batch_destroy(b);
        if (rc) return rc;
        double s = 0.0;
        for (int64_t v : gb) s += (double)v;
        ms[t] = el / trial_iterations;
        mean[t] = s / count;
        apb[t] = w;
    }
    const double best_mean = *std::min_element(mean.begin(), mean.end());
    int32_t pick = -1;
    for (int32_t t = 0; t < nks; ++t)
        if (mean[t] <= best_mean * (1.0 + tolerance) && (pick < 0 || ms[t] < ms[pick])) pick = t;
    *chosen_k = ks[pick];
    if (report)
        for (int32_t t = 0; t < nks; ++t) {
            report[3 * t] = (float)ms[t];
            report[3 * t + 1] = (float)mean[t];
            report[3 * t + 2] = (float)apb[t];
        }
    return ACO_OK;
}

// Island model (the reference's broker build is an empty extension point,
// inc/Colony.h:142-149): the batch holds copies of one instance, island i on
// Philox stream i; every `exchange_every` iterations the shortest best-so-far
// tour over all islands becomes every island's best-so-far (the rank-based
// deposit then reinforces it everywhere).
extern "C" int aco_batch_exchange_best(aco_batch* b) {
    if (!b) return ACO_E_ARG;
    if (!b->initialized) return bfail(b, ACO_E_STATE, "aco_batch_exchange_best before aco_batch_initialize");
    for (int32_t i = 1; i < b->count; ++i)
        if (b->n[i] != b->n[0]) return bfail(b, ACO_E_ARG, "islands must be copies of one instance (n differs)");
    BHIP(b, aco::launch_batch_exchange(b->buf, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return ACO_OK;
}

extern "C" int aco_batch_run_islands(aco_batch* b, int32_t iterations, int32_t exchange_every, float* elapsed_ms) {
    if (!b) return ACO_E_ARG;
    if (iterations < 0 || exchange_every < 1) return bfail(b, ACO_E_ARG, "bad iteration count or exchange interval");
    float total = 0.0f;
    for (int32_t left = iterations; left > 0;) {
        const int32_t step = left < exchange_every ? left : exchange_every;
        float ms = 0.0f;
        int rc = aco_batch_run(b, step, &ms);
        if (rc) return rc;
        total += ms;
        left -= step;
        if ((rc = aco_batch_exchange_best(b))) return rc;
    }
    if (elapsed_ms) *elapsed_ms = total;
    return ACO_OK;
}
