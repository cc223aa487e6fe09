// gfx950 kernels of the Ant-System hot path.
//
// Memory layout in HBM (one colony):
//   dist, tau[2], choice   n rows x P columns (P = n rounded up to 64), int32 /
//                          fp32; padding columns hold 0 (tau, choice).
//   nn, choice_nn          n x k, row-major (candidate lists and their weights).
//   tours                  m x n int32; lengths m int64.
// Every kernel is written for wave64: one wave per matrix row in the dense
// sweeps, one wave per ant in construction.
//
// Bit-exactness: this file is compiled with -ffp-contract=off, so every fp32
// expression is evaluated as written, one IEEE operation at a time, matching
// the CPU oracle.

#include <hip/hip_ext.h>

#include "aco_device.h"
#include "aco_kernels.h"

namespace aco {

// tau^alpha rounded once from fp64 (the oracle rounds the same fp64 power);
// out of line so that the alpha == 1 fast path stays small.
// tau^alpha for alpha != 1, defined as 2^(alpha log2 tau) evaluated in fp64
// by a fixed sequence of IEEE operations (no library calls, no fused
// operations), so the CPU oracle computes the same bits: log2 from the
// exponent plus 2 atanh((m-1)/(m+1)) / ln 2 for the mantissa m in
// [sqrt(1/2), sqrt(2)) (odd series to t^23, |t| <= 0.172), 2^z from the
// integer part as an exponent and e^(f ln 2), f in [0, 1), by its Taylor
// series to the 18th power.  Relative error ~1e-15 before the rounding to
// fp32.  Same operations in oracle/aco_oracle.c (orc_pow_alpha).
__device__ __noinline__ float pow_alpha(float p, float alpha) {
    if (!(p > 0.0f)) return p == 0.0f ? 0.0f : __builtin_nanf("");
    double x = (double)p;  // a float: normal in fp64
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    int32_t e = (int32_t)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double t = (m - 1.0) / (m + 1.0);
    const double t2 = t * t;
    double s = 1.0 / 23.0;
    s = s * t2 + 1.0 / 21.0;
    s = s * t2 + 1.0 / 19.0;
    s = s * t2 + 1.0 / 17.0;
    s = s * t2 + 1.0 / 15.0;
    s = s * t2 + 1.0 / 13.0;
    s = s * t2 + 1.0 / 11.0;
    s = s * t2 + 1.0 / 9.0;
    s = s * t2 + 1.0 / 7.0;
    s = s * t2 + 1.0 / 5.0;
    s = s * t2 + 1.0 / 3.0;
    s = s * t2 + 1.0;
    const double lg = (double)e + ((2.0 * t) * s) * 1.4426950408889634;
    const double z = (double)alpha * lg;
    const double k = floor(z);
    if (k < -1022.0) return 0.0f;
    if (k > 1023.0) return __builtin_inff();
    const double g = (z - k) * 0.6931471805599453;
    double q = 1.0 / 6402373705728000.0;  // 1/18!
    q = q * g + 1.0 / 355687428096000.0;
    q = q * g + 1.0 / 20922789888000.0;
    q = q * g + 1.0 / 1307674368000.0;
    q = q * g + 1.0 / 87178291200.0;
    q = q * g + 1.0 / 6227020800.0;
    q = q * g + 1.0 / 479001600.0;
    q = q * g + 1.0 / 39916800.0;
    q = q * g + 1.0 / 3628800.0;
    q = q * g + 1.0 / 362880.0;
    q = q * g + 1.0 / 40320.0;
    q = q * g + 1.0 / 5040.0;
    q = q * g + 1.0 / 720.0;
    q = q * g + 1.0 / 120.0;
    q = q * g + 1.0 / 24.0;
    q = q * g + 1.0 / 6.0;
    q = q * g + 0.5;
    q = q * g + 1.0;
    q = q * g + 1.0;
    const double scale = __longlong_as_double((long long)((uint64_t)((int64_t)k + 1023) << 52));
    return (float)(q * scale);
}

__device__ __forceinline__ float apply_alpha(float p, float alpha) { return alpha == 1.0f ? p : pow_alpha(p, alpha); }

// ---------------------------------------------------------------------------
// EUC_2D, exact integer form of (int)(sqrt(dx^2+dy^2)+0.5) (Utils::
// round_distance src/Utils.cu:212-218): the r with (2r-1)^2 <= 4D < (2r+1)^2.
__device__ __forceinline__ int32_t euc2d_exact(int64_t dx, int64_t dy) {
    const int64_t D = dx * dx + dy * dy;
    int64_t r = (int64_t)(sqrt((double)D) + 0.5);
    while (r > 0 && (2 * r - 1) * (2 * r - 1) > 4 * D) --r;
    while ((2 * r + 1) * (2 * r + 1) <= 4 * D) ++r;
    return (int32_t)r;
}

// The same distance for coordinate differences below 2^21 in magnitude, in
// fp64 without 64-bit integer products and without a loop: D = dx^2 + dy^2 <
// 2^43 is exact; s = sqrt(fp32(D)) (v_sqrt_f32, 1 ulp) is within 0.53 of
// sqrt(D) < 2^21.5, so r = rint(s) is within 1 of nint(sqrt(D)), and one step
// with the exact residual e = D - r^2 corrects it (nint(sqrt(D)) = r iff
// -r < e <= r, and r = 0 iff D = 0).
__device__ __forceinline__ int32_t euc2d_fast(int32_t dx, int32_t dy) {
    const double x = (double)dx, y = (double)dy;
    const double D = __builtin_fma(x, x, y * y);
    const double r = __builtin_rint((double)__builtin_amdgcn_sqrtf((float)D));
    const double e = __builtin_fma(-r, r, D);
    const double up = e > r ? 1.0 : 0.0, down = (r > 0.0 && e <= -r) ? 1.0 : 0.0;
    return (int32_t)(r + up - down);
}

__global__ __launch_bounds__(256) void k_dist_from_xy(const int32_t* __restrict__ xy, int32_t n, int32_t P,
                                                      int32_t* __restrict__ dist) {
    const int32_t i = blockIdx.y;
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= P) return;
    int32_t d = ACO_MAX_DISTANCE;
    if (j < n && j != i)
        d = euc2d_exact((int64_t)xy[2 * i] - xy[2 * j], (int64_t)xy[2 * i + 1] - xy[2 * j + 1]);
    dist[(size_t)i * P + j] = d;
}

__global__ __launch_bounds__(256) void k_eta_full(const int32_t* __restrict__ dist, const float* __restrict__ eta,
                                                  int32_t eta_len, int32_t n, int32_t P, float* __restrict__ out) {
    const int32_t i = blockIdx.y;
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= P) return;
    const size_t x = (size_t)i * P + j;
    out[x] = j < n ? eta_of(eta, eta_len, dist[x]) : 0.0f;
}

__global__ __launch_bounds__(256) void k_fix_diagonal(int32_t* __restrict__ dist, int32_t n, int32_t P) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dist[(size_t)i * P + i] = ACO_MAX_DISTANCE;
}

// Largest off-diagonal distance: rows strided over the grid, one atomic per
// block (a single address takes every block's atomic, so few blocks).
__global__ __launch_bounds__(256) void k_dist_max(const int32_t* __restrict__ dist, int32_t n, int32_t P,
                                                  int32_t* __restrict__ out) {
    __shared__ int32_t red[4];
    int32_t v = 0;
    for (int32_t i = blockIdx.x; i < n; i += gridDim.x)
        for (int32_t j = threadIdx.x; j < n; j += 256) {
            const int32_t d = dist[(size_t)i * P + j];
            if (j != i && d != ACO_MAX_DISTANCE) v = max(v, d);
        }
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(red[0], red[1]), max(red[2], red[3])));
}

// ---------------------------------------------------------------------------
// Nearest-neighbour lists: one wave per row, k rounds of a wave-wide minimum
// over the (d, j) keys strictly above the previous one -> the stable (d, j)
// order of dev_compute_nnlist (inc/Colony.cuh:101-123).  k < n excludes the
// city itself; k == n is the full row (self last, d = MAX_DISTANCE).
__device__ __forceinline__ void nn_row(const int32_t* __restrict__ row, int32_t n, int32_t k, int32_t i,
                                       int32_t* __restrict__ out, int lane) {
    const bool include_self = (k == n);
    uint64_t last = 0;
    for (int32_t r = 0; r < k; ++r) {
        uint64_t best = ~0ull;
        for (int32_t j = lane; j < n; j += 64) {
            if (j == i && !include_self) continue;
            const uint64_t key = ((uint64_t)(uint32_t)row[j] << 32) | (uint32_t)j;
            if ((r == 0 || key > last) && key < best) best = key;
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)best, off);
            const uint32_t hi = __shfl_xor((uint32_t)(best >> 32), off);
            const uint64_t o = ((uint64_t)hi << 32) | lo;
            best = o < best ? o : best;
        }
        last = best;
        if (lane == 0) out[r] = (int32_t)(uint32_t)best;
    }
}

__global__ __launch_bounds__(256) void k_nn_lists(const int32_t* __restrict__ dist, int32_t n, int32_t P,
                                                  int32_t k, int32_t* __restrict__ nn) {
    const int32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    nn_row(dist + (size_t)i * P, n, k, i, nn + (size_t)i * k, threadIdx.x & 63);
}

__global__ __launch_bounds__(256) void k_fill_rows(float* __restrict__ a, int32_t n, int32_t P, float v) {
    const int32_t i = blockIdx.y;
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j < P) a[(size_t)i * P + j] = (j < n) ? v : 0.0f;
}

// Nearest-neighbour tour from `start` (ties to the lowest index): C_nn of the
// initial-pheromone formulas (Colony::greedyDistance, a stub in the
// reference, src/Colony.cu:1053-1055).  One block; the visited bitmap in LDS.
// With candidate lists (nn, k < n, stable (d, j) order) wave 0 takes the
// first unvisited candidate of the current city, the same city the full
// (d, j) minimum would give, step after step without block barriers; only
// when every candidate is visited does the whole block scan the row.
__device__ void greedy_block(const int32_t* __restrict__ dist, int32_t n, int32_t P, int32_t start,
                             int64_t* __restrict__ out, uint32_t* vis, uint64_t* red, int32_t& s_next,
                             int32_t& s_step, const int32_t* __restrict__ nn, int32_t k) {
    const int32_t words = (n + 31) / 32;
    for (int32_t w = threadIdx.x; w < words; w += blockDim.x) vis[w] = 0;
    __syncthreads();
    if (threadIdx.x == 0) vis[start >> 5] |= 1u << (start & 31);
    __syncthreads();
    int64_t len = 0;  // thread 0
    int32_t c = start;
    int32_t s = 1;
    while (s < n) {
        if (nn != nullptr) {
            if (threadIdx.x < 64) {
                const int lane = threadIdx.x;
                while (s < n) {
                    const int32_t j = lane < k ? nn[(size_t)c * k + lane] : 0;
                    const unsigned long long m = __ballot(lane < k && !((vis[j >> 5] >> (j & 31)) & 1u));
                    if (!m) break;
                    const int32_t nx = readlane_i(j, __ffsll((long long)m) - 1);
                    if (lane == 0) {
                        len += dist[(size_t)c * P + nx];
                        vis[nx >> 5] |= 1u << (nx & 31);
                    }
                    wave_lds_sync();
                    c = nx;
                    ++s;
                }
                if (lane == 0) {
                    s_next = c;
                    s_step = s;
                }
            }
            __syncthreads();
            c = s_next;
            s = s_step;
            if (s >= n) break;
        }
        uint64_t best = ~0ull;
        for (int32_t j = threadIdx.x; j < n; j += blockDim.x) {
            if ((vis[j >> 5] >> (j & 31)) & 1u) continue;
            const uint64_t key = ((uint64_t)(uint32_t)dist[(size_t)c * P + j] << 32) | (uint32_t)j;
            best = key < best ? key : best;
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)best, off);
            const uint32_t hi = __shfl_xor((uint32_t)(best >> 32), off);
            const uint64_t o = ((uint64_t)hi << 32) | lo;
            best = o < best ? o : best;
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t b = red[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = red[w] < b ? red[w] : b;
            s_next = (int32_t)(uint32_t)b;
            len += (int64_t)(b >> 32);
            vis[s_next >> 5] |= 1u << (s_next & 31);
        }
        __syncthreads();
        c = s_next;
        ++s;
    }
    if (threadIdx.x == 0) *out = len + dist[(size_t)c * P + start];
}

__global__ __launch_bounds__(1024) void k_greedy(const int32_t* __restrict__ dist, int32_t n, int32_t P,
                                                 int32_t start, int64_t* __restrict__ out,
                                                 const int32_t* __restrict__ nn, int32_t k) {
    extern __shared__ __align__(16) uint32_t sh[];
    __shared__ int32_t s_next, s_step;
    const int32_t words = (n + 31) / 32;
    greedy_block(dist, n, P, start, out, sh, reinterpret_cast<uint64_t*>(sh + ((words + 1) & ~1)), s_next, s_step,
                 nn, k);
}

// ---------------------------------------------------------------------------
// Fused dense sweep, one wave per row (4 rows per 256-thread block), 16-byte
// coalesced loads and stores.  Per element:
//   choice   = tau^alpha * eta(d)          (dev_compute_selection, inc/Colony.cuh:91-99)
//   tau_next = max((1-rho) tau, 1e-18)     (dev_evaporate_pheromones, src/AntSystem.cu:26-32)
// and per row the mean of tau (dev_compute_pheromone_mean, inc/Colony.cuh:164-190),
// lane l summing columns 256s + 4l + q in order, lanes combined by the DPP scan.
// Algorithmic traffic: 16 bytes per element (read tau, dist; write choice, tau_next).
// One row, four 256-element chunks at a time: all tau/dist loads of the
// group are issued first, then the heuristic gathers, then the stores (the
// output pointers may alias nothing the group still reads).  The row-mean
// order is unchanged: lane l adds its columns 256s + 4l + q in s, q order.
// Top-list keys: a larger key is a larger weight, then a lower index.
__device__ __forceinline__ uint64_t top_key(float v, int32_t j) {
    return ((uint64_t)__float_as_uint(v) << 32) | (uint32_t)(0xffffffffu - (uint32_t)j);
}

// Descending bitonic sort of buf[0 .. TOP_BUF) by one wave.
__device__ __forceinline__ void wave_sort_desc(uint64_t* buf, int lane) {
    for (int k = 2; k <= TOP_BUF; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            wave_lds_sync();
#pragma unroll
            for (int t = 0; t < TOP_BUF / 64; ++t) {
                const int i = t * 64 + lane;
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t x = buf[i], y = buf[ixj];
                    const bool desc = (i & k) == 0;
                    if (desc ? (x < y) : (x > y)) {
                        buf[i] = y;
                        buf[ixj] = x;
                    }
                }
            }
        }
    }
    wave_lds_sync();
}

// Keeps the TOP_K largest keys of buf[0 .. cnt) in buf[0 .. TOP_K), sorted;
// returns the TOP_K-th weight (keys below it can no longer enter).
__device__ __forceinline__ float top_compact(uint64_t* buf, int32_t& cnt, int lane) {
    for (int32_t x = cnt + lane; x < TOP_BUF; x += 64) buf[x] = 0ull;
    wave_sort_desc(buf, lane);
    cnt = TOP_K;
    return __uint_as_float((uint32_t)(buf[TOP_K - 1] >> 32));
}

// Streaming selection of a row's top list during the sweep.  The previous
// sweep's TOP_K-th weight w, times (1 - rho) and a rounding margin, is a lower
// bound of this sweep's TOP_K-th weight (evaporation scales every weight by
// (1 - rho), deposits only add), so only weights >= that bound are kept;
// fewer than TOP_K of them means the bound did not hold (pheromone set from
// outside) and the row gets no list this iteration and a full selection the
// next.
struct TopSel {
    uint64_t* buf;
    float thr;
    int32_t cnt;
};

__device__ __forceinline__ void top_offer(TopSel& t, float v, int32_t j, bool in, int lane) {
    const bool pass = in && v >= t.thr;  // NaN never passes
    const unsigned long long m = __ballot(pass);
    if (!m) return;
    if (t.cnt + __popcll(m) > TOP_BUF) t.thr = fmaxf(t.thr, top_compact(t.buf, t.cnt, lane));
    if (pass) t.buf[t.cnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = top_key(v, j);
    t.cnt += __popcll(m);
    wave_lds_sync();
}

__device__ __forceinline__ void top_finish(const SweepArgs& a, TopSel& t, int32_t i, int lane) {
    uint16_t* out = a.top + (size_t)i * TOP_K;
    if (t.cnt >= TOP_K) {
        const float w = top_compact(t.buf, t.cnt, lane);
        for (int32_t r = lane; r < TOP_K; r += 64) out[r] = (uint16_t)(0xffffffffu - (uint32_t)t.buf[r]);
        if (lane == 0) a.top_thr[i] = w;
    } else {
        if (lane == 0) {
            out[0] = TOP_INVALID;
            a.top_thr[i] = 0.0f;
        }
    }
}

// One row's sweep state: pointers, the row sum (lane l adds its columns
// 256 s + 4 l + q of every 1024-column chunk in s, q order) and the top-list
// selection.
struct SweepRow {
    const float* __restrict__ trow;
    const int32_t* __restrict__ drow;
    const float* __restrict__ erow;  // eta_full row (null: from distances)
    float* __restrict__ crow;
    float* __restrict__ nrow;
    int32_t xi, yi;                  // the row city's coordinates (a.xy set)
    float acc;
    TopSel top;
    int32_t nnj;                     // candidate `lane` of the row and its heuristic, loaded
    float nne;                       // before the row (lanes < k)
};

__device__ __forceinline__ SweepRow sweep_row_begin(const SweepArgs& a, int32_t i, uint64_t* topbuf) {
    const size_t row = (size_t)i * a.P;
    SweepRow r;
    r.trow = a.tau + row;
    r.drow = a.dist + row;
    r.erow = a.eta_full != nullptr ? a.eta_full + row : nullptr;
    r.crow = a.choice + row;
    r.nrow = a.tau_next + row;
    r.xi = a.xy != nullptr ? a.xy[2 * i] : 0;
    r.yi = a.xy != nullptr ? a.xy[2 * i + 1] : 0;
    r.acc = 0.0f;
    r.top = TopSel{topbuf, 0.0f, 0};
    if (a.top != nullptr) {
        // lower bound of this sweep's TOP_K-th weight: (1 - rho) w (1 - u)^3
        // >= fl(fl(w (1 - rho)) (1 - 2^-20)) for the unit roundoff u = 2^-24
        r.top.thr = (a.top_thr[i] * a.one_minus_rho) * 0.99999904632568359375f;
    }
    return r;
}

__device__ __forceinline__ void sweep_row_end(const SweepArgs& a, int32_t i, int lane, SweepRow& r) {
    if (a.top != nullptr) top_finish(a, r.top, i, lane);
    if (a.want_mean) {
        const float total = readlane_f(wave_inclusive_scan(r.acc), 63);
        if (lane == 0) a.mean[i] = total / (float)a.n;
    }
    if (a.nn != nullptr) {
        for (int32_t q = lane; q < a.k; q += 64) {
            const int32_t j = q < 64 ? r.nnj : a.nn[(size_t)i * a.k + q];
            const float e = q < 64 ? r.nne : a.eta_nn[(size_t)i * a.k + q];
            const float t = r.trow[j];
            a.choice_nn[(size_t)i * a.k + q] = apply_alpha(t, a.alpha) * e;
            if (a.tau_nn != nullptr) a.tau_nn[(size_t)i * a.k + q] = t;
        }
    }
}

// Columns g0 .. g0 + 1023 of row i, any n (columns >= n are skipped): the
// batch's sweep and the last chunk of a single colony's rows.
// halt: the iteration is skipped; tested after the chunk's loads are issued
// and before anything is stored (false: nothing stored).
__device__ __forceinline__ bool sweep_chunk(const SweepArgs& a, int32_t i, int32_t g0, int lane, SweepRow& r,
                                            bool halt = false) {
    // Distances from the (L2-resident) coordinates instead of the matrix
    // row: 4n bytes of HBM reads become 8n bytes of L2 reads.
    const bool from_xy = r.erow == nullptr && a.xy != nullptr;
    float4 t[4];
    int4 d[4];
    float e[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int32_t j0 = g0 + 256 * s + 4 * lane;
        t[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        d[s] = make_int4(0, 0, 0, 0);
        e[s][0] = e[s][1] = e[s][2] = e[s][3] = 0.0f;
        if (j0 < a.n) {
            t[s] = *reinterpret_cast<const float4*>(r.trow + j0);
            if (r.erow != nullptr) {
                const float4 ev = *reinterpret_cast<const float4*>(r.erow + j0);
                e[s][0] = ev.x;
                e[s][1] = ev.y;
                e[s][2] = ev.z;
                e[s][3] = ev.w;
            } else if (from_xy) {
                // cities j0..j0+3 (the coordinate array is padded to a
                // multiple of four cities)
                const int4 p0 = *reinterpret_cast<const int4*>(a.xy + 2 * j0);
                const int4 p1 = *reinterpret_cast<const int4*>(a.xy + 2 * j0 + 4);
                const int32_t px[4] = {p0.x, p0.z, p1.x, p1.z}, py[4] = {p0.y, p0.w, p1.y, p1.w};
                int32_t dv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int32_t dq = euc2d_fast(r.xi - px[q], r.yi - py[q]);
                    dv[q] = (j0 + q == i) ? ACO_MAX_DISTANCE : dq;
                }
                d[s] = make_int4(dv[0], dv[1], dv[2], dv[3]);
            } else {
                d[s] = *reinterpret_cast<const int4*>(r.drow + j0);
            }
        }
    }
    if (r.erow == nullptr) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int32_t dv[4] = {d[s].x, d[s].y, d[s].z, d[s].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // 1 / d^beta computed (the table's value) rather than
                // gathered: the gathers were most of the sweep's L2 requests
                if (a.eta_pow)
                    e[s][q] = (uint32_t)dv[q] < (uint32_t)a.eta_len ? eta_power_value(a.eta_pow, dv[q]) : 0.0f;
                else
                    e[s][q] = eta_of(a.eta, a.eta_len, dv[q]);
            }
        }
    }
    if (halt) return false;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int32_t j0 = g0 + 256 * s + 4 * lane;
        if (j0 >= a.n) continue;
        const float tv[4] = {t[s].x, t[s].y, t[s].z, t[s].w};
        float cv[4], ev[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool in = (j0 + q) < a.n;
            const float x = in ? tv[q] : 0.0f;
            r.acc = r.acc + x;
            cv[q] = in ? apply_alpha(x, a.alpha) * e[s][q] : 0.0f;
            ev[q] = in ? fmaxf(a.one_minus_rho * x, ACO_MIN_PHEROMONE) : 0.0f;
        }
        *reinterpret_cast<float4*>(r.crow + j0) = make_float4(cv[0], cv[1], cv[2], cv[3]);
        *reinterpret_cast<float4*>(r.nrow + j0) = make_float4(ev[0], ev[1], ev[2], ev[3]);
    }
    if (a.top != nullptr) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int32_t j0 = g0 + 256 * s + 4 * lane;
            const float tv[4] = {t[s].x, t[s].y, t[s].z, t[s].w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool in = (j0 + q) < a.n;
                // the weight just stored (alpha = 1: tau times eta)
                const float v = in ? tv[q] * e[s][q] : 0.0f;
                top_offer(r.top, v, j0 + q, in, lane);
            }
        }
    }
    return true;
}

__device__ __forceinline__ void sweep_row(const SweepArgs& a, int32_t i, int lane, uint64_t* topbuf) {
    SweepRow r = sweep_row_begin(a, i, topbuf);
    r.nnj = 0;
    r.nne = 0.0f;
    if (a.nn != nullptr && lane < a.k) {
        r.nnj = a.nn[(size_t)i * a.k + lane];
        r.nne = a.eta_nn[(size_t)i * a.k + lane];
    }
    for (int32_t g0 = 0; g0 < a.n; g0 += 1024) sweep_chunk(a, i, g0, lane, r);
    sweep_row_end(a, i, lane, r);
}

// eta_power_value(PW, d) for d < eta_len (0 beyond: the diagonal), with the
// reciprocal of p = d^PW (an integer in [1, 2^53)) as the compiler's IEEE
// division computes it minus its scaling and special-case steps, which never
// act on such p: two Newton steps from v_rcp_f64, then the final
// residual-corrected product, rounded once to fp32.  The library checks it
// against the host table at creation (k_eta_check) and uses the table
// instead on any difference.
template <int PW>
__device__ __forceinline__ float eta_pow_fast(int32_t d, int32_t eta_len) {
    const double dd = (double)(d > 0 ? d : 1);
    double p = dd;
    if (PW >= 2) p = p * dd;
    if (PW >= 3) p = p * dd;
    double y = __builtin_amdgcn_rcp(p);
    double e = __builtin_fma(-p, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-p, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double res = __builtin_fma(-p, y, 1.0);
    const float v = (float)__builtin_fma(res, y, y);
    const float w = d == 0 ? (float)(1 << PW) : v;
    return (uint32_t)d < (uint32_t)eta_len ? w : 0.0f;
}

// Columns g0 .. g0 + 1023 < n of row i of a single colony, without range
// tests or branches per element: every load of the chunk is issued first;
// the heuristic comes from the eta_full row (SRC 0), the distance matrix
// row (SRC 1) or the coordinates (SRC 2), as eta_pow_fast<PW> (PW 1..3) or
// a table gather (PW 0).  Same values, stores and row-sum order as
// sweep_chunk.
template <int SRC, int PW, bool ALPHA1>
__device__ __forceinline__ bool sweep_chunk_full(const SweepArgs& a, int32_t i, int32_t g0, int lane, SweepRow& r,
                                                 bool halt = false) {
    // Non-temporal loads and stores for the row streams (each read or
    // written once per sweep); the coordinates are re-read by every row and
    // stay cached.
    typedef float f4v __attribute__((ext_vector_type(4)));
    typedef int i4v __attribute__((ext_vector_type(4)));
    float4 t[4], ef[4];
    int4 dl[4], p0[4], p1[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int32_t j0 = g0 + 256 * s + 4 * lane;
        const f4v tv = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(r.trow + j0));
        t[s] = make_float4(tv.x, tv.y, tv.z, tv.w);
        if (SRC == 0) {
            const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(r.erow + j0));
            ef[s] = make_float4(v.x, v.y, v.z, v.w);
        }
        if (SRC == 1) {
            const i4v v = __builtin_nontemporal_load(reinterpret_cast<const i4v*>(r.drow + j0));
            dl[s] = make_int4(v.x, v.y, v.z, v.w);
        }
        if (SRC == 2) {
            p0[s] = *reinterpret_cast<const int4*>(a.xy + 2 * j0);
            p1[s] = *reinterpret_cast<const int4*>(a.xy + 2 * j0 + 4);
        }
    }
    if (halt) return false;  // as sweep_chunk
    float w[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int32_t j0 = g0 + 256 * s + 4 * lane;
        const float tv[4] = {t[s].x, t[s].y, t[s].z, t[s].w};
        float e[4];
        if (SRC == 0) {
            e[0] = ef[s].x;
            e[1] = ef[s].y;
            e[2] = ef[s].z;
            e[3] = ef[s].w;
        } else {
            int32_t dv[4];
            if (SRC == 1) {
                dv[0] = dl[s].x;
                dv[1] = dl[s].y;
                dv[2] = dl[s].z;
                dv[3] = dl[s].w;
            } else {
                const int32_t px[4] = {p0[s].x, p0[s].z, p1[s].x, p1[s].z};
                const int32_t py[4] = {p0[s].y, p0[s].w, p1[s].y, p1[s].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int32_t dq = euc2d_fast(r.xi - px[q], r.yi - py[q]);
                    dv[q] = (j0 + q == i) ? ACO_MAX_DISTANCE : dq;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                e[q] = PW > 0 ? eta_pow_fast<PW>(dv[q], a.eta_len) : eta_of(a.eta, a.eta_len, dv[q]);
        }
        float cv[4], ev[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float x = tv[q];
            r.acc = r.acc + x;
            cv[q] = (ALPHA1 ? x : pow_alpha(x, a.alpha)) * e[q];
            ev[q] = fmaxf(a.one_minus_rho * x, ACO_MIN_PHEROMONE);
            w[s][q] = x * e[q];
        }
        __builtin_nontemporal_store(f4v{cv[0], cv[1], cv[2], cv[3]}, reinterpret_cast<f4v*>(r.crow + j0));
        __builtin_nontemporal_store(f4v{ev[0], ev[1], ev[2], ev[3]}, reinterpret_cast<f4v*>(r.nrow + j0));
    }
    if (a.top != nullptr) {
        // one ballot for the chunk: offers only where some weight reaches
        // the threshold (the kept set, the TOP_K largest keys, does not
        // depend on the order of the offers)
        bool any = false;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int q = 0; q < 4; ++q) any = any || w[s][q] >= r.top.thr;
        if (__ballot(any)) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int q = 0; q < 4; ++q) top_offer(r.top, w[s][q], g0 + 256 * s + 4 * lane + q, true, lane);
        }
    }
    return true;
}

template <int SRC, int PW, bool ALPHA1>
__device__ __forceinline__ void sweep_row_fast(const SweepArgs& a, int32_t i, int lane, uint64_t* topbuf,
                                               int32_t nnj, float nne, bool halt) {
    SweepRow r = sweep_row_begin(a, i, topbuf);
    r.nnj = nnj;
    r.nne = nne;
    const int32_t full_end = a.n & ~1023;
    for (int32_t g0 = 0; g0 < full_end; g0 += 1024)
        if (!sweep_chunk_full<SRC, PW, ALPHA1>(a, i, g0, lane, r, halt)) return;
    if (full_end < a.n && !sweep_chunk(a, i, full_end, lane, r, halt)) return;
    sweep_row_end(a, i, lane, r);
}

template <int SRC, int PW, bool ALPHA1>
__global__ __launch_bounds__(256) void k_sweep(const SweepArgs a) {
    extern __shared__ __align__(16) uint64_t topbuf[];  // TOP_BUF keys per wave when a.top is set
    const int lane = threadIdx.x & 63;
    stamp_start(a.stamp);
    const int32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
    // the row's candidate entries are requested first: their round trip
    // overlaps the halted flag's and the row's instead of following them
    int32_t nnj = 0;
    float nne = 0.0f;
    if (a.nn != nullptr && i < a.n && lane < a.k) {
        nnj = a.nn[(size_t)i * a.k + lane];
        nne = a.eta_nn[(size_t)i * a.k + lane];
    }
    // A halted run skips whole iterations: the sweep, first kernel of an
    // iteration, publishes the decision for the other three.  The flag is
    // tested after the row's loads are issued and before its first store,
    // so its round trip overlaps theirs.
    const int32_t h = a.state != nullptr ? a.state->halted : 0;
    if (i < a.n)
        sweep_row_fast<SRC, PW, ALPHA1>(a, i, lane, topbuf + (size_t)(threadIdx.x >> 6) * TOP_BUF, nnj, nne,
                                        h != 0);
    if (a.state != nullptr && blockIdx.x == 0 && threadIdx.x == 0) a.state->skip = h;
    if (h) return;
    // updateAntAges (src/Colony.cu:353-364): every ant ages by one at the
    // start of an iteration, before construction reads the ages.
    if (a.age != nullptr)
        for (int32_t k = blockIdx.x * 256 + threadIdx.x; k < a.m; k += gridDim.x * 256) a.age[k] += 1;
}

// Bit-for-bit check of eta_pow_fast<PW> against the table of
// eta_power_value: flag = 1 on any difference.
template <int PW>
__global__ __launch_bounds__(256) void k_eta_check(const float* __restrict__ eta, int32_t eta_len,
                                                   int32_t* __restrict__ flag) {
    const int32_t d = blockIdx.x * 256 + threadIdx.x;
    if (d < eta_len && __float_as_uint(eta_pow_fast<PW>(d, eta_len)) != __float_as_uint(eta[d])) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
// Construction helpers.
struct AntCtx {
    bool young;
    float ratio;
};

// compute_fittness (inc/Colony.cuh:14-31) for aging ants younger than
// min_age after convergence: tau pulled toward the row mean,
// ((tau - mean) * (age/min_age)) + mean, times eta.
__device__ __forceinline__ float young_formula(float tau, float mc, float ratio, float eta, float alpha) {
    const float p = ((tau - mc) * ratio) + mc;
    return apply_alpha(p, alpha) * eta;
}

__device__ __forceinline__ float young_weight(const ConstructArgs& a, int32_t c, int32_t j, float ratio) {
    const size_t xy = (size_t)c * a.P + j;
    const float e = a.eta_full != nullptr ? a.eta_full[xy] : eta_of(a.eta, a.eta_len, a.dist[xy]);
    return young_formula(a.tau[xy], a.mean[c], ratio, e, a.alpha);
}

__device__ __forceinline__ bool is_visited(const uint32_t* vis, int32_t j) {
    return (vis[j >> 5] >> (j & 31)) & 1u;
}

// Marks city j visited: a fire-and-forget LDS OR (ds_or_b32) by one lane;
// the wave's later LDS reads are ordered after it.
__device__ __forceinline__ void mark_visited(uint32_t* vis, int32_t j, int lane) {
    if (lane == 0) atomicOr(vis + (j >> 5), 1u << (j & 31));
}

// Tour staging in VGPRs: lane (s & 63) keeps the city of step s; a full chunk
// of 64 steps leaves as one coalesced store.
__device__ __forceinline__ void stage_city(int32_t* tour, int32_t& tbuf, int32_t s, int32_t city, int32_t n,
                                           int lane) {
    if (lane == (s & 63)) tbuf = city;
    if ((s & 63) == 63 || s == n - 1) {
        const int32_t base = s & ~63;
        if (lane <= (s & 63)) tour[base + lane] = tbuf;
    }
}

// Random words of one ant.  Lane l holds the four Philox words of block
// 64*(s/256) + l; step s uses word (s/64)%4 of lane s%64, so a step costs one
// lane read, the word switches every 64 steps and a refill every 256.
struct AntRng {
    uint32_t w0, w1, w2, w3;
    uint32_t cur;
};

__device__ __forceinline__ void rng_refill(AntRng& r, int32_t s, uint32_t ant, uint32_t iter, const ConstructArgs& a,
                                           int lane) {
    const uint32_t blk = (uint32_t)((s >> 8) << 6) + (uint32_t)lane;
    const Philox4 q = philox4x32_10(blk, ant, iter, a.stream, a.seed_lo, a.seed_hi);
    r.w0 = q.x0;
    r.w1 = q.x1;
    r.w2 = q.x2;
    r.w3 = q.x3;
    r.cur = q.x0;
}

// Before using step s >= 1 (uniform branch, taken every 64 steps).
__device__ __forceinline__ void rng_step(AntRng& r, int32_t s, uint32_t ant, uint32_t iter, const ConstructArgs& a,
                                         int lane) {
    if (__builtin_expect((s & 63) != 0, 1)) return;
    if ((s & 255) == 0) {
        rng_refill(r, s, ant, iter, a, lane);
    } else {  // words rotate: w1 serves steps 64..127 of the chunk, w2 the next 64, ...
        r.cur = r.w1;
        r.w1 = r.w2;
        r.w2 = r.w3;
    }
}

__device__ __forceinline__ uint32_t rng_word(const AntRng& r, int32_t s) {
    return (uint32_t)readlane_i((int)r.cur, s & 63);
}

// choose_best_next: argmax of the selection weight over every unvisited city,
// ties to the lowest index.  Lanes stride the row four loads at a time; exact
// comparisons make the result independent of the reduction order.
__device__ __forceinline__ int32_t fallback_argmax(const ConstructArgs& a, const uint32_t* vis, int32_t c,
                                                bool young, float ratio, int lane, unsigned long long& scanned,
                                                unsigned long long& listed) {
    float best = -1.0f;
    int32_t bj = 0x7fffffff;
    const float* crow = a.choice + (size_t)c * a.P;
    if (!young && a.top != nullptr) {
        // The row's top list holds its TOP_K largest weights in argmax order
        // (weight descending, index ascending), and every city outside it
        // comes later in that order: its first unvisited entry is the argmax
        // over all unvisited cities.
        const uint16_t* tl = a.top + (size_t)c * TOP_K;
        int32_t e[TOP_K / 64];
#pragma unroll
        for (int q = 0; q < TOP_K / 64; ++q) e[q] = __builtin_nontemporal_load(tl + 64 * q + lane);
        if (readlane_i(e[0], 0) != (int32_t)TOP_INVALID) {
#pragma unroll
            for (int q = 0; q < TOP_K / 64; ++q) {
                const unsigned long long m = __ballot(!is_visited(vis, e[q]));
                if (m) {
                    ++listed;
                    if (lane == 0) ++scanned;
                    return readlane_i(e[q], __ffsll((long long)m) - 1);
                }
            }
        }
    }
    if (!young) {
        // 16-byte coalesced loads, 16 per lane in flight (4096 cities per
        // batch, one memory round trip each); the row is padded to a
        // multiple of 64 floats.
        // Non-temporal loads: a fallback row is read once, and streaming it
        // through the caches would evict the candidate rows every step reads.
        typedef float f4v __attribute__((ext_vector_type(4)));
        for (int32_t j0 = 0; j0 < a.n; j0 += 4096) {
            float4 x[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int32_t j = j0 + 256 * q + 4 * lane;
                if (j < a.n) {
                    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(crow + j));
                    x[q] = make_float4(v.x, v.y, v.z, v.w);
                } else {
                    x[q] = make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int32_t j = j0 + 256 * q + 4 * lane;
                if (j < a.n) {
                    const uint32_t word = vis[j >> 5] >> (j & 31);  // 4 cities in one word
                    const float xv[4] = {x[q].x, x[q].y, x[q].z, x[q].w};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (j + t < a.n && !((word >> t) & 1u)) {
                            ++scanned;
                            if (xv[t] > best) {
                                best = xv[t];
                                bj = j + t;
                            }
                        }
                    }
                }
            }
        }
    } else {
        for (int32_t j = lane; j < a.n; j += 64) {
            if (is_visited(vis, j)) continue;
            ++scanned;
            const float x = young_weight(a, c, j, ratio);
            if (x > best) {
                best = x;
                bj = j;
            }
        }
    }
    int32_t next = wave_argmax_index(best, bj);
    if (next == 0x7fffffff) {  // only NaN weights left: lowest unvisited city
        for (int32_t j0 = 0; j0 < a.n; j0 += 64) {
            const int32_t j = j0 + lane;
            const unsigned long long m = __ballot(j < a.n && !is_visited(vis, j));
            if (m) {
                next = j0 + __ffsll((long long)m) - 1;
                break;
            }
        }
    }
    return next;
}

__device__ __forceinline__ void finish_counters(const ConstructArgs& a, unsigned long long fb,
                                                unsigned long long scanned, int lane, unsigned long long listed = 0) {
    for (int off = 32; off >= 1; off >>= 1) scanned += __shfl_xor(scanned, off);
    if (lane == 0) {
        if (fb) atomicAdd(&a.state->fb_steps, fb);
        if (scanned) atomicAdd(&a.state->fb_scanned, scanned);
        if (listed) atomicAdd(&a.state->fb_listed, listed);
    }
}

__device__ __forceinline__ bool ant_is_young(const ConstructArgs& a, int32_t ant, int32_t converged) {
    return a.aging && converged && a.valid[ant] && a.age[ant] < a.min_age;
}

// Wave-64 prefix sum restricted to 32 lanes: the first five steps of
// wave_inclusive_scan.  When lanes 32..63 hold zeros it yields the same bits
// in lanes 0..31 as the full scan, and lane 31 equals the full scan's lane 63.
__device__ __forceinline__ float wave_inclusive_scan32(float v) {
    v = v + dpp_src<0x111, 0xf>(v);
    v = v + dpp_src<0x112, 0xf>(v);
    v = v + dpp_src<0x114, 0xf>(v);
    v = v + dpp_src<0x118, 0xf>(v);
    v = v + dpp_src<0x142, 0xa>(v);
    return v;
}

// Exact tour length of one ant whose tour the wave has just stored: lane l
// sums the edges i -> i+1 for i = l, l + 64, ...; sixteen distance loads per
// lane are in flight at a time.
__device__ __forceinline__ void store_tour_length(const ConstructArgs& a, const int32_t* tour, int32_t ant,
                                                  int lane) {
    __threadfence_block();  // the wave's tour stores before its lanes read them back
    int64_t len = 0;
    for (int32_t i0 = 0; i0 < a.n; i0 += 64 * 16) {
        int32_t x[16], y[16], d[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int32_t i = i0 + 64 * q + lane;
            x[q] = i < a.n ? tour[i] : 0;
            y[q] = i < a.n ? tour[i + 1 == a.n ? 0 : i + 1] : 0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int32_t i = i0 + 64 * q + lane;
            d[q] = i < a.n ? a.dist[(size_t)x[q] * a.P + y[q]] : 0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) len += d[q];
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)len, off);
        const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)len >> 32), off);
        len += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    if (lane == 0) a.lengths[ant] = len;
}

// One roulette step: the first lane whose inclusive prefix of the weights w
// (0 for visited or absent candidates) exceeds u * total and whose weight is
// positive gives the next city j (selectCities src/Colony.cu:821-851); with
// no such lane the last positive lane (u * total rounded up to the total),
// and with a zero total the fallback's city.  Ballots come straight from
// the compare masks; the common case falls through and the rare ones are
// laid apart.
template <bool K32, typename Fallback>
__device__ __forceinline__ int32_t roulette_pick(float w, int32_t j, float u, Fallback&& fallback) {
    const unsigned long long pos = __builtin_amdgcn_ballot_w64(w > 0.0f);
    const float p = K32 ? wave_inclusive_scan32(w) : wave_inclusive_scan(w);
    // u * total read from the last lane of p * u: one lane read instead of a
    // read and a move before the multiply, the same product.  The empty asm
    // keeps the compiler from folding the multiply back behind the read.
    float pu = p * u;
    asm volatile("" : "+v"(pu));
    const float rr = readlane_f(pu, K32 ? 31 : 63);
    const unsigned long long hit = __builtin_amdgcn_ballot_w64(p > rr) & pos;
    // s_ff1 gives the lowest set bit, -1 for no hit (the lane read then
    // selects lane 63 and is replaced below)
    int32_t first;
    asm("s_ff1_i32_b64 %0, %1" : "=s"(first) : "s"(hit));
    int32_t next = readlane_i(j, first);
    if (__builtin_expect(hit == 0ull, 0)) {
        const float total = readlane_f(p, K32 ? 31 : 63);
        next = total > 0.0f ? readlane_i(j, 63 - __clzll((long long)pos)) : fallback();
    }
    return next;
}

// ---------------------------------------------------------------------------
// Tour construction, candidate-list mode, global-memory variant (any n).
// One ant per wave64 for all n-1 steps (the reference launches 3(n-1)
// kernels per iteration, src/Colony.cu:713-755).  Per step: lane r < k holds
// candidate r of the current city and its weight (0 if visited, per the LDS
// bitmap); the DPP scan gives the roulette prefix and a ballot the pick; a
// zero total falls back to the argmax over the unvisited cities.  The next
// city's candidate row is requested as soon as the pick is known, so its
// memory latency overlaps the step's bookkeeping, and no other vector load
// sits in the per-step chain: vector memory returns in issue order, so a
// distance load per step would make every candidate load wait for it.  The
// tour length is summed after the last step from the stored tour.
// K32: k <= 32, the roulette scan stops at 32 lanes.
template <bool K32>
__device__ __forceinline__ void construct_cand_ant(const ConstructArgs& a, int32_t ant, bool young, float ratio,
                                                   uint32_t* vis, uint32_t iter, int lane) {
    for (int32_t w = lane; w < a.words; w += 64) vis[w] = 0u;
    AntRng rng;
    rng_refill(rng, 0, (uint32_t)ant, iter, a, lane);
    const int32_t start = word_to_city(rng_word(rng, 0), a.n);
    mark_visited(vis, start, lane);
    const bool in = lane < a.k;
    // lanes >= k keep candidate 0 with weight 0
    auto load_row = [&](int32_t city, int32_t& j, float& cw) {
        j = 0;
        cw = 0.0f;
        if (in) {
            const size_t e = (size_t)city * a.k + lane;
            j = a.nn16 != nullptr ? (int32_t)a.nn16[e] : a.nn[e];
            cw = young ? young_formula(a.tau_nn[e], a.mean[city], ratio, a.eta_nn[e], a.alpha) : a.choice_nn[e];
        }
    };
    int32_t* tour = a.tours + (size_t)ant * a.n;
    int32_t tbuf = start;
    int32_t c = start;
    int32_t j;
    float cw;
    load_row(start, j, cw);
    unsigned long long fb = 0, scanned = 0, listed = 0;
    // steps in chunks of 64 (see construct_staged_ant)
    for (int32_t s0 = 0; s0 < a.n; s0 += 64) {
        if (s0 > 0) rng_step(rng, s0, (uint32_t)ant, iter, a, lane);
        const float uc = word_to_uniform(rng.cur);
        const int32_t s_end = min(s0 + 64, a.n);
        for (int32_t s = s0 > 0 ? s0 : 1; s < s_end; ++s) {
            const float u = readlane_f(uc, s & 63);
            const float w = is_visited(vis, j) ? 0.0f : cw;
            const int32_t next = roulette_pick<K32>(w, j, u, [&]() {
                ++fb;
                return fallback_argmax(a, vis, c, young, ratio, lane, scanned, listed);
            });
            load_row(next, j, cw);
            mark_visited(vis, next, lane);
            tbuf = (lane == s - s0) ? next : tbuf;
            c = next;
        }
        // non-temporal: the tours (m x n x 4 bytes per iteration) would
        // otherwise stream through L2 past the candidate rows every step reads
        if (lane < s_end - s0) __builtin_nontemporal_store(tbuf, tour + s0 + lane);
    }
    finish_counters(a, fb, scanned, lane, listed);
    store_tour_length(a, tour, ant, lane);
}

__global__ __launch_bounds__(512) void k_construct_cand(const ConstructArgs a) {
    extern __shared__ __align__(16) uint32_t lds[];
    stamp_start(a.stamp);
    if (a.state->skip) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int32_t ant = blockIdx.x * (blockDim.x >> 6) + wave;
    if (ant >= a.m) return;
    if (a.aging && !a.valid[ant]) return;
    uint32_t* vis = lds + (size_t)wave * a.words;
    const uint32_t iter = a.state->iter;
    const bool young = ant_is_young(a, ant, a.state->converged);
    const float ratio = young ? (float)a.age[ant] / (float)a.min_age : 0.0f;
    if (a.k <= 32)
        construct_cand_ant<true>(a, ant, young, ratio, vis, iter, lane);
    else
        construct_cand_ant<false>(a, ant, young, ratio, vis, iter, lane);
}

// ---------------------------------------------------------------------------
// Tour construction, candidate-list mode, LDS-staged variant (n*k*6 bytes fit
// in LDS).  Every block first stages the candidate lists (16-bit) and their
// weights for the whole instance, so each step reads only LDS.  Normal
// blocks stage choice_nn and run W ants each; the trailing young_blocks
// blocks each take one aging ant younger than min_age after convergence and
// stage that ant's pulled-toward-the-mean weights instead (they differ per
// age), so young ants also run at LDS speed.  Each 64-step chunk of a tour
// is written when it ends, and the lengths of its edges are loaded then and
// summed a chunk later (exact int64 tour lengths).

// Argmax fallback with the ant's bitmap in registers (lane src + w holds word
// w), spread over the whole wave: lane l reads cities CPL l .. CPL l + CPL - 1
// (CPL / 4 16-byte loads per row; n <= 64 CPL; rows are padded to a multiple
// of 64) and takes their bits from the owner lane with one bpermute.  The
// weight is choice, or for a young aging ant the pulled-toward-the-mean
// pheromone times the heuristic (per-edge heuristic rows when the colony
// keeps them, else the distance row and table gathers).  Each lane keeps
// the largest weight of its unvisited cities (strictly greater in ascending
// order: lowest index on ties, a NaN weight never wins) and the wave takes
// the largest (weight, lowest index) key: the same result as the sequential
// scan.  Call with every lane active.
template <int CPL>
__device__ __forceinline__ int32_t fallback_argmax_spread(const ConstructArgs& a, uint32_t vreg, int src, int32_t c,
                                                          bool young, float ratio, int lane,
                                                          unsigned long long& scanned) {
    const int32_t j0 = CPL * lane;
    const uint32_t word = (uint32_t)__builtin_amdgcn_ds_bpermute((src + (j0 >> 5)) << 2, (int)vreg);
    uint32_t free_bits = 0u;
    float best = -1.0f;
    int32_t bj = 0x7fffffff;
    if (j0 < a.n) {
        constexpr uint32_t FULL = CPL == 32 ? 0xffffffffu : ((1u << CPL) - 1u);
        free_bits = ~(word >> (j0 & 31)) & FULL;
        if (a.n - j0 < CPL) free_bits &= (1u << (a.n - j0)) - 1u;
        scanned += __popc(free_bits);
        const size_t row = (size_t)c * a.P + j0;
        float v[CPL];
        if (!young) {
            const float4* src4 = reinterpret_cast<const float4*>(a.choice + row);
#pragma unroll
            for (int q = 0; q < CPL / 4; ++q) {
                const float4 t = src4[q];
                v[4 * q] = t.x;
                v[4 * q + 1] = t.y;
                v[4 * q + 2] = t.z;
                v[4 * q + 3] = t.w;
            }
        } else {
            // all loads first, then the weights
            const float mc = a.mean[c];
            float t[CPL], e[CPL];
            const float4* t4 = reinterpret_cast<const float4*>(a.tau + row);
            if (a.eta_full != nullptr) {
                const float4* e4 = reinterpret_cast<const float4*>(a.eta_full + row);
#pragma unroll
                for (int q = 0; q < CPL / 4; ++q) {
                    const float4 tv = t4[q];
                    const float4 ev = e4[q];
                    t[4 * q] = tv.x;
                    t[4 * q + 1] = tv.y;
                    t[4 * q + 2] = tv.z;
                    t[4 * q + 3] = tv.w;
                    e[4 * q] = ev.x;
                    e[4 * q + 1] = ev.y;
                    e[4 * q + 2] = ev.z;
                    e[4 * q + 3] = ev.w;
                }
            } else {
                int32_t d[CPL];
                const int4* d4 = reinterpret_cast<const int4*>(a.dist + row);
#pragma unroll
                for (int q = 0; q < CPL / 4; ++q) {
                    const float4 tv = t4[q];
                    const int4 dv = d4[q];
                    t[4 * q] = tv.x;
                    t[4 * q + 1] = tv.y;
                    t[4 * q + 2] = tv.z;
                    t[4 * q + 3] = tv.w;
                    d[4 * q] = dv.x;
                    d[4 * q + 1] = dv.y;
                    d[4 * q + 2] = dv.z;
                    d[4 * q + 3] = dv.w;
                }
#pragma unroll
                for (int q = 0; q < CPL; ++q) e[q] = eta_of(a.eta, a.eta_len, d[q]);
            }
            if (a.alpha == 1.0f) {  // every city's weight without branches, the visited ones masked below
#pragma unroll
                for (int q = 0; q < CPL; ++q) v[q] = young_formula(t[q], mc, ratio, e[q], 1.0f);
            } else {
#pragma unroll
                for (int q = 0; q < CPL; ++q)
                    v[q] = ((free_bits >> q) & 1u) ? young_formula(t[q], mc, ratio, e[q], a.alpha) : -1.0f;
            }
        }
        // visited cities weigh -1 (never above the initial best); ascending
        // cities, strictly greater: the lowest index among equal maxima, and
        // a NaN weight never wins.  Selects, not branches: a branch per city
        // costs two exec-mask updates and a jump.
#pragma unroll
        for (int q = 0; q < CPL; ++q) v[q] = ((free_bits >> q) & 1u) ? v[q] : -1.0f;
        int32_t bq = 0;  // offset in the lane's cities (inline constants; j0 added once)
#pragma unroll
        for (int q = 0; q < CPL; ++q) {
            const bool gt = v[q] > best;
            best = gt ? v[q] : best;
            bq = gt ? q : bq;
        }
        bj = j0 + bq;  // read only when best > -1
    }
    int32_t next = wave_argmax_index(best, bj);
    if (next == 0x7fffffff) {  // only NaN weights left: lowest unvisited city
        const unsigned long long m = __ballot(free_bits != 0u);
        const int l = __ffsll((long long)m) - 1;
        next = CPL * l + __ffs(readlane_i((int)free_bits, l)) - 1;
    }
    return next;
}

// The register-bitmap fallback at the narrowest spread that covers the row
// (n <= 64 MAXCPL; wave-uniform).
template <int MAXCPL = 32>
__device__ __forceinline__ int32_t fallback_argmax_regs(const ConstructArgs& a, uint32_t vreg, int src, int32_t c,
                                                        bool young, float ratio, int lane,
                                                        unsigned long long& scanned) {
    if (MAXCPL == 8 || a.n <= 512) return fallback_argmax_spread<8>(a, vreg, src, c, young, ratio, lane, scanned);
    if (MAXCPL == 16 || a.n <= 1024)
        return fallback_argmax_spread<16>(a, vreg, src, c, young, ratio, lane, scanned);
    return fallback_argmax_spread<32>(a, vreg, src, c, young, ratio, lane, scanned);
}

// One ant on LDS-staged candidate data.  VREG_VIS: the visited set lives in
// one VGPR per lane (lane l holds cities 32l..32l+31; n <= 2048) and is read
// across lanes with ds_bpermute; otherwise an LDS bitmap.  K32: k <= 32, the
// roulette scan stops at 32 lanes.  The common step is: two LDS reads, the
// visited test, the scan, one ballot and one lane read.
template <bool VREG_VIS, bool K32>
__device__ __forceinline__ void construct_staged_ant(const ConstructArgs& a, int32_t ant, bool young, float ratio,
                                                     const float* wl, const uint16_t* nl, uint32_t* vis_lds,
                                                     uint32_t iter, int lane) {
    uint32_t vreg = 0u;
    if (!VREG_VIS)
        for (int32_t w = lane; w < a.words; w += 64) vis_lds[w] = 0u;
    AntRng rng;
    rng_refill(rng, 0, (uint32_t)ant, iter, a, lane);
    const int32_t start = word_to_city(rng_word(rng, 0), a.n);
    if (VREG_VIS) {
        vreg |= (lane == (start >> 5)) ? (1u << (start & 31)) : 0u;
    } else {
        mark_visited(vis_lds, start, lane);
    }
    // lanes >= k read the zero weight staged after the n*k candidate entries
    // row stride per lane (0 for lanes >= k): a 24-bit multiply, n k < 2^24
    const uint32_t kk = lane < a.k ? (uint32_t)a.k : 0u;
    const int32_t loff = lane < a.k ? lane : a.n * a.k;
    int32_t tbuf = start;
    int32_t c = start;
    unsigned long long fb = 0, scanned = 0, listed = 0;
    // the candidate row of the current city is loaded one step ahead: the
    // LDS reads for step s+1 are issued as soon as step s has chosen, before
    // its bookkeeping
    // per-lane row bases; an entry address is one 24-bit multiply-add
    const char* nlb = reinterpret_cast<const char*>(nl + loff);
    const char* wlb = reinterpret_cast<const char*>(wl + loff);
    const uint32_t kk2 = 2u * kk, kk4 = 4u * kk;
    int32_t j = *reinterpret_cast<const uint16_t*>(nlb + __umul24((uint32_t)c, kk2));
    float cw = *reinterpret_cast<const float*>(wlb + __umul24((uint32_t)c, kk4));
    // steps in chunks of 64: a chunk's random words sit one per lane in
    // rng.cur and its cities one per lane in tbuf, so the step itself tests
    // neither the refill nor the tour flush
    int32_t* tour = a.tours + (size_t)ant * a.n;
    int32_t prev_last = start;  // the city before the chunk's first position
    int32_t dpend = 0;          // the previous chunk's edge lengths, loads in flight
    int64_t len = 0;
    for (int32_t s0 = 0; s0 < a.n; s0 += 64) {
        if (s0 > 0) rng_step(rng, s0, (uint32_t)ant, iter, a, lane);
        const float uc = word_to_uniform(rng.cur);  // the chunk's uniforms, one per lane
        const int32_t s_end = min(s0 + 64, a.n);
        for (int32_t s = s0 > 0 ? s0 : 1; s < s_end; ++s) {
            const float u = readlane_f(uc, s & 63);
            __builtin_assume((uint32_t)j < 65536u);  // a 16-bit LDS entry
            float w;
            if (VREG_VIS) {
                // ds_bpermute takes a byte address and drops its two low bits:
                // j >> 3 addresses lane j >> 5, the owner of j's bitmap word; the
                // signed bit field extract of bit j mod 32 gives an all-ones mask
                // for a visited city, which clears the weight's bits
                const int32_t word = __builtin_amdgcn_ds_bpermute(j >> 3, (int)vreg);
                // the current city joins the set only now, while the lane
                // read is in flight: it is none of its own candidates
                if (lane == (c >> 5)) vreg |= 1u << (c & 31);
                const int32_t vmask = __builtin_amdgcn_sbfe(word, (uint32_t)j, 1u);
                w = __int_as_float(__float_as_int(cw) & ~vmask);
            } else {
                w = is_visited(vis_lds, j) ? 0.0f : cw;
                mark_visited(vis_lds, c, lane);  // likewise after the read
            }
            const int32_t next = roulette_pick<K32>(w, j, u, [&]() {
                ++fb;
                if (!VREG_VIS) return fallback_argmax(a, vis_lds, c, young, ratio, lane, scanned, listed);
                return fallback_argmax_regs(a, vreg, 0, c, young, ratio, lane, scanned);
            });
            const int32_t jn = *reinterpret_cast<const uint16_t*>(nlb + __umul24((uint32_t)next, kk2));
            const float cwn = *reinterpret_cast<const float*>(wlb + __umul24((uint32_t)next, kk4));
            tbuf = (lane == s - s0) ? next : tbuf;
            c = next;
            j = jn;
            cw = cwn;
        }
        // The chunk's tour entries and the lengths of its edges (position
        // p - 1 -> p, one per lane) are requested now and summed at the end
        // of the next chunk, so the loads complete during its steps instead
        // of one memory round trip per 64 edges after the last step.
        const int32_t cnt = s_end - s0;
        int32_t pred = __shfl_up(tbuf, 1);
        pred = lane == 0 ? prev_last : pred;
        len += dpend;
        dpend = (lane < cnt && s0 + lane > 0) ? a.dist[(size_t)pred * a.P + tbuf] : 0;
        if (lane < cnt) __builtin_nontemporal_store(tbuf, tour + s0 + lane);
        prev_last = readlane_i(tbuf, cnt - 1);
    }
    finish_counters(a, fb, scanned, lane, listed);
    // the closing edge, last city -> start
    len += dpend;
    if (lane == 0) len += a.dist[(size_t)prev_last * a.P + start];
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)len, off);
        const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)len >> 32), off);
        len += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    if (lane == 0) a.lengths[ant] = len;
}

// Dispatch on the register-bitmap and 32-lane cases (wave-uniform).
__device__ __forceinline__ void construct_staged_dispatch(const ConstructArgs& a, int32_t ant, bool young,
                                                          float ratio, const float* wl, const uint16_t* nl,
                                                          uint32_t* vis, uint32_t iter, int lane) {
    if (a.n <= 2048) {
        if (a.k <= 32)
            construct_staged_ant<true, true>(a, ant, young, ratio, wl, nl, vis, iter, lane);
        else
            construct_staged_ant<true, false>(a, ant, young, ratio, wl, nl, vis, iter, lane);
    } else {
        if (a.k <= 32)
            construct_staged_ant<false, true>(a, ant, young, ratio, wl, nl, vis, iter, lane);
        else
            construct_staged_ant<false, false>(a, ant, young, ratio, wl, nl, vis, iter, lane);
    }
}

// LDS layout of the staged construction: weights (n*k fp32, then one zero
// entry that lanes >= k read), candidate lists (n*k uint16 and one entry),
// each region padded to 16 bytes, then per wave the visited bitmap and the
// tour.  staged_lds_bytes() is the host-side size of the same layout.
struct StagedLayout {
    float* wl;
    uint16_t* nl;
    uint32_t* per;
};

__device__ __forceinline__ StagedLayout staged_layout(uint32_t* lds, int32_t nk) {
    StagedLayout L;
    L.wl = reinterpret_cast<float*>(lds);
    L.nl = reinterpret_cast<uint16_t*>(L.wl + ((nk + 4) & ~3));
    L.per = reinterpret_cast<uint32_t*>(L.nl + ((nk + 8) & ~7));
    return L;
}

// The entries read by lanes >= k; staging never writes index nk.
__device__ __forceinline__ void staged_zero_entry(const StagedLayout& L, int32_t nk) {
    if (threadIdx.x == 0) {
        L.wl[nk] = 0.0f;
        L.nl[nk] = 0;
    }
}

// Block-wide copies global -> LDS, 16 bytes per thread and load where the
// addresses allow (hipMalloc buffers are 256-byte aligned).
// 16-byte vectors, four loads in flight per thread before the LDS stores.
template <typename V>
__device__ __forceinline__ void stage_vec(V* d4, const V* __restrict__ s4, int32_t v) {
    const int32_t stride = blockDim.x;
    int32_t x = threadIdx.x;
    for (; x + 3 * stride < v; x += 4 * stride) {
        const V a0 = s4[x], a1 = s4[x + stride], a2 = s4[x + 2 * stride], a3 = s4[x + 3 * stride];
        d4[x] = a0;
        d4[x + stride] = a1;
        d4[x + 2 * stride] = a2;
        d4[x + 3 * stride] = a3;
    }
    for (; x < v; x += stride) d4[x] = s4[x];
}

__device__ __forceinline__ void stage_u16(uint16_t* dst, const uint16_t* __restrict__ src, int32_t count) {
    const int32_t v = count / 8;
    stage_vec(reinterpret_cast<uint4*>(dst), reinterpret_cast<const uint4*>(src), v);
    for (int32_t x = 8 * v + threadIdx.x; x < count; x += blockDim.x) dst[x] = src[x];
}

// The candidate lists (count uint16) and their weights (count fp32) in one
// pass: up to 16 loads of 16 bytes per thread in flight, from both tables,
// before the first LDS store, so a block of 512 threads stages an n = 1000,
// k = 20 instance (120 KB) in one memory round trip instead of one per four
// loads per thread and table.
__device__ __forceinline__ void stage_lists_weights(uint16_t* nl, const uint16_t* __restrict__ nn16, float* wl,
                                                    const float* __restrict__ w, int32_t count) {
    constexpr int U = 16;
    const int32_t v1 = count / 8, v2 = count / 4, v = v1 + v2;
    const uint4* s1 = reinterpret_cast<const uint4*>(nn16);
    const uint4* s2 = reinterpret_cast<const uint4*>(w);
    uint4* d1 = reinterpret_cast<uint4*>(nl);
    uint4* d2 = reinterpret_cast<uint4*>(wl);
    const int32_t stride = blockDim.x;
    for (int32_t x0 = threadIdx.x; x0 < v; x0 += U * stride) {
        uint4 t[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {  // indices past the end reload the last vector
            const int32_t x = min(x0 + q * stride, v - 1);
            t[q] = *(x < v1 ? s1 + x : s2 + (x - v1));
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int32_t x = x0 + q * stride;
            if (x < v1)
                d1[x] = t[q];
            else if (x < v)
                d2[x - v1] = t[q];
        }
    }
    for (int32_t x = 8 * v1 + threadIdx.x; x < count; x += stride) nl[x] = nn16[x];
    for (int32_t x = 4 * v2 + threadIdx.x; x < count; x += stride) wl[x] = w[x];
}

// Young-ant weights for the whole instance, four candidates per 16-byte
// vector.  A thread first issues the loads of U vectors (tau and heuristic)
// and of their rows' means, then computes and stores them: one memory round
// trip per U vectors per thread instead of one per mean, which kept the
// young ant's block from starting its tour for tens of microseconds.
__device__ __forceinline__ void stage_young(float* wl, const ConstructArgs& a, float ratio, int32_t nk) {
    constexpr int U = 8;
    const int32_t v = nk / 4;
    const float4* t4 = reinterpret_cast<const float4*>(a.tau_nn);
    const float4* e4 = reinterpret_cast<const float4*>(a.eta_nn);
    float4* w4 = reinterpret_cast<float4*>(wl);
    const int32_t stride = blockDim.x;
    for (int32_t x0 = threadIdx.x; x0 < v; x0 += U * stride) {
        float4 t[U], e[U];
        float mc[U][4];
#pragma unroll
        for (int q = 0; q < U; ++q) {  // indices past the end reload the last vector
            const int32_t x = min(x0 + q * stride, v - 1);
            t[q] = t4[x];
            e[q] = e4[x];
#pragma unroll
            for (int i = 0; i < 4; ++i) mc[q][i] = a.mean[(4 * x + i) / a.k];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int32_t x = x0 + q * stride;
            if (x < v) {
                float4 w;
                w.x = young_formula(t[q].x, mc[q][0], ratio, e[q].x, a.alpha);
                w.y = young_formula(t[q].y, mc[q][1], ratio, e[q].y, a.alpha);
                w.z = young_formula(t[q].z, mc[q][2], ratio, e[q].z, a.alpha);
                w.w = young_formula(t[q].w, mc[q][3], ratio, e[q].w, a.alpha);
                w4[x] = w;
            }
        }
    }
    for (int32_t x = 4 * v + threadIdx.x; x < nk; x += stride)
        wl[x] = young_formula(a.tau_nn[x], a.mean[x / a.k], ratio, a.eta_nn[x], a.alpha);
}

// Index of the y-th young ant (ant order), -1 if there are at most y; wave
// 0 only.  Each lane loads the validity and age of 16 ants at a time before
// any ballot, one memory round trip per 1024 ants.
__device__ __forceinline__ int32_t find_young_ant(const ConstructArgs& a, int32_t y, int32_t converged, int lane) {
    int32_t seen = 0;
    for (int32_t a0 = 0; a0 < a.m; a0 += 1024) {
        int32_t vl[16], ag[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {  // both loads unconditional (ants past m read ant m - 1)
            const int32_t k = min(a0 + 64 * q + lane, a.m - 1);
            vl[q] = a.valid[k];
            ag[q] = a.age[k];
        }
        bool yg[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)  // ant_is_young
            yg[q] = a0 + 64 * q + lane < a.m && a.aging && converged && vl[q] && ag[q] < a.min_age;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            unsigned long long mk = __ballot(yg[q]);
            const int cnt = __popcll(mk);
            if (seen + cnt > y) {
                for (int t = 0; t < y - seen; ++t) mk &= mk - 1;
                return a0 + 64 * q + __ffsll((long long)mk) - 1;
            }
            seen += cnt;
        }
    }
    return -1;
}

__global__ __launch_bounds__(512) void k_construct_staged(const ConstructArgs a) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ int32_t s_ant;
    stamp_start(a.stamp);
    // a skipped iteration builds nothing; normal blocks stage first (LDS
    // only) and test the flag after, so its round trip overlaps the staging
    const int32_t skip = a.state->skip;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int32_t nk = a.n * a.k;
    const StagedLayout L = staged_layout(lds, nk);
    float* wl = L.wl;
    uint16_t* nl = L.nl;
    uint32_t* vis = L.per + (size_t)wave * a.words;
    const uint32_t iter = a.state->iter;
    const int32_t converged = a.state->converged;
    const int32_t normal_blocks = gridDim.x - a.young_blocks;
    const bool young_block = (int32_t)blockIdx.x >= normal_blocks;
    if (young_block && (skip || !a.aging || !converged)) return;

    // Normal blocks run one round (W ants).
    if (!young_block) {
        stage_lists_weights(nl, a.nn16, wl, a.choice_nn, nk);
        staged_zero_entry(L, nk);
        __syncthreads();
        if (skip) return;
        const int32_t ant = blockIdx.x * nwaves + wave;
        if (ant < a.m && !(a.aging && (!a.valid[ant] || ant_is_young(a, ant, converged))))
            construct_staged_dispatch(a, ant, false, 0.0f, wl, nl, vis, iter, lane);
        return;
    }

    // Young block y runs the young ants y, y + young_blocks, ... one per
    // round on wave 0.
    stage_u16(nl, a.nn16, nk);
    staged_zero_entry(L, nk);
    for (int32_t y = blockIdx.x - normal_blocks;; y += a.young_blocks) {
        if (wave == 0) {
            const int32_t found = find_young_ant(a, y, converged, lane);
            if (lane == 0) s_ant = found;
        }
        __syncthreads();
        const int32_t ant = s_ant;
        if (ant < 0) return;
        const float ratio = (float)a.age[ant] / (float)a.min_age;
        stage_young(wl, a, ratio, nk);
        __syncthreads();
        if (wave == 0) construct_staged_dispatch(a, ant, true, ratio, wl, nl, vis, iter, lane);
        __syncthreads();
    }
}

// 16-bit candidate lists and the heuristic on candidate edges (static; built
// once after the NN lists).
__global__ __launch_bounds__(256) void k_nn_aux(const int32_t* __restrict__ nn, const int32_t* __restrict__ dist,
                                                const float* __restrict__ eta, int32_t eta_len, int32_t n, int32_t P,
                                                int32_t k, uint16_t* __restrict__ nn16, float* __restrict__ eta_nn) {
    const int32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= n * k) return;
    const int32_t i = x / k;
    const int32_t j = nn[x];
    if (nn16) nn16[x] = (uint16_t)j;
    eta_nn[x] = eta_of(eta, eta_len, dist[(size_t)i * P + j]);
}

// ---------------------------------------------------------------------------
// Tour construction, full-row mode: the reference's candidate set
// (calculateProbabilityOfSelection src/Colony.cu:795-818: every city in the
// row's NN order, visited ones weighted 0).  The row is processed in chunks
// of 64 lanes; chunk prefixes come from the DPP scan and are offset by the
// running chunk total T (sequential over chunks).  Fallback: the farthest
// unvisited city (select_local_best inc/Colony.cuh:33-43).  LDS per ant:
// bitmap, weights and prefixes of the whole row, chunk offsets.
// Selection weights of row c in city order into rowbuf (LDS), 0 for visited
// cities: eight 16-byte loads per lane in flight (2048 cities per pass).
// Young aging ants get the pulled-toward-the-mean weights.
__device__ __forceinline__ void stage_weight_row(const ConstructArgs& a, int32_t c, const AntCtx& ac,
                                                 const uint32_t* vis, float* rowbuf, int lane) {
    const size_t row = (size_t)c * a.P;
    for (int32_t b0 = 0; b0 < a.n; b0 += 2048) {
        float4 x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int32_t j = b0 + 256 * q + 4 * lane;
            x[q] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < a.n) {
                if (!ac.young) {
                    x[q] = *reinterpret_cast<const float4*>(a.choice + row + j);
                } else {
                    const float4 t = *reinterpret_cast<const float4*>(a.tau + row + j);
                    const int4 d = *reinterpret_cast<const int4*>(a.dist + row + j);
                    const float mc = a.mean[c];
                    x[q].x = young_formula(t.x, mc, ac.ratio, eta_of(a.eta, a.eta_len, d.x), a.alpha);
                    x[q].y = young_formula(t.y, mc, ac.ratio, eta_of(a.eta, a.eta_len, d.y), a.alpha);
                    x[q].z = young_formula(t.z, mc, ac.ratio, eta_of(a.eta, a.eta_len, d.z), a.alpha);
                    x[q].w = young_formula(t.w, mc, ac.ratio, eta_of(a.eta, a.eta_len, d.w), a.alpha);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int32_t j = b0 + 256 * q + 4 * lane;
            if (j < a.n) {
                const uint32_t word = vis[j >> 5] >> (j & 31);  // cities j .. j+3 share a word
                float4 v = x[q];
                v.x = (word & 1u) ? 0.0f : v.x;
                v.y = (word & 2u) ? 0.0f : v.y;
                v.z = (word & 4u) ? 0.0f : v.z;
                v.w = (word & 8u) ? 0.0f : v.w;
                *reinterpret_cast<float4*>(rowbuf + j) = v;
            }
        }
    }
}

__global__ __launch_bounds__(512) void k_construct_full(const ConstructArgs a) {
    extern __shared__ __align__(16) uint32_t lds[];
    stamp_start(a.stamp);
    if (a.state->skip) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int32_t ant = blockIdx.x * (blockDim.x >> 6) + wave;
    if (ant >= a.m) return;
    const int32_t nchunk = (a.n + 63) / 64;
    uint32_t* vis = lds + (size_t)wave * full_lds_words(a.n, a.words);
    float* rowbuf = reinterpret_cast<float*>(vis + full_row_offset(a.words));
    float* wbuf = rowbuf + (size_t)nchunk * 64;
    float* pbuf = wbuf + (size_t)nchunk * 64;
    int32_t* jbuf = reinterpret_cast<int32_t*>(pbuf + (size_t)nchunk * 64);
    float* tp = reinterpret_cast<float*>(jbuf + (size_t)nchunk * 64);
    const uint32_t iter = a.state->iter;

    if (a.aging && !a.valid[ant]) return;
    AntCtx ac{false, 0.0f};
    ac.young = ant_is_young(a, ant, a.state->converged);
    ac.ratio = ac.young ? (float)a.age[ant] / (float)a.min_age : 0.0f;
    for (int32_t w = lane; w < a.words; w += 64) vis[w] = 0u;
    AntRng rng;
    rng_refill(rng, 0, (uint32_t)ant, iter, a, lane);
    const int32_t start = word_to_city(rng_word(rng, 0), a.n);
    mark_visited(vis, start, lane);

    int32_t* tour = a.tours + (size_t)ant * a.n;
    int32_t tbuf = start;
    int32_t c = start;
    unsigned long long fb = 0;
    for (int32_t s = 1; s < a.n; ++s) {
        rng_step(rng, s, (uint32_t)ant, iter, a, lane);
        const float u = word_to_uniform(rng_word(rng, s));
        const int32_t* nrow = a.nn + (size_t)c * a.n;
        // FULL_GROUP chunks of 64 candidates at a time.  The list entries of
        // the first group are requested together with the weight row, which
        // is read in city order (16-byte loads), zeroed where visited and
        // kept in LDS; the list order then gathers from LDS, so a step costs
        // one memory round trip for n <= 64 FULL_GROUP.  The chunk scans run
        // in chunk order; T accumulates the chunk totals sequentially.
        int32_t jj[FULL_GROUP];
        auto load_entries = [&](int32_t g0) {
#pragma unroll
            for (int i = 0; i < FULL_GROUP; ++i) {
                const int32_t r = (g0 + i) * 64 + lane;
                jj[i] = r < a.n ? nrow[r] : -1;
            }
        };
        load_entries(0);
        stage_weight_row(a, c, ac, vis, rowbuf, lane);
        float T = 0.0f;
        for (int32_t g0 = 0; g0 < nchunk; g0 += FULL_GROUP) {
            if (g0 > 0) load_entries(g0);
            float w[FULL_GROUP];
#pragma unroll
            for (int i = 0; i < FULL_GROUP; ++i) w[i] = jj[i] >= 0 ? rowbuf[jj[i]] : 0.0f;
#pragma unroll
            for (int i = 0; i < FULL_GROUP; ++i) {
                if (g0 + i < nchunk) {
                    const int32_t r = (g0 + i) * 64 + lane;
                    const float p = wave_inclusive_scan(w[i]);
                    wbuf[r] = w[i];
                    pbuf[r] = p;
                    jbuf[r] = jj[i];
                    if (lane == 0) tp[g0 + i] = T;
                    T = T + readlane_f(p, 63);
                }
            }
        }
        if (lane == 0) tp[nchunk] = T;
        // The pick is the first (chunk, lane) with tp[q] + prefix > r and a
        // positive weight.  tp[q + 1] is the same sum for the chunk's last
        // lane and the sums grow along a chunk, so the first chunk whose
        // tp[q + 1] exceeds r holds it (its first lane above r has a
        // positive weight: the sum rose there); one ballot over the chunk
        // ends, one over that chunk.
        int32_t next = -1;
        if (T > 0.0f) {
            const float rr = u * T;
            for (int32_t q0 = 0; q0 < nchunk; q0 += 64) {
                const int32_t q = q0 + lane;
                const unsigned long long qa = __ballot(q < nchunk && tp[q < nchunk ? q + 1 : 0] > rr);
                if (qa) {
                    const int32_t qs = q0 + __ffsll((long long)qa) - 1;
                    const int32_t r = qs * 64 + lane;
                    const float P = tp[qs] + pbuf[r];
                    const unsigned long long hit = __ballot(P > rr && wbuf[r] > 0.0f);
                    if (hit) next = jbuf[qs * 64 + __ffsll((long long)hit) - 1];
                    break;
                }
            }
        }
        if (next < 0) {
            ++fb;
            for (int32_t q = nchunk - 1; q >= 0; --q) {
                const int32_t r = q * 64 + lane;
                const unsigned long long m = __ballot(r < a.n && !is_visited(vis, jbuf[r < a.n ? r : 0]));
                if (m) {
                    next = jbuf[q * 64 + 63 - __clzll((long long)m)];
                    break;
                }
            }
        }
        next = uniform_i(next);
        mark_visited(vis, next, lane);
        stage_city(tour, tbuf, s, next, a.n, lane);
        c = next;
    }
    if (a.n == 1 && lane == 0) tour[0] = start;
    if (lane == 0 && fb) atomicAdd(&a.state->fb_steps, fb);
    store_tour_length(a, tour, ant, lane);
}

// ---------------------------------------------------------------------------
// Iteration bookkeeping in one 1024-thread block (computeAntDistances
// src/Colony.cu:869-1047, CPU semantics): aging performance counters,
// iteration best (first strict minimum), best-so-far update and tour copy,
// deposit list (AS: every valid ant, dense by ant index; RBAS: the w-1 best
// ranked ants plus the best-so-far tour), then aging retire/recruit
// (src/Colony.cu:366-423).
struct LenKey {
    int64_t len;
    int32_t ant;
};

__device__ __forceinline__ bool key_less(const LenKey& x, const LenKey& y) {
    return x.len < y.len || (x.len == y.len && x.ant < y.ant);
}

// (len, ant) minimum over the wave; every lane gets it.
__device__ __forceinline__ LenKey wave_min_key(LenKey v) {
    for (int off = 32; off >= 1; off >>= 1) {
        LenKey o;
        const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)v.len, off);
        const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)v.len >> 32), off);
        o.len = (int64_t)(((uint64_t)hi << 32) | lo);
        o.ant = __shfl_xor(v.ant, off);
        if (key_less(o, v)) v = o;
    }
    return v;
}

__device__ __forceinline__ LenKey block_min_key(LenKey v, LenKey* red) {
    v = wave_min_key(v);
    const int nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    LenKey b = red[0];
    for (int w = 1; w < nw; ++w)
        if (key_less(red[w], b)) b = red[w];
    return b;
}

__device__ __forceinline__ int32_t block_min_i32(int32_t v, int32_t* red) {
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int32_t b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = min(b, red[w]);
    return b;
}

// Largest w-1 of the two-round rank selection in stats_block.
constexpr int32_t RANK_MAX = 16;

// `taken` holds (m+31)/32 words (RBAS ranking); red/redi one slot per wave;
// red_rank (16 x RANK_MAX) the per-wave rank candidates.
__device__ __forceinline__ void stats_block_general(const StatsArgs& a, LenKey* red, int32_t* redi, uint32_t* taken,
                                    int32_t& s_improved, int32_t& s_recruit, LenKey* red_rank) {
    DevState* st = a.state;
    const LenKey none{(int64_t)ACO_MAX_DISTANCE - 1, 0x7fffffff};

    LenKey mine = none;
    for (int32_t k = threadIdx.x; k < a.m; k += blockDim.x) {
        if (a.aging && !a.valid[k]) continue;
        const int64_t L = a.lengths[k];
        if (a.aging) {
            a.perf[k] = (L == a.prev_len[k]) ? a.perf[k] + 1 : 0;
            a.prev_len[k] = L;
        }
        const LenKey x{L, k};
        if (key_less(x, mine)) mine = x;
    }
    const LenKey ib = block_min_key(mine, red);
    if (threadIdx.x == 0) {
        // iterBestDist starts at MAX_DISTANCE - 1 and only a strictly smaller
        // length is taken (src/Colony.cu:872,1015).
        const bool found = ib.len < (int64_t)ACO_MAX_DISTANCE - 1;
        st->iter_best = found ? ib.len : (int64_t)ACO_MAX_DISTANCE - 1;
        st->iter_best_ant = found ? ib.ant : -1;
        const bool improved = found && ib.len < st->glob_best;
        if (improved) {
            st->glob_best = ib.len;
            st->gperf = 0;
        } else {
            st->gperf += 1;
        }
        s_improved = improved ? ib.ant : -1;
    }
    __syncthreads();
    const int32_t imp = s_improved;
    if (imp >= 0)
        for (int32_t i = threadIdx.x; i < a.n; i += blockDim.x) a.best_tour[i] = a.tours[(size_t)imp * a.n + i];

    if (a.variant == 0) {  // AS: dense, one entry per ant
        for (int32_t k = threadIdx.x; k < a.m; k += blockDim.x) {
            const bool v = !a.aging || a.valid[k];
            a.dep_val[k] = v ? (float)(1.0 / (double)a.lengths[k]) : 0.0f;
        }
        if (threadIdx.x == 0) st->ndep = a.m;
    } else {  // RBAS, general w and m: w-1 block-wide passes
        const int32_t words = (a.m + 31) / 32;
        for (int32_t w = threadIdx.x; w < words; w += blockDim.x) taken[w] = 0u;
        __syncthreads();
        int32_t nd = 0;
        for (int32_t r = 1; r <= a.w - 1; ++r) {
            LenKey mk = none;
            for (int32_t k = threadIdx.x; k < a.m; k += blockDim.x) {
                if (a.aging && !a.valid[k]) continue;
                if ((taken[k >> 5] >> (k & 31)) & 1u) continue;
                const LenKey x{a.lengths[k], k};
                if (key_less(x, mk)) mk = x;
            }
            // `none` has len MAX-1: a valid ant of that length cannot exist.
            const LenKey b = block_min_key(mk, red);
            if (b.ant == 0x7fffffff) break;
            if (threadIdx.x == 0) {
                taken[b.ant >> 5] |= 1u << (b.ant & 31);
                a.dep_ant[nd] = b.ant;
                a.dep_val[nd] = (float)((double)(a.w - r) / (double)b.len);
            }
            ++nd;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (st->glob_best < (int64_t)ACO_MAX_DISTANCE) {
                a.dep_ant[nd] = -1;
                a.dep_val[nd] = (float)((double)a.w / (double)st->glob_best);
                ++nd;
            }
            st->ndep = nd;
        }
    }

    if (a.aging) {
        // retireAnts: first valid ant among the first m/2 at max_age.
        int32_t cand = 0x7fffffff;
        for (int32_t k = threadIdx.x; k < a.m / 2; k += blockDim.x)
            if (a.valid[k] && a.age[k] >= a.max_age) cand = min(cand, k);
        const int32_t ret = block_min_i32(cand, redi);
        if (threadIdx.x == 0 && ret != 0x7fffffff) {
            a.valid[ret] = 0;
            st->num_valid -= 1;
            st->converged = 1;
        }
        // recruiteAnts: first invalid ant, when the colony stagnates or is small.
        if (threadIdx.x == 0) s_recruit = (st->gperf >= a.global_perf || st->num_valid < a.min_pop);
        __syncthreads();
        if (s_recruit) {
            int32_t inv = 0x7fffffff;
            for (int32_t k = threadIdx.x; k < a.m; k += blockDim.x)
                if (!a.valid[k]) inv = min(inv, k);
            const int32_t rc = block_min_i32(inv, redi);
            if (threadIdx.x == 0 && rc != 0x7fffffff) {
                a.valid[rc] = 1;
                a.age[rc] = 0;
                st->num_valid += 1;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        st->iter += 1;
        // TSP::terminationCondition's optimum test (src/TSP.cpp:168-171) on the
        // iteration best, evaluated on the device
        if (st->stop_at >= 0 && st->iter_best <= st->stop_at) st->halted = 1;
    }
}

// Block minimum of 64-bit keys: wave minima through LDS (one slot per wave).
__device__ __forceinline__ uint64_t block_min_u64(uint64_t v, uint64_t* red) {
    v = wave_min_u64(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = min(b, red[w]);
    return b;
}

// Wave minimum of 32-bit keys: one v_min_u32 per step of the scan pattern
// (sources outside the pattern read ~0), lane 63 read by every lane.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_min_src_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)~0u, (int)v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, dpp_min_src_u32<0x111, 0xf>(v));
    v = min(v, dpp_min_src_u32<0x112, 0xf>(v));
    v = min(v, dpp_min_src_u32<0x114, 0xf>(v));
    v = min(v, dpp_min_src_u32<0x118, 0xf>(v));
    v = min(v, dpp_min_src_u32<0x142, 0xa>(v));
    v = min(v, dpp_min_src_u32<0x143, 0xc>(v));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// (len, ant) keys packed as len << 20 | ant: the packed order is the
// (len, ant) order for ant < 2^20 and len < 2^44 (n <= 2^18 cities of
// distance <= 2^26).
__device__ __forceinline__ uint64_t pack_key(int64_t len, int32_t ant) { return ((uint64_t)len << 20) | (uint32_t)ant; }
__device__ __forceinline__ int64_t key_len(uint64_t k) { return (int64_t)(k >> 20); }
__device__ __forceinline__ int32_t key_ant(uint64_t k) { return (int32_t)(k & 0xfffff); }

// The same (len, ant) order in the key width the statistics block works in:
// 64-bit keys in general, 32-bit keys (len << 10 | ant) when m <= 1024 and
// every valid ant's length is below 2^22, where each reduction step is one
// v_min_u32 instead of a 64-bit compare and two selects.
template <typename K>
struct StatKey;
template <>
struct StatKey<uint64_t> {
    static constexpr uint64_t NONE = ~0ull;
    static __device__ __forceinline__ uint64_t pack(int64_t len, int32_t ant) { return pack_key(len, ant); }
    static __device__ __forceinline__ int64_t len(uint64_t k) { return key_len(k); }
    static __device__ __forceinline__ int32_t ant(uint64_t k) { return key_ant(k); }
    static __device__ __forceinline__ uint64_t wave_min(uint64_t v) { return wave_min_u64(v); }
};
template <>
struct StatKey<uint32_t> {
    static constexpr uint32_t NONE = ~0u;
    static constexpr int64_t LEN_LIMIT = 1 << 22;
    static __device__ __forceinline__ uint32_t pack(int64_t len, int32_t ant) {
        return ((uint32_t)len << 10) | (uint32_t)ant;
    }
    static __device__ __forceinline__ int64_t len(uint32_t k) { return (int64_t)(k >> 10); }
    static __device__ __forceinline__ int32_t ant(uint32_t k) { return (int32_t)(k & 1023u); }
    static __device__ __forceinline__ uint32_t wave_min(uint32_t v) { return wave_min_u32(v); }
};

template <typename K>
__device__ __forceinline__ K block_min_key(K v, K* red) {
    v = StatKey<K>::wave_min(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    K b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = min(b, red[w]);
    return b;
}

// The same statistics with every ant's length, validity and age loaded into
// registers once (at most four ants per thread) and the iteration state kept
// in LDS: the dependent chain is one round of global loads instead of one
// per phase.  Used when m <= 4 x block size and (rank-based) w - 1 <= RANK_MAX.
struct StatsShared {
    int64_t glob_best;
    int32_t gperf, num_valid, improved, recruit;
};

// Per-thread view of its (at most four) ants, loaded once.
struct StatsAnts {
    int64_t L[4];
    int32_t AG[4];
    bool IN[4], V[4];
};

// Rank-based statistics with one reduction pass: every wave leaves its
// smallest max(w - 1, 1) keys and its aging candidates in LDS, then wave 0
// takes the colony's smallest keys in order (the first is the iteration
// best), updates the best-so-far, writes the deposit list (one division per
// lane) and makes the aging decision; after one more barrier the block
// copies an improved tour.  The same results as the per-phase reductions of
// the Ant System path (src/Colony.cu:996-1043 order).
template <typename K>
__device__ __forceinline__ void stats_block_ranked(const StatsArgs& a, const StatsAnts& s, K* red, K* red_rank,
                                                   StatsShared& sh) {
    using SK = StatKey<K>;
    DevState* st = a.state;
    const int32_t B = blockDim.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = (int)(B >> 6);
    const int32_t t = a.w - 1;     // ranked ants in the deposit list (<= RANK_MAX)
    const int32_t tt = max(t, 1);  // keys taken per wave; the first is also the iteration best
    K ks[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ks[q] = s.V[q] ? SK::pack(s.L[q], (int32_t)(threadIdx.x + q * B)) : SK::NONE;
    for (int32_t r = 0; r < tt; ++r) {
        const K wmin = SK::wave_min(min(min(ks[0], ks[1]), min(ks[2], ks[3])));
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ks[q] == wmin) ks[q] = SK::NONE;
        if (lane == 0) red_rank[wave * RANK_MAX + r] = wmin;
    }
    // retireAnts / recruiteAnts candidates (see the Ant System path below)
    uint32_t* red32 = reinterpret_cast<uint32_t*>(red);
    if (a.aging) {
        uint32_t cand = ~0u, inv = ~0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t k = threadIdx.x + q * B;
            if ((int32_t)k < a.m / 2 && s.V[q] && s.AG[q] >= a.max_age) cand = min(cand, k);
            if (s.IN[q] && !s.V[q]) inv = min(inv, k);
        }
        cand = wave_min_u32(cand);
        inv = wave_min_u32(inv);
        if (lane == 0) {
            red32[wave] = cand;
            red32[nw + wave] = inv;
        }
    }
    __syncthreads();
    if (wave == 0) {
        const int32_t nc = nw * RANK_MAX;  // candidate slots, up to four per lane
        K cs[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t x = lane + 64 * q;
            cs[q] = (x < nc && (x % RANK_MAX) < tt) ? red_rank[x] : SK::NONE;
        }
        K mykey = SK::NONE, ib = SK::NONE;  // lane r: the r-th smallest key
        int32_t found = 0;
        for (int32_t r = 0; r < tt; ++r) {
            const K wmin = SK::wave_min(min(min(cs[0], cs[1]), min(cs[2], cs[3])));
            if (wmin == SK::NONE) break;  // fewer valid ants
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (cs[q] == wmin) cs[q] = SK::NONE;
            mykey = lane == r ? wmin : mykey;
            ib = r == 0 ? wmin : ib;
            ++found;
        }
        // iterBestDist starts at MAX_DISTANCE - 1 and only a strictly smaller
        // length is taken (src/Colony.cu:872,1015)
        const bool fnd = ib != SK::NONE && SK::len(ib) < (int64_t)ACO_MAX_DISTANCE - 1;
        const int64_t ibest = fnd ? SK::len(ib) : (int64_t)ACO_MAX_DISTANCE - 1;
        const bool improved = fnd && ibest < sh.glob_best;
        const int64_t gb = improved ? ibest : sh.glob_best;
        const int32_t gperf = improved ? 0 : sh.gperf + 1;
        // deposit list: the t best ranked ants with weights w - r, then the
        // best-so-far tour with weight w
        const int32_t nr = min(t, found);
        if (lane < nr) {
            a.dep_ant[lane] = SK::ant(mykey);
            a.dep_val[lane] = (float)((double)(a.w - (lane + 1)) / (double)SK::len(mykey));
        }
        if (lane == 0) {
            st->iter_best = ibest;
            st->iter_best_ant = fnd ? SK::ant(ib) : -1;
            st->glob_best = gb;
            st->gperf = gperf;
            sh.glob_best = gb;
            sh.gperf = gperf;
            sh.improved = improved ? SK::ant(ib) : -1;
            int32_t nd = nr;
            if (gb < (int64_t)ACO_MAX_DISTANCE) {
                a.dep_ant[nd] = -1;
                a.dep_val[nd] = (float)((double)a.w / (double)gb);
                ++nd;
            }
            st->ndep = nd;
            if (a.aging) {
                uint32_t ret = red32[0], first_inv = red32[nw];
                for (int w = 1; w < nw; ++w) {
                    ret = min(ret, red32[w]);
                    first_inv = min(first_inv, red32[nw + w]);
                }
                int32_t num_valid = sh.num_valid;
                if (ret != ~0u) {
                    a.valid[ret] = 0;
                    num_valid -= 1;
                    st->converged = 1;
                    first_inv = min(first_inv, ret);
                }
                if ((gperf >= a.global_perf || num_valid < a.min_pop) && first_inv != ~0u) {
                    a.valid[first_inv] = 1;
                    a.age[first_inv] = 0;
                    num_valid += 1;
                }
                sh.num_valid = num_valid;
                st->num_valid = num_valid;
            }
            st->iter += 1;
            // TSP::terminationCondition's optimum test (src/TSP.cpp:168-171)
            if (st->stop_at >= 0 && ibest <= st->stop_at) st->halted = 1;
        }
    }
    __syncthreads();
    const int32_t imp = sh.improved;
    if (imp >= 0)
        for (int32_t i = threadIdx.x; i < a.n; i += B) a.best_tour[i] = a.tours[(size_t)imp * a.n + i];
}

template <typename K>
__device__ __forceinline__ void stats_block_fast(const StatsArgs& a, const StatsAnts& s, K* red, K* red_rank,
                                                 StatsShared& sh) {
    if (a.variant != 0) {
        stats_block_ranked<K>(a, s, red, red_rank, sh);
        return;
    }
    using SK = StatKey<K>;
    DevState* st = a.state;
    const int32_t B = blockDim.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    bool V[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) V[q] = s.V[q];
    K mine = SK::NONE;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (V[q]) mine = min(mine, SK::pack(s.L[q], (int32_t)(threadIdx.x + q * B)));
    const K ib = block_min_key(mine, red);
    // iterBestDist starts at MAX_DISTANCE - 1 and only a strictly smaller
    // length is taken (src/Colony.cu:872,1015).
    const bool found = ib != SK::NONE && SK::len(ib) < (int64_t)ACO_MAX_DISTANCE - 1;
    const int64_t ibest = found ? SK::len(ib) : (int64_t)ACO_MAX_DISTANCE - 1;
    if (threadIdx.x == 0) {
        st->iter_best = ibest;
        st->iter_best_ant = found ? SK::ant(ib) : -1;
        const bool improved = found && ibest < sh.glob_best;
        if (improved) {
            sh.glob_best = ibest;
            sh.gperf = 0;
        } else {
            sh.gperf += 1;
        }
        st->glob_best = sh.glob_best;
        st->gperf = sh.gperf;
        sh.improved = improved ? SK::ant(ib) : -1;
    }
    __syncthreads();
    const int32_t imp = sh.improved;
    if (imp >= 0)
        for (int32_t i = threadIdx.x; i < a.n; i += B) a.best_tour[i] = a.tours[(size_t)imp * a.n + i];

    // AS: dense, one entry per ant
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (s.IN[q]) a.dep_val[threadIdx.x + q * B] = V[q] ? (float)(1.0 / (double)s.L[q]) : 0.0f;
    if (threadIdx.x == 0) st->ndep = a.m;

    if (a.aging) {
        // retireAnts: the first valid ant among the first m/2 at max_age;
        // recruiteAnts: the first invalid ant (the retired one included),
        // when the colony stagnates or is small.  Both minima in one block
        // reduction: the first invalid ant after the retirement is the
        // smaller of the first one before it and the retired ant.
        uint32_t cand = ~0u, inv = ~0u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t k = threadIdx.x + q * B;
            if ((int32_t)k < a.m / 2 && V[q] && s.AG[q] >= a.max_age) cand = min(cand, k);
            if (s.IN[q] && !V[q]) inv = min(inv, k);
        }
        cand = wave_min_u32(cand);
        inv = wave_min_u32(inv);
        uint32_t* red32 = reinterpret_cast<uint32_t*>(red);
        const int nw = (int)(blockDim.x >> 6);
        __syncthreads();
        if (lane == 0) {
            red32[wave] = cand;
            red32[nw + wave] = inv;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t ret = red32[0], first_inv = red32[nw];
            for (int w = 1; w < nw; ++w) {
                ret = min(ret, red32[w]);
                first_inv = min(first_inv, red32[nw + w]);
            }
            if (ret != ~0u) {
                a.valid[ret] = 0;
                sh.num_valid -= 1;
                st->converged = 1;
                first_inv = min(first_inv, ret);
            }
            if ((sh.gperf >= a.global_perf || sh.num_valid < a.min_pop) && first_inv != ~0u) {
                a.valid[first_inv] = 1;
                a.age[first_inv] = 0;
                sh.num_valid += 1;
            }
            st->num_valid = sh.num_valid;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        st->iter += 1;
        // TSP::terminationCondition's optimum test (src/TSP.cpp:168-171) on the
        // iteration best, evaluated on the device
        if (st->stop_at >= 0 && ibest <= st->stop_at) st->halted = 1;
    }
}

// Loads the ants once, updates the stagnation counters and picks the key
// width (a block-wide vote on the 32-bit limits).
__device__ __forceinline__ void stats_block_fast_entry(const StatsArgs& a, uint64_t* red, uint64_t* red_rank,
                                                       StatsShared& sh) {
    DevState* st = a.state;
    const int32_t B = blockDim.x;
    if (threadIdx.x == 0) {
        sh.glob_best = st->glob_best;
        sh.gperf = st->gperf;
        sh.num_valid = st->num_valid;
    }
    StatsAnts s;
    int64_t PL[4];
    int32_t PF[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int32_t k = threadIdx.x + q * B;
        s.IN[q] = k < a.m;
        s.L[q] = s.IN[q] ? a.lengths[k] : 0;
        s.V[q] = s.IN[q] && (!a.aging || a.valid[k]);
        s.AG[q] = (s.IN[q] && a.aging) ? a.age[k] : 0;
        PL[q] = (s.IN[q] && a.aging) ? a.prev_len[k] : 0;
        PF[q] = (s.IN[q] && a.aging) ? a.perf[k] : 0;
    }
    bool wide = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!s.V[q]) continue;
        const int32_t k = threadIdx.x + q * B;
        if (a.aging) {
            a.perf[k] = (s.L[q] == PL[q]) ? PF[q] + 1 : 0;
            a.prev_len[k] = s.L[q];
        }
        wide = wide || s.L[q] < 0 || s.L[q] >= StatKey<uint32_t>::LEN_LIMIT;
    }
    wide = __syncthreads_or(wide ? 1 : 0) != 0 || a.m > 1024;
    if (wide)
        stats_block_fast<uint64_t>(a, s, red, red_rank, sh);
    else
        stats_block_fast<uint32_t>(a, s, reinterpret_cast<uint32_t*>(red), reinterpret_cast<uint32_t*>(red_rank), sh);
}

__device__ __forceinline__ void stats_block(const StatsArgs& a, LenKey* red, int32_t* redi, uint32_t* taken, int32_t& s_improved,
                            int32_t& s_recruit, LenKey* red_rank, StatsShared& sh) {
    if (a.m <= 4 * (int32_t)blockDim.x && a.m < (1 << 20) && a.n <= (1 << 18) &&
        (a.variant == 0 || a.w - 1 <= RANK_MAX))
        stats_block_fast_entry(a, reinterpret_cast<uint64_t*>(red), reinterpret_cast<uint64_t*>(red_rank), sh);
    else
        stats_block_general(a, red, redi, taken, s_improved, s_recruit, red_rank);
}

__global__ __launch_bounds__(1024) void k_stats(const StatsArgs a) {
    __shared__ LenKey red[16];
    __shared__ int32_t redi[16];
    __shared__ int32_t s_improved, s_recruit;
    extern __shared__ uint32_t taken[];  // RBAS: ants already ranked
    __shared__ LenKey red_rank[16 * RANK_MAX];
    __shared__ StatsShared sh;
    stamp_start(a.stamp);
    if (a.state->skip) return;
    // the ants that built a tour this iteration: the valid ones (validity
    // changes only below, after the deposit list is made)
    if (threadIdx.x == 0) a.state->tours += a.aging ? (unsigned long long)a.state->num_valid : (unsigned long long)a.m;
    stats_block(a, red, redi, taken, s_improved, s_recruit, red_rank, sh);
    if (a.host_stats != nullptr) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const DevState* st = a.state;
            a.host_stats[0] = st->iter_best;
            a.host_stats[1] = st->glob_best;
            a.host_stats[2] = a.aging ? (int64_t)st->num_valid : (int64_t)a.m;
            __threadfence_system();
        }
    }
}

// ---------------------------------------------------------------------------
// Ant System deposit with fp32 atomics (the path the n = 20000 workload
// names, and AS beyond the adjacency deposit's size): one block per list
// entry, atomics on both (i,j) and (j,i) of every tour edge
// (dev_update_pheromones src/AntSystem.cu:105-128, which only updated one
// direction; the CPU branch src/AntSystem.cu:296-316 is symmetric); the
// rank-based deposit adds its tours in list order instead (k_deposit_rows).
// One depositing tour per block.  The iteration state and the block's
// deposit entry are read together (the entry arrays hold m + w + 2 slots per
// colony, so every block's slot exists), and each thread loads the ends of
// up to U edges before its atomics: two memory round trips before the first
// atomic instead of one per phase and per edge batch.
__device__ __forceinline__ void deposit_tour(const DepositArgs& a, int32_t e, bool check_skip) {
    constexpr int U = 8;
    const int32_t skip = check_skip ? a.state->skip : 0;
    const int32_t ndep = a.state->ndep;
    const int32_t ant = a.dense ? e : a.dep_ant[e];
    const float d = a.dep_val[e];
    if (skip || e >= ndep || d == 0.0f) return;
    const int32_t* t = ant >= 0 ? a.tours + (size_t)ant * a.n : a.best_tour;
    const int32_t stride = blockDim.x;
    for (int32_t i0 = threadIdx.x; i0 < a.n; i0 += U * stride) {
        int32_t x[U], y[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {  // indices past the end reload the last edge
            const int32_t i = min(i0 + q * stride, a.n - 1);
            x[q] = t[i];
            y[q] = t[(i + 1 == a.n) ? 0 : i + 1];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            if (i0 + q * stride < a.n) {
                unsafeAtomicAdd(a.tau_next + (size_t)x[q] * a.P + y[q], d);
                unsafeAtomicAdd(a.tau_next + (size_t)y[q] * a.P + x[q], d);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_deposit(const DepositArgs a) {
    stamp_start(a.stamp);
    deposit_tour(a, blockIdx.x, true);
}

// Rank-based deposit with one writer per pheromone cell.  A block owns a
// band of rows x0 .. x0 + rows - 1 (rows <= ROWS) of tau_next.  For each
// pass of up to DEP_PASS list entries, the waves read the entries' tours 64
// positions at a time and note, for every city of the band, its predecessor
// and successor in each tour (neighbours by shuffle, the wave's two end
// lanes load theirs).  Then four lanes per row add the pass's deposits to
// the row's cells in list order, each cell loaded once and stored once by
// one lane: ((tau + d_0) + d_1) + ..., the order of the CPU loop that adds
// the ranked tours one after another (src/AntSystem.cu:296-316 per tour;
// the list is built by the statistics kernel).  No atomics, so the sums, and
// with them the colony's run, are the same from launch to launch.  Small
// bands spread the cell loads, which miss in L2, over many CUs; every block
// reads the whole tours.  Every thread of a 1024-thread block calls it (it
// synchronises the block).
constexpr int32_t DEP_PASS = 8;
template <int ROWS>
struct DepositRowsShared {
    int32_t nb[DEP_PASS][ROWS][2];
    float dv[DEP_PASS];
    const int32_t* tp[DEP_PASS];
};
// Phase 1 of a pass (entries 0 .. ne - 1 with weights dv and tours tp in
// LDS): wave w takes positions 64 s + lane, s = w, w + nw, ..., of every
// entry at once (DEP_PASS loads in flight per lane) and notes the band's
// cities' neighbours in nb.  Every thread calls it.
template <int ROWS>
__device__ __forceinline__ void deposit_rows_scan(int32_t n, int32_t x0, int32_t rows, int32_t ne,
                                                  DepositRowsShared<ROWS>& sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int32_t nseg = (n + 63) >> 6;
    for (int32_t s0 = wave; s0 < nseg; s0 += nw) {
        const int32_t i = (s0 << 6) + lane;
        const bool last = lane == 63 || i == n - 1;  // successor outside the wave's 64
        int32_t c[DEP_PASS], pe[DEP_PASS], se[DEP_PASS];
#pragma unroll
        for (int e = 0; e < DEP_PASS; ++e) {
            const bool in = e < ne && i < n && sh.dv[e] != 0.0f;
            const int32_t* t = sh.tp[e];
            c[e] = in ? t[i] : -1;
            pe[e] = (in && lane == 0) ? t[i == 0 ? n - 1 : i - 1] : -1;
            se[e] = (in && last) ? t[i + 1 == n ? 0 : i + 1] : -1;
        }
#pragma unroll
        for (int e = 0; e < DEP_PASS; ++e) {
            const int32_t up = __shfl_up(c[e], 1), down = __shfl_down(c[e], 1);
            const int32_t r = c[e] - x0;
            if (c[e] >= 0 && (uint32_t)r < (uint32_t)rows) {
                sh.nb[e][r][0] = lane == 0 ? pe[e] : up;
                sh.nb[e][r][1] = last ? se[e] : down;
            }
        }
    }
}

// Phase 2 of a pass: four lanes per row; the row's 2 ne (target, deposit)
// slots; a target's first slot loads the cell, sums every slot naming it in
// list order and stores it; the row's k-th first slot is lane k mod 4's.
template <int ROWS>
__device__ __forceinline__ void deposit_rows_cells(float* tau_next, int32_t P, int32_t x0, int32_t rows, int32_t ne,
                                                   const DepositRowsShared<ROWS>& sh) {
    if ((int32_t)threadIdx.x >= 4 * rows) return;
    const int32_t r = threadIdx.x >> 2, sub = threadIdx.x & 3;
    float* row = tau_next + (size_t)(x0 + r) * P;
    int32_t j[2 * DEP_PASS];
    float d[2 * DEP_PASS], v[2 * DEP_PASS];
#pragma unroll
    for (int q = 0; q < 2 * DEP_PASS; ++q) {
        const int e = q >> 1;
        const bool on = e < ne && sh.dv[e] != 0.0f;
        j[q] = on ? sh.nb[e][r][q & 1] : -1;
        d[q] = on ? sh.dv[e] : 0.0f;
    }
    uint32_t mine = 0;
    int32_t cnt = 0;
#pragma unroll
    for (int q = 0; q < 2 * DEP_PASS; ++q) {
        bool f = j[q] >= 0;
#pragma unroll
        for (int q2 = 0; q2 < q; ++q2) f = f && j[q2] != j[q];
        mine |= (uint32_t)(f && (cnt & 3) == sub) << q;
        cnt += f;
    }
#pragma unroll
    for (int q = 0; q < 2 * DEP_PASS; ++q) v[q] = ((mine >> q) & 1) ? row[j[q]] : 0.0f;
#pragma unroll
    for (int q = 0; q < 2 * DEP_PASS; ++q) {
        if (!((mine >> q) & 1)) continue;
        float s = v[q];
#pragma unroll
        for (int q2 = q; q2 < 2 * DEP_PASS; ++q2)
            if (j[q2] == j[q]) s += d[q2];
        row[j[q]] = s;
    }
}

template <int ROWS>
__device__ __forceinline__ void deposit_rows_band(const DepositArgs& a, int32_t x0, int32_t rows, bool check_skip,
                                                  DepositRowsShared<ROWS>& sh) {
    static_assert(4 * ROWS <= 1024, "four lanes per row");
    // the first pass's list slots are read with the list length and the skip
    // flag, in one round trip (slots past ndep are read but never used)
    int32_t ant = -1;
    float dval = 0.0f;
    if ((int32_t)threadIdx.x < min(DEP_PASS, a.list_cap)) {
        ant = a.dep_ant[threadIdx.x];
        dval = a.dep_val[threadIdx.x];
    }
    const int32_t ndep = a.state->ndep;
    if (check_skip && a.state->skip) return;
    for (int32_t e0 = 0; e0 < ndep; e0 += DEP_PASS) {
        const int32_t ne = min(DEP_PASS, ndep - e0);
        if (e0 > 0 && (int32_t)threadIdx.x < ne) {
            ant = a.dep_ant[e0 + threadIdx.x];
            dval = a.dep_val[e0 + threadIdx.x];
        }
        if ((int32_t)threadIdx.x < ne) {
            sh.dv[threadIdx.x] = dval;
            sh.tp[threadIdx.x] = ant >= 0 ? a.tours + (size_t)ant * a.n : a.best_tour;
        }
        __syncthreads();
        deposit_rows_scan<ROWS>(a.n, x0, rows, ne, sh);
        __syncthreads();
        deposit_rows_cells<ROWS>(a.tau_next, a.P, x0, rows, ne, sh);
        __syncthreads();
    }
}

// Bands of 16 rows up to 4096 cities (every block rereads the whole tours
// from L2), 256 rows beyond.
template <int ROWS>
__global__ __launch_bounds__(1024) void k_deposit_rows(const DepositArgs a) {
    stamp_start(a.stamp);
    __shared__ DepositRowsShared<ROWS> sh;
    const int32_t x0 = blockIdx.x * ROWS;
    deposit_rows_band<ROWS>(a, x0, min(ROWS, a.n - x0), true, sh);
}

// Rank-based statistics and deposit in one launch (m <= 1024, w <= DEP_PASS,
// n <= 4096; 1024-thread blocks over 16-row bands).  Every block ranks the
// ants itself (one key per thread, the same keys and order as
// stats_block_ranked) and builds the deposit list in LDS; the best-so-far
// entry is the iteration-best ant's tour when this iteration improves the
// best, else the stored best tour.  After scanning its band's tours a block
// counts itself done reading (fuse_count); the block that completes the
// count, when every block has read the iteration's lengths, validity,
// tours, best tour and state, makes the statistics' writes: tours counter,
// per-ant stagnation, iteration best and best-so-far, the best tour's copy,
// the deposit list, aging and the iteration number.  Every block adds its
// band's deposits.  One launch and one list computation per block instead
// of a one-block statistics launch ahead of the deposit's.
struct FusedShared {
    uint64_t red_rank[16 * RANK_MAX];  // per-wave candidates (K-wide)
    uint32_t red32[32];                // aging candidates
    int64_t ibest, gb;
    int32_t ib_ant, improved_ant, gperf, ndep, last;
    int32_t list_ant[DEP_PASS];
};

template <typename K>
__device__ __forceinline__ void fused_rank_list(const StatsArgs& a, int64_t L, bool V, int32_t age, int64_t gb_old,
                                                int32_t gperf_old, FusedShared& fs, DepositRowsShared<16>& sh) {
    using SK = StatKey<K>;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = (int)(blockDim.x >> 6);
    K* red_rank = reinterpret_cast<K*>(fs.red_rank);
    const int32_t t = a.w - 1, tt = max(t, 1);
    K k0 = V ? SK::pack(L, (int32_t)threadIdx.x) : SK::NONE;
    for (int32_t r = 0; r < tt; ++r) {
        const K wmin = SK::wave_min(k0);
        k0 = k0 == wmin ? SK::NONE : k0;
        if (lane == 0) red_rank[wave * RANK_MAX + r] = wmin;
    }
    if (a.aging) {  // retireAnts / recruiteAnts candidates, for the last block
        const int32_t tid = threadIdx.x;
        const bool in = tid < a.m;
        uint32_t cand = (tid < a.m / 2 && V && age >= a.max_age) ? (uint32_t)tid : ~0u;
        uint32_t inv = (in && !V) ? (uint32_t)tid : ~0u;
        cand = wave_min_u32(cand);
        inv = wave_min_u32(inv);
        if (lane == 0) {
            fs.red32[wave] = cand;
            fs.red32[nw + wave] = inv;
        }
    }
    __syncthreads();
    if (wave == 0) {
        const int32_t nc = nw * RANK_MAX;
        K cs[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t x = lane + 64 * q;
            cs[q] = (x < nc && (x % RANK_MAX) < tt) ? red_rank[x] : SK::NONE;
        }
        K mykey = SK::NONE, ib = SK::NONE;
        int32_t found = 0;
        for (int32_t r = 0; r < tt; ++r) {
            const K wmin = SK::wave_min(min(min(cs[0], cs[1]), min(cs[2], cs[3])));
            if (wmin == SK::NONE) break;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (cs[q] == wmin) cs[q] = SK::NONE;
            mykey = lane == r ? wmin : mykey;
            ib = r == 0 ? wmin : ib;
            ++found;
        }
        const bool fnd = ib != SK::NONE && SK::len(ib) < (int64_t)ACO_MAX_DISTANCE - 1;
        const int64_t ibest = fnd ? SK::len(ib) : (int64_t)ACO_MAX_DISTANCE - 1;
        const bool improved = fnd && ibest < gb_old;
        const int64_t gb = improved ? ibest : gb_old;
        const int32_t nr = min(t, found);
        if (lane < nr) {
            const int32_t ant = SK::ant(mykey);
            sh.dv[lane] = (float)((double)(a.w - (lane + 1)) / (double)SK::len(mykey));
            sh.tp[lane] = a.tours + (size_t)ant * a.n;
            fs.list_ant[lane] = ant;
        }
        if (lane == 0) {
            int32_t nd = nr;
            if (gb < (int64_t)ACO_MAX_DISTANCE) {
                sh.dv[nd] = (float)((double)a.w / (double)gb);
                sh.tp[nd] = improved ? a.tours + (size_t)SK::ant(ib) * a.n : a.best_tour;
                fs.list_ant[nd] = -1;
                ++nd;
            }
            fs.ndep = nd;
            fs.ibest = ibest;
            fs.ib_ant = fnd ? SK::ant(ib) : -1;
            fs.improved_ant = improved ? SK::ant(ib) : -1;
            fs.gb = gb;
            fs.gperf = improved ? 0 : gperf_old + 1;
        }
    }
    __syncthreads();
}

// The statistics' writes (src/Colony.cu:996-1043 order), by the last block
// of k_stats_deposit; L, V: this thread's ant as read at the start.
// Per-thread values every block loads with its ant's length (only the last
// block uses them): no memory round trip before the statistics' writes.
struct FusedPre {
    int64_t prev_len, stop_at;
    unsigned long long tours;
    uint32_t iter;
    int32_t perf, age, num_valid, tour_word;  // tour_word: entry tid of the improved ant's tour
};

__device__ __forceinline__ void fused_stats_writes(const StatsArgs& a, int64_t L, bool V, const FusedPre& pre,
                                                   FusedShared& fs, const DepositRowsShared<16>& sh) {
    DevState* st = a.state;
    const int32_t tid = threadIdx.x;
    const int nw = (int)(blockDim.x >> 6);
    // the ants that built a tour this iteration: the valid ones, before aging
    if (tid == 0) st->tours = pre.tours + (a.aging ? (unsigned long long)pre.num_valid : (unsigned long long)a.m);
    if (a.aging && V) {  // Colony::antperformance
        a.perf[tid] = L == pre.prev_len ? pre.perf + 1 : 0;
        a.prev_len[tid] = L;
    }
    if (tid == 0) {
        st->iter_best = fs.ibest;
        st->iter_best_ant = fs.ib_ant;
        st->glob_best = fs.gb;
        st->gperf = fs.gperf;
        st->ndep = fs.ndep;
    }
    if (tid < fs.ndep) {
        a.dep_ant[tid] = fs.list_ant[tid];
        a.dep_val[tid] = sh.dv[tid];
    }
    const int32_t imp = fs.improved_ant;
    if (imp >= 0) {
        if (tid < a.n) a.best_tour[tid] = pre.tour_word;
        for (int32_t i = tid + blockDim.x; i < a.n; i += blockDim.x) a.best_tour[i] = a.tours[(size_t)imp * a.n + i];
    }
    if (a.aging) {  // retireAnts / recruiteAnts (candidates from fused_rank_list), as stats_block_ranked
        if (tid == 0) {
            uint32_t ret = fs.red32[0], first_inv = fs.red32[nw];
            for (int w = 1; w < nw; ++w) {
                ret = min(ret, fs.red32[w]);
                first_inv = min(first_inv, fs.red32[nw + w]);
            }
            int32_t num_valid = pre.num_valid;
            if (ret != ~0u) {
                a.valid[ret] = 0;
                num_valid -= 1;
                st->converged = 1;
                first_inv = min(first_inv, ret);
            }
            if ((fs.gperf >= a.global_perf || num_valid < a.min_pop) && first_inv != ~0u) {
                a.valid[first_inv] = 1;
                a.age[first_inv] = 0;
                num_valid += 1;
            }
            st->num_valid = num_valid;
        }
    }
    if (tid == 0) {
        st->iter = pre.iter + 1;
        // TSP::terminationCondition's optimum test (src/TSP.cpp:168-171)
        if (pre.stop_at >= 0 && fs.ibest <= pre.stop_at) st->halted = 1;
        if (a.host_stats != nullptr) {
            a.host_stats[0] = fs.ibest;
            a.host_stats[1] = fs.gb;
            a.host_stats[2] = a.aging ? (int64_t)st->num_valid : (int64_t)a.m;
            __threadfence_system();
        }
        st->fuse_count = 0;
    }
}

__global__ __launch_bounds__(1024) void k_stats_deposit(const StatsArgs a, const DepositArgs d) {
    if (a.stamp != nullptr && blockIdx.x == 0 && threadIdx.x == 0) a.stamp[0] = wall_clock64();
    __shared__ DepositRowsShared<16> sh;
    __shared__ FusedShared fs;
    DevState* st = a.state;
    const int32_t tid = threadIdx.x;
    const bool in = tid < a.m;
    FusedPre pre{};
    const int64_t L = in ? a.lengths[tid] : 0;
    const bool vl = in && (!a.aging || a.valid[tid]);
    if (a.aging && in) {
        pre.prev_len = a.prev_len[tid];
        pre.perf = a.perf[tid];
        pre.age = a.age[tid];
    }
    const int64_t gb_old = st->glob_best;
    const int32_t gperf_old = st->gperf;
    pre.num_valid = st->num_valid;
    pre.tours = st->tours;
    pre.stop_at = st->stop_at;
    pre.iter = st->iter;
    if (st->skip) return;
    const bool V = vl;
    const bool wide = __syncthreads_or((V && (L < 0 || L >= StatKey<uint32_t>::LEN_LIMIT)) ? 1 : 0) != 0;
    if (wide)
        fused_rank_list<uint64_t>(a, L, V, pre.age, gb_old, gperf_old, fs, sh);
    else
        fused_rank_list<uint32_t>(a, L, V, pre.age, gb_old, gperf_old, fs, sh);
    // the deposit phase's time from block 0 holding its list
    if (a.stamp != nullptr && blockIdx.x == 0 && tid == 0) a.stamp[1] = wall_clock64();
    // the improved ant's tour, requested now for the last block's copy
    if (fs.improved_ant >= 0 && tid < a.n) pre.tour_word = a.tours[(size_t)fs.improved_ant * a.n + tid];
    const int32_t ne = fs.ndep;
    const int32_t x0 = blockIdx.x * 16, rows = min(16, a.n - x0);
    deposit_rows_scan<16>(a.n, x0, rows, ne, sh);
    // This block's reads of the iteration's inputs are over (their values
    // sit in LDS and registers).  The count needs no fence: no store of the
    // last block can reach a read that has returned.  The block's last
    // thread counts (its wave adds no deposits: the 64 lanes of wave 0 do),
    // so the atomic's round trip overlaps the deposits instead of queueing
    // ahead of their loads.
    __syncthreads();
    const bool counter = tid == (int32_t)blockDim.x - 1;
    uint32_t ticket = 0;
    if (counter) ticket = atomicAdd(&st->fuse_count, 1u);
    deposit_rows_cells<16>(d.tau_next, d.P, x0, rows, ne, sh);
    if (counter) fs.last = ticket == gridDim.x - 1;
    __syncthreads();
    if (fs.last) fused_stats_writes(a, L, V, pre, fs, sh);
}

// Ant System deposit in two passes with no global atomics; the same sum as
// deposit_tour (src/AntSystem.cu:295-317): tau_xy += dep_val[a] for every
// ant a whose tour uses edge (x, y), in both directions.
// Pass 1: ant a writes succ_a(x) to slot 2a and pred_a(y) to slot 2a+1 of
// rows x and y of its tour edges (x, y), one wave per ant, 64 edges per
// store.  Plain 16-bit stores: every slot has exactly one writer.  A block
// takes 32 consecutive ants, whose slots share each row's cache lines, so
// the partial-line stores meet in one L2.
// Pass 1, staged (A ants a0 .. a0 + A - 1 of one instance per block, rows
// x0 .. x0 + X - 1 of the adjacency): the ants write their slots of those
// rows into an LDS copy of their 2A columns, which the block then stores row
// by row in whole 16-byte chunks, so each row's slots of these ants reach
// memory in one piece instead of 2A scattered 2-byte stores that each cost a
// line write.  Small instances take every row at once (X = n); large ones
// split the rows into bands, each block reading its ants' tours once per
// band.  Columns of ants past m or without a deposit hold garbage, which the
// gather never reads (dep_val 0 or beyond ndep).
__device__ __forceinline__ void adj_scatter_staged(const DepositArgs& a, int32_t a0, int32_t A, int32_t x0, int32_t X,
                                                   uint16_t* st) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int32_t w2 = 2 * A;
    const int32_t S = adj_stage_stride(A);
    const int32_t rows = min(X, a.n - x0);
    for (int32_t j = wave; j < A; j += nw) {
        const int32_t ant = a0 + j;
        if (ant >= a.m || ant >= a.state->ndep || a.dep_val[ant] == 0.0f) continue;
        const int32_t* t = a.tours + (size_t)ant * a.n;
        // eight edges per lane per pass, their tour loads issued together
        for (int32_t i0 = 0; i0 < a.n; i0 += 512) {
            int32_t xs[8], ys[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int32_t i = i0 + 64 * q + lane;
                xs[q] = i < a.n ? t[i] : -1;
                ys[q] = i < a.n ? t[(i + 1 == a.n) ? 0 : i + 1] : -1;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int32_t x = xs[q], y = ys[q];
                if (x < 0) continue;
                if ((uint32_t)(x - x0) < (uint32_t)rows) st[(x - x0) * S + 2 * j] = (uint16_t)y;
                if ((uint32_t)(y - x0) < (uint32_t)rows) st[(y - x0) * S + 2 * j + 1] = (uint16_t)x;
            }
        }
    }
    __syncthreads();
    const int32_t pitch = adj_pitch(a.m);
    const int32_t chunks = w2 / 8;                // 16-byte chunks of a row's columns
    const int32_t lim = min(w2, pitch - 2 * a0);  // columns inside the row (a multiple of 8)
    for (int32_t idx = threadIdx.x; idx < rows * chunks; idx += blockDim.x) {
        const int32_t r = idx / chunks, c = idx - r * chunks;
        if (8 * c >= lim) continue;
        // the padded LDS row is only 4-byte aligned: four dword reads
        const uint32_t* src = reinterpret_cast<const uint32_t*>(st + r * S + 8 * c);
        *reinterpret_cast<uint4*>(a.adj + (size_t)(x0 + r) * pitch + 2 * a0 + 8 * c) =
            make_uint4(src[0], src[1], src[2], src[3]);
    }
}

// Adds each lane's pending (target t, sum v) to the LDS row: the targets of
// the first two leading lanes are summed over the whole wave first (one DPP
// scan each) and added once, the rest with an LDS atomic per lane.  Once the
// colony agrees on a row's neighbours most lanes hold the same target, whose
// atomics would otherwise serialise on one LDS address.  Every lane calls it.
__device__ __forceinline__ void combine_add(float* delta, int32_t t, float v, int lane) {
#pragma unroll
    for (int round = 0; round < 2; ++round) {
        const unsigned long long act = __ballot(t >= 0);
        if (!act) return;
        const int lead = __ffsll((long long)act) - 1;
        const int32_t tl = readlane_i(t, lead);
        const bool mine = t == tl;
        const float x = readlane_f(wave_inclusive_scan(mine ? v : 0.0f), 63);
        if (lane == lead) atomicAdd(&delta[tl], x);
        if (mine) t = -1;
    }
    if (t >= 0) atomicAdd(&delta[t], v);
}

// Pass 2, one row x per `nthr` threads (thread t of the row): the row's
// slots are summed into the LDS row `delta` (zero on entry, zero again on
// return) by their ants' dep_val (slots of ants with dep_val 0 are stale and
// skipped), then added once to tau_next where nonzero.  Each thread reads
// the slots of 8 consecutive ants as two 16-byte loads and sums their 16
// targets (successors and predecessors alike) into two (target, sum)
// accumulators, a third target going straight to LDS; the accumulators are
// then combined over the wave (combine_add).  Once the colony agrees on a
// row's two neighbours, almost every deposit lands in the accumulators and
// the row takes a few LDS atomics per wave.  (Float sums then associate
// differently from a per-slot order; both orders are within the fp32
// rounding the deposit's tests allow.)  Every lane of a wave calls it.
__device__ __forceinline__ void adj_gather_row(const DepositArgs& a, int32_t x, float* delta, int t, int nthr) {
    const uint16_t* row = a.adj + (size_t)x * adj_pitch(a.m);
    const int32_t nd = min(a.m, a.state->ndep);
    const int lane = t & 63;
    const int32_t wave0 = t - lane;  // nthr is a multiple of 64: whole waves share each pass
    for (int32_t base = 0; base < nd; base += 8 * nthr) {
        if (base + 8 * wave0 >= nd) break;  // no ant of this wave in this pass (wave-uniform)
        const int32_t a0 = base + 8 * t;
        const bool act = a0 < nd;
        uint4 v0 = make_uint4(0, 0, 0, 0), v1 = make_uint4(0, 0, 0, 0);
        if (act) {
            v0 = *reinterpret_cast<const uint4*>(row + 2 * a0);
            v1 = *reinterpret_cast<const uint4*>(row + 2 * a0 + 8);
        }
        float d[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) d[q] = (act && a0 + q < nd) ? a.dep_val[a0 + q] : 0.0f;
        const uint32_t w32[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        int32_t t1 = -1, t2 = -1;
        float s1 = 0.0f, s2 = 0.0f;
        auto put = [&](int32_t tg, float v) {
            if (tg == t1) {
                s1 += v;
            } else if (tg == t2) {
                s2 += v;
            } else if (t1 < 0) {
                t1 = tg;
                s1 = v;
            } else if (t2 < 0) {
                t2 = tg;
                s2 = v;
            } else {
                atomicAdd(&delta[tg], v);
            }
        };
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (d[q] == 0.0f) continue;  // past nd, or an ant without a deposit: stale slots
            put((int32_t)(w32[q] & 0xffffu), d[q]);
            put((int32_t)(w32[q] >> 16), d[q]);
        }
        combine_add(delta, t1, s1, lane);
        combine_add(delta, t2, s2, lane);
    }
}


// Second half of pass 2: the row's sums go to tau_next and the LDS row is
// cleared for the next row.
__device__ __forceinline__ void adj_apply_row(const DepositArgs& a, int32_t x, float* delta, int t, int nthr) {
    float* tn = a.tau_next + (size_t)x * a.P;
    for (int32_t j = t; j < a.n; j += nthr) {
        const float v = delta[j];
        if (v != 0.0f) {
            tn[j] = tn[j] + v;
            delta[j] = 0.0f;
        }
    }
}

// Block (g, band) = blockIdx.x % groups, / groups: ants g A .. g A + A - 1,
// rows band X .. band X + X - 1.  16 waves: one or two tours per wave.
__global__ __launch_bounds__(1024) void k_adj_scatter_staged(const DepositArgs a, int32_t A, int32_t X) {
    extern __shared__ __align__(16) uint16_t adj_stage[];
    stamp_start(a.stamp);
    if (a.state->skip) return;
    const int32_t groups = (a.m + A - 1) / A;
    const int32_t g = blockIdx.x % groups, band = blockIdx.x / groups;
    adj_scatter_staged(a, g * A, A, band * X, X, adj_stage);
}

// One block of 1024 threads per row: the LDS row holds n floats (n <=
// ADJ_MAX_N); zeroing and applying it take n / 1024 passes.
__global__ __launch_bounds__(1024) void k_adj_gather(const DepositArgs a) {
    extern __shared__ float delta_row[];
    if (a.state->skip) return;
    for (int32_t j = threadIdx.x; j < a.n; j += blockDim.x) delta_row[j] = 0.0f;
    __syncthreads();
    adj_gather_row(a, blockIdx.x, delta_row, threadIdx.x, blockDim.x);
    __syncthreads();
    adj_apply_row(a, blockIdx.x, delta_row, threadIdx.x, blockDim.x);
}

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Batched instances.  Every kernel maps its global index (row, ant, instance
// or work item) to an instance and rebuilds that instance's argument struct,
// then runs the same device code as the single-instance kernels; instance i
// uses Philox stream i, so its tours equal a single colony's with stream_id i.
__device__ __forceinline__ SweepArgs batch_sweep_args(const BatchBuffers& b, const InstMeta& im) {
    SweepArgs a{};
    a.tau = b.tau_cur + im.mat_off;
    a.dist = b.dist + im.mat_off;
    a.eta = b.eta;
    a.eta_len = b.eta_len;
    a.nn = b.nn + im.nk_off;
    a.k = b.k;
    a.choice = b.choice + im.mat_off;
    a.choice_nn = b.choice_nn + im.nk_off;
    a.eta_nn = b.eta_nn + im.nk_off;
    a.tau_nn = nullptr;
    a.tau_next = b.tau_next + im.mat_off;
    a.mean = nullptr;
    a.age = nullptr;
    a.m = im.m;
    a.n = im.n;
    a.P = im.P;
    a.alpha = b.alpha;
    a.one_minus_rho = b.one_minus_rho;
    a.want_mean = 0;
    a.eta_pow = b.eta_pow;
    return a;
}

__device__ __forceinline__ ConstructArgs batch_construct_args(const BatchBuffers& b, int32_t inst,
                                                              const InstMeta& im) {
    ConstructArgs a{};
    a.nn = b.nn + im.nk_off;
    a.nn16 = b.nn16 + im.nk_off;
    a.choice_nn = b.choice_nn + im.nk_off;
    a.tau_nn = nullptr;
    a.eta_nn = b.eta_nn + im.nk_off;
    a.choice = b.choice + im.mat_off;
    a.dist = b.dist + im.mat_off;
    a.tau = b.tau_cur + im.mat_off;
    a.mean = nullptr;
    a.eta = b.eta;
    a.eta_len = b.eta_len;
    a.tours = b.tours + im.tour_off;
    a.lengths = b.lengths + im.ant_off;
    a.age = nullptr;
    a.valid = nullptr;
    a.state = b.state + inst;
    a.n = im.n;
    a.P = im.P;
    a.m = im.m;
    a.k = b.k;
    a.aging = 0;
    a.min_age = 1;
    a.alpha = b.alpha;
    a.seed_lo = b.seed_lo;
    a.seed_hi = b.seed_hi;
    a.stream = (uint32_t)inst;
    a.words = im.words;
    a.young_blocks = 0;
    return a;
}

__device__ __forceinline__ DepositArgs batch_deposit_args(const BatchBuffers& b, int32_t inst, const InstMeta& im) {
    DepositArgs a{};
    a.tours = b.tours + im.tour_off;
    a.best_tour = b.best_tours + im.city_off;
    a.dep_ant = b.dep_ant + im.dep_off;
    a.dep_val = b.dep_val + im.dep_off;
    a.state = b.state + inst;
    a.tau_next = b.tau_next + im.mat_off;
    a.n = im.n;
    a.P = im.P;
    a.dense = b.variant == 0;
    a.adj = b.adj ? b.adj + im.adj_off : nullptr;
    a.m = im.m;
    a.list_cap = im.m + b.w + 2;
    return a;
}

__global__ __launch_bounds__(256) void kb_dist(const BatchBuffers b) {
    const int32_t r = blockIdx.y;
    const InstMeta im = b.meta[b.row_inst[r]];
    const int32_t i = r - im.city_off;
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= im.P) return;
    const int32_t* xy = b.xy + 2 * (size_t)im.city_off;
    int32_t d = ACO_MAX_DISTANCE;
    if (j < im.n && j != i)
        d = euc2d_exact((int64_t)xy[2 * i] - xy[2 * j], (int64_t)xy[2 * i + 1] - xy[2 * j + 1]);
    b.dist[im.mat_off + (size_t)i * im.P + j] = d;
}

__global__ __launch_bounds__(256) void kb_nn(const BatchBuffers b) {
    const int32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= b.total_rows) return;
    const InstMeta im = b.meta[b.row_inst[r]];
    const int32_t i = r - im.city_off;
    nn_row(b.dist + im.mat_off + (size_t)i * im.P, im.n, b.k, i, b.nn + im.nk_off + (size_t)i * b.k,
           threadIdx.x & 63);
}

__global__ __launch_bounds__(256) void kb_nn_aux(const BatchBuffers b) {
    const int32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= b.total_rows) return;
    const InstMeta im = b.meta[b.row_inst[r]];
    const int32_t i = r - im.city_off;
    for (int32_t q = lane; q < b.k; q += 64) {
        const size_t x = im.nk_off + (size_t)i * b.k + q;
        const int32_t j = b.nn[x];
        b.nn16[x] = (uint16_t)j;
        b.eta_nn[x] = eta_of(b.eta, b.eta_len, b.dist[im.mat_off + (size_t)i * im.P + j]);
    }
}

// One block per instance: C_nn, then tau0 for the instance's policy.
__global__ __launch_bounds__(1024) void kb_greedy(const BatchBuffers b) {
    extern __shared__ __align__(16) uint32_t sh[];
    __shared__ int32_t s_next, s_step;
    const int32_t inst = blockIdx.x;
    const InstMeta im = b.meta[inst];
    const int32_t words = (im.n + 31) / 32;
    greedy_block(b.dist + im.mat_off, im.n, im.P, 0, b.cnn + inst, sh,
                 reinterpret_cast<uint64_t*>(sh + ((words + 1) & ~1)), s_next, s_step, nullptr, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double cnn = (double)b.cnn[inst];
        float t0 = b.tau0_const;
        if (b.tau0_policy == 1) t0 = (float)((double)im.m / cnn);
        if (b.tau0_policy == 2) t0 = (float)(0.5 * b.w * (b.w - 1) / ((double)b.rho * cnn));
        b.tau0[inst] = t0;
    }
}

__global__ __launch_bounds__(256) void kb_fill_tau(const BatchBuffers b) {
    const int32_t r = blockIdx.y;
    const int32_t inst = b.row_inst[r];
    const InstMeta im = b.meta[inst];
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= im.P) return;
    const size_t x = im.mat_off + (size_t)(r - im.city_off) * im.P + j;
    b.tau_cur[x] = j < im.n ? b.tau0[inst] : 0.0f;
    b.tau_next[x] = 0.0f;
}

__global__ __launch_bounds__(256) void kb_sweep(const BatchBuffers b) {
    const int32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= b.total_rows) return;
    const InstMeta im = b.meta[b.row_inst[r]];
    sweep_row(batch_sweep_args(b, im), r - im.city_off, threadIdx.x & 63, nullptr);
}

// The persistent construction's work queue: QUEUE_BINS bins of items, each
// bin holding whole instances (longest first, bins balanced by work).
// Blocks are dispatched round-robin over the XCDs, so block b takes items
// from bin b mod 8 first: an instance's candidate data, choice rows and
// distances are then read through one XCD's L2.  A block whose bin is empty
// moves on to the next bins.  Thread 0 only; `bin` starts at the block's
// home bin and only grows.  Returns the item index, or -1 once every bin is
// empty.
__device__ __forceinline__ int32_t batch_next_item(const BatchBuffers& b, int32_t& bin) {
    const int32_t home = blockIdx.x % QUEUE_BINS;
    for (; bin < home + QUEUE_BINS; ++bin) {
        const int32_t h = bin % QUEUE_BINS;
        const int32_t k = (int32_t)atomicAdd(&b.queue[h], 1u);
        if (k < b.bin_off[h + 1] - b.bin_off[h]) return b.bin_off[h] + k;
    }
    return -1;
}

// One persistent launch: blocks take work items (instance, W ants) from the
// binned queue, stage the instance's candidate lists and weights in LDS and
// construct W tours (one per wave).
template <bool VREG_VIS>
__device__ __forceinline__ void kb_construct_body(const BatchBuffers& b, uint32_t* lds, int32_t& s_item) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int32_t bin = blockIdx.x % QUEUE_BINS;
    for (;;) {
        if (threadIdx.x == 0) s_item = batch_next_item(b, bin);
        __syncthreads();
        const int32_t item = s_item;
        __syncthreads();
        if (item < 0) return;
        const int2 it = b.items[item];
        const InstMeta im = b.meta[it.x];
        const ConstructArgs a = batch_construct_args(b, it.x, im);
        const int32_t nk = im.n * b.k;
        const StagedLayout L = staged_layout(lds, nk);
        float* wl = L.wl;
        uint16_t* nl = L.nl;
        uint32_t* vis = L.per + (size_t)wave * im.words;
        stage_lists_weights(nl, a.nn16, wl, a.choice_nn, nk);
        staged_zero_entry(L, nk);
        __syncthreads();
        const int32_t ant = it.y + wave;
        if (ant < im.m) {
            if (a.k <= 32)
                construct_staged_ant<VREG_VIS, true>(a, ant, false, 0.0f, wl, nl, vis, a.state->iter, lane);
            else
                construct_staged_ant<VREG_VIS, false>(a, ant, false, 0.0f, wl, nl, vis, a.state->iter, lane);
        }
        __syncthreads();
    }
}

// Register bitmap (n <= 2048), at most 64 VGPRs (eight waves per SIMD):
// with the candidate data staged the loop is VALU-issue bound once the SIMDs
// are full, and the few spills this costs sit on the Philox refill path.
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void kb_construct_vreg(
    const BatchBuffers b) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ int32_t s_item;
    kb_construct_body<true>(b, lds, s_item);
}

// LDS bitmap (any n; at most 8 ants per block): its fallback keeps 16 row
// loads in flight per lane, so the register budget stays at 256.
__global__ __launch_bounds__(512) void kb_construct_lds(const BatchBuffers b) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ int32_t s_item;
    kb_construct_body<false>(b, lds, s_item);
}

// ---------------------------------------------------------------------------
// Several ants per wave (batched instances with small k and n): G ants share
// a wave, each on its own group of L = 64 / G lanes, when k <= L and n <= 32 L.
// Lane i of group g holds candidate i of ant g's current row and word i of
// its visited bitmap, so every step does for G ants at once what the one-ant
// loop does for one: the scan, ballots and compares stay inside each group
// (the DPP row shifts never cross a 16-lane row, and for L = 32 the
// row_bcast:15 step joins rows 0-1 and 2-3 only), which gives every group's
// prefixes the same bits as the one-ant scan.  The random words are the
// one-ant stream's: step s uses Philox block 64 (s / 256) + (s mod 64),
// word (s / 64) mod 4; lane i of group g computes blocks i, i + L, ... of its
// ant's 256-step chunk, and the step's word is read from lane s mod L of the
// current queue entry.  Tours go to global memory in L-step segments, and
// the edge lengths of each segment are loaded then and summed a segment
// later, so no tour is staged in LDS.
__device__ __forceinline__ int32_t bperm_i(int32_t src_lane, int32_t v) {
    return __builtin_amdgcn_ds_bpermute(src_lane << 2, v);
}

template <int G>
__device__ __forceinline__ float group_scan(float v) {
    if constexpr (G == 4) {
        v = v + dpp_src<0x111, 0xf>(v);
        v = v + dpp_src<0x112, 0xf>(v);
        v = v + dpp_src<0x114, 0xf>(v);
        v = v + dpp_src<0x118, 0xf>(v);
        return v;
    } else {
        return wave_inclusive_scan32(v);
    }
}

// The group's last lane's value in every lane of the group.
template <int G>
__device__ __forceinline__ float group_last(float p, int lane) {
    if constexpr (G == 4) {
        // row_newbcast:15: each 16-lane row reads its own lane 15
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(p), 0x15F, 0xf, 0xf, false));
    } else {
        const float t0 = readlane_f(p, 31);
        const float t1 = readlane_f(p, 63);
        return lane < 32 ? t0 : t1;
    }
}

template <int G>
__device__ __forceinline__ void construct_group_wave(const ConstructArgs& a, int32_t ant0, const int2* cand,
                                                     uint32_t iter, int lane) {
    constexpr int L = 64 / G;
    const int g = lane / L;
    const int i = lane & (L - 1);
    const int gb = g * L;
    const int32_t ant = ant0 + g;
    const bool valid = ant < a.m;

    // The uniforms of the current 256-step chunk in step order, consumed
    // from q[0] (4 G registers; the VGPR budget at this kernel's occupancy
    // has room, and each Philox block is computed once).
    float q[4 * G];
    uint32_t first = 0;  // step 0's word (lane i = 0 of the first block)
    auto refill = [&](int32_t s) {
        const uint32_t blk0 = ((uint32_t)s >> 8) << 6;
#pragma unroll
        for (int b = 0; b < G; ++b) {
            const Philox4 r = philox4x32_10(blk0 + b * L + i, (uint32_t)ant, iter, a.stream, a.seed_lo, a.seed_hi);
            if (b == 0) first = r.x0;
            q[b] = word_to_uniform(r.x0);
            q[G + b] = word_to_uniform(r.x1);
            q[2 * G + b] = word_to_uniform(r.x2);
            q[3 * G + b] = word_to_uniform(r.x3);
        }
    };
    auto advance = [&]() {
#pragma unroll
        for (int t = 0; t + 1 < 4 * G; ++t) q[t] = q[t + 1];
    };
    refill(0);
    const int32_t start = word_to_city((uint32_t)bperm_i(gb, (int32_t)first), a.n);
    uint32_t vreg = (i == (start >> 5)) ? (1u << (start & 31)) : 0u;
    // row stride per lane (0 for lanes >= k): a 24-bit multiply, n k < 2^24
    const uint32_t kk = i < a.k ? (uint32_t)a.k : 0u;
    const int32_t loff = i < a.k ? i : a.n * a.k;
    int32_t c = start;
    // per-lane row base; an entry address is one 24-bit multiply-add
    const char* candb = reinterpret_cast<const char*>(cand + loff);
    const uint32_t kk8 = 8u * kk;
    int2 cj = *reinterpret_cast<const int2*>(candb + __umul24((uint32_t)c, kk8));
    int32_t tbuf = start;
    int32_t prev_last = 0;
    int64_t len = 0;
    int32_t pend = 0;
    unsigned long long fb = 0, scanned = 0;
    int32_t* tour = a.tours + (size_t)ant * a.n;
    // steps in L-step segments: the uniform refill or advance comes before
    // a segment and the tour write after it, not tested at every step
    for (int32_t s0 = 0; s0 < a.n; s0 += L) {
        if (s0 > 0) {
            if ((s0 & 255) == 0)
                refill(s0);
            else
                advance();
        }
        const int32_t s_end = min(s0 + L, a.n);
        for (int32_t s = s0 > 0 ? s0 : 1; s < s_end; ++s) {
            const int32_t j = cj.x;
            const float cw = __int_as_float(cj.y);
            const float u = __int_as_float(
                __builtin_amdgcn_ds_bpermute((gb << 2) + ((s & (L - 1)) << 2), __float_as_int(q[0])));
            // j >> 3 addresses lane j >> 5 of the group (bpermute drops the two
            // low address bits; gb * 4 and j >> 3 share no bits for n <= 32 L);
            // the bit field extract reads bit j mod 32
            // (the signed extract of bit j mod 32 is an all-ones mask for a
            // visited city, which clears the weight's bits)
            const int32_t word = __builtin_amdgcn_ds_bpermute((gb << 2) + (j >> 3), (int)vreg);
            const int32_t vmask = __builtin_amdgcn_sbfe(word, (uint32_t)j, 1u);
            const float w = __int_as_float(__float_as_int(cw) & ~vmask);
            const unsigned long long pos = __builtin_amdgcn_ballot_w64(w > 0.0f);
            const float p = group_scan<G>(w);
            const float tot = group_last<G>(p, lane);
            const float rr = u * tot;
            const unsigned long long hit = __builtin_amdgcn_ballot_w64(p > rr) & pos;
            // first hit of the group; no hit gives L (another group's lane, the
            // result is then replaced below)
            int32_t f;
            if constexpr (L == 16)
                f = __builtin_ctz(((uint32_t)(hit >> gb) & 0xffffu) | 0x10000u);
            else
                f = __builtin_ctzll((hit >> gb) | (1ull << 32));
            int32_t next = bperm_i(gb + f, j);
            const unsigned long long nohit = __builtin_amdgcn_ballot_w64(f == L);
            if (__builtin_expect(nohit != 0ull, 0)) {
    #pragma unroll 1
                for (int gg = 0; gg < G; ++gg) {
                    if (!((nohit >> (gg * L)) & 1ull)) continue;
                    const float tg = readlane_f(tot, gg * L);
                    const uint32_t pg = (uint32_t)((pos >> (gg * L)) & (L == 32 ? 0xffffffffull : 0xffffull));
                    int32_t ng;
                    if (tg > 0.0f && pg != 0u) {  // rounding: the group's last positive lane
                        ng = readlane_i(j, gg * L + 31 - __clz((int)pg));
                    } else {
                        const bool vg = ant0 + gg < a.m;
                        fb += vg ? 1 : 0;
                        unsigned long long sc = 0;
                        // the group's row over the whole wave (n <= 1024 for G = 2, 512 for G = 4)
                        ng = fallback_argmax_regs<G == 4 ? 8 : 16>(a, vreg, gg * L, readlane_i(c, gg * L), false,
                                                                   0.0f, lane, sc);
                        scanned += vg ? sc : 0ull;
                    }
                    next = (g == gg) ? ng : next;
                }
            }
            cj = *reinterpret_cast<const int2*>(candb + __umul24((uint32_t)next, kk8));
            vreg |= (i == (next >> 5)) ? (1u << (next & 31)) : 0u;
            tbuf = (i == s - s0) ? next : tbuf;
            c = next;
        }
        {
            // segment s0 .. s_end - 1: write it, add the previous segment's
            // edge lengths and load this one's (edge into entry s0 + i)
            const int32_t cnt = s_end - s0;
            // with every lane active: bpermute reads 0 from an inactive lane,
            // and a conditional operand would run it only where taken
            const int32_t left = bperm_i(lane - 1, tbuf);
            const int32_t from = i == 0 ? prev_last : left;
            len += pend;
            pend = 0;
            if (valid && i < cnt) {
                tour[s0 + i] = tbuf;
                if (s0 + i > 0) pend = a.dist[(size_t)from * a.P + tbuf];
            }
            prev_last = bperm_i(gb + L - 1, tbuf);
        }
    }
    len += pend;
    if (valid && i == 0) len += a.dist[(size_t)c * a.P + start];
#pragma unroll
    for (int off = L / 2; off >= 1; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)(uint64_t)len, off);
        const uint32_t hi = __shfl_xor((uint32_t)((uint64_t)len >> 32), off);
        len += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    if (valid && i == 0) a.lengths[ant] = len;
    finish_counters(a, fb, scanned, lane);
}

// Candidate table of the grouped construction: entry e = (city, weight bits)
// of candidate e mod k of row e / k, then one (0, 0) entry that lanes >= k
// read; one 8-byte LDS read per step.
__device__ __forceinline__ void stage_cand(int2* cand, const uint16_t* __restrict__ nn16,
                                           const float* __restrict__ w, int32_t nk) {
    for (int32_t x = threadIdx.x; x < nk; x += blockDim.x) cand[x] = make_int2((int32_t)nn16[x], __float_as_int(w[x]));
    if (threadIdx.x == 0) cand[nk] = make_int2(0, 0);
}

// Persistent batched construction with G ants per wave: work items of
// W * G ants, candidate data staged as in kb_construct, no per-wave LDS.
template <int G>
__global__ __launch_bounds__(1024) void kb_construct_group(const BatchBuffers b) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ int32_t s_item;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int32_t bin = blockIdx.x % QUEUE_BINS;
    for (;;) {
        if (threadIdx.x == 0) s_item = batch_next_item(b, bin);
        __syncthreads();
        const int32_t item = s_item;
        __syncthreads();
        if (item < 0) return;
        const int2 it = b.items[item];
        const InstMeta im = b.meta[it.x];
        const ConstructArgs a = batch_construct_args(b, it.x, im);
        const int32_t nk = im.n * b.k;
        int2* cand = reinterpret_cast<int2*>(lds);
        stage_cand(cand, a.nn16, a.choice_nn, nk);
        __syncthreads();
        const int32_t ant0 = it.y + wave * G;
        if (ant0 < im.m) construct_group_wave<G>(a, ant0, cand, a.state->iter, lane);
        __syncthreads();
    }
}

// One block per instance; block 0 also re-arms the construction queue bins.
__global__ __launch_bounds__(256) void kb_stats(const BatchBuffers b) {
    __shared__ LenKey red[16];
    __shared__ int32_t redi[16];
    __shared__ int32_t s_improved, s_recruit;
    extern __shared__ uint32_t taken[];
    __shared__ LenKey red_rank[16 * RANK_MAX];
    __shared__ StatsShared sh;
    const int32_t inst = blockIdx.x;
    const InstMeta im = b.meta[inst];
    StatsArgs a{};
    a.tours = b.tours + im.tour_off;
    a.lengths = b.lengths + im.ant_off;
    a.valid = nullptr;
    a.age = nullptr;
    a.perf = nullptr;
    a.prev_len = nullptr;
    a.best_tour = b.best_tours + im.city_off;
    a.dep_ant = b.dep_ant + im.dep_off;
    a.dep_val = b.dep_val + im.dep_off;
    a.state = b.state + inst;
    a.n = im.n;
    a.m = im.m;
    a.variant = b.variant;
    a.w = b.w < im.n ? b.w : im.n;
    a.aging = 0;
    stats_block(a, red, redi, taken, s_improved, s_recruit, red_rank, sh);
    if (inst == 0 && threadIdx.x < QUEUE_BINS) b.queue[threadIdx.x] = 0u;
}

// Island exchange (batched copies of one instance): every island whose
// best-so-far is longer than the shortest one over all islands (ties: the
// lowest island) takes that tour and length.  One block per island; the
// winner is never written, so the copies race with nothing.
__global__ __launch_bounds__(256) void kb_exchange(const BatchBuffers b) {
    __shared__ LenKey red[4];
    const int32_t me = blockIdx.x;
    LenKey v{INT64_MAX, 0x7fffffff};
    for (int32_t i = threadIdx.x; i < b.count; i += blockDim.x) {
        const LenKey k{b.state[i].glob_best, i};
        if (key_less(k, v)) v = k;
    }
    const LenKey win = block_min_key(v, red);
    if (win.ant == me || b.state[me].glob_best <= win.len) return;
    const InstMeta wi = b.meta[win.ant];
    const InstMeta mi = b.meta[me];
    for (int32_t t = threadIdx.x; t < mi.n; t += blockDim.x)
        b.best_tours[mi.city_off + t] = b.best_tours[wi.city_off + t];
    if (threadIdx.x == 0) b.state[me].glob_best = win.len;
}

// One staged scatter block per (instance, first ant) item; 16 waves, so
// each wave reads two of the item's 32 tours.
__global__ __launch_bounds__(1024) void kb_adj_scatter(const BatchBuffers b) {
    extern __shared__ __align__(16) uint16_t adj_stage_b[];
    const int2 it = b.adj_items[blockIdx.x];
    const InstMeta im = b.meta[it.x];
    adj_scatter_staged(batch_deposit_args(b, it.x, im), it.y, b.adj_ants, 0, im.n, adj_stage_b);
}

// R consecutive rows per wave, four waves per block; each wave's LDS holds R
// rows of nmax floats (R = gather_rows(nmax)).  When a wave's rows belong to
// one instance (all but the waves that straddle an instance boundary) the
// instance's metadata, its ant weights and the slots of all its rows are
// loaded together, the runs of equal targets summed into the LDS rows, and
// the rows' nonzero sums added to tau with their loads issued together:
// a few memory round trips per wave instead of several per row.  Other
// waves take their rows one at a time.
template <int R>
__global__ __launch_bounds__(256) void kb_adj_gather(const BatchBuffers b, int32_t nmax) {
    constexpr int GATHER_ROWS = R;
    extern __shared__ float delta_rows[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* dl = delta_rows + (size_t)wave * GATHER_ROWS * nmax;
    const int32_t r0 = (blockIdx.x * 4 + wave) * GATHER_ROWS;
    if (r0 >= b.total_rows) return;
    const int32_t inst = b.row_inst[r0];
    const bool one = r0 + GATHER_ROWS <= b.total_rows && b.row_inst[r0 + GATHER_ROWS - 1] == inst;
    if (!one) {
        for (int q = 0; q < GATHER_ROWS && r0 + q < b.total_rows; ++q) {
            const int32_t r = r0 + q;
            const int32_t iq = b.row_inst[r];
            const InstMeta im = b.meta[iq];
            const DepositArgs a = batch_deposit_args(b, iq, im);
            for (int32_t j = lane; j < a.n; j += 64) dl[j] = 0.0f;
            wave_lds_sync();
            adj_gather_row(a, r - im.city_off, dl, lane, 64);
            wave_lds_sync();
            adj_apply_row(a, r - im.city_off, dl, lane, 64);
            wave_lds_sync();
        }
        return;
    }
    const InstMeta im = b.meta[inst];
    const DepositArgs a = batch_deposit_args(b, inst, im);
    const int32_t x0 = r0 - im.city_off;
    for (int32_t j = lane; j < GATHER_ROWS * a.n; j += 64) dl[(j / a.n) * nmax + j % a.n] = 0.0f;
    wave_lds_sync();
    const size_t pitch = (size_t)adj_pitch(a.m);
    const int32_t nd = min(a.m, a.state->ndep);
    for (int32_t base = 0; base < nd; base += 512) {
        const int32_t a0 = base + 8 * lane;
        const bool act = a0 < nd;
        float d[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) d[q] = (act && a0 + q < nd) ? a.dep_val[a0 + q] : 0.0f;
        uint4 v[GATHER_ROWS][2];
#pragma unroll
        for (int r = 0; r < GATHER_ROWS; ++r) {
            v[r][0] = make_uint4(0, 0, 0, 0);
            v[r][1] = make_uint4(0, 0, 0, 0);
            if (act) {
                const uint16_t* row = a.adj + (size_t)(x0 + r) * pitch + 2 * a0;
                v[r][0] = *reinterpret_cast<const uint4*>(row);
                v[r][1] = *reinterpret_cast<const uint4*>(row + 8);
            }
        }
#pragma unroll
        for (int r = 0; r < GATHER_ROWS; ++r) {
            const uint32_t w32[8] = {v[r][0].x, v[r][0].y, v[r][0].z, v[r][0].w,
                                     v[r][1].x, v[r][1].y, v[r][1].z, v[r][1].w};
            float* delta = dl + (size_t)r * nmax;
            // two (target, sum) accumulators over the lane's 16 deposits
            // (successor and predecessor of 8 ants): the row's two tour
            // neighbours once the colony agrees, whichever direction each
            // ant travelled; a third target goes straight to LDS
            int32_t t1 = -1, t2 = -1;
            float s1 = 0.0f, s2 = 0.0f;
            auto put = [&](int32_t t, float x) {
                if (t == t1) {
                    s1 += x;
                } else if (t == t2) {
                    s2 += x;
                } else if (t1 < 0) {
                    t1 = t;
                    s1 = x;
                } else if (t2 < 0) {
                    t2 = t;
                    s2 = x;
                } else {
                    atomicAdd(&delta[t], x);
                }
            };
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (d[q] == 0.0f) continue;  // past nd, or an ant without a deposit: stale slots
                put((int32_t)(w32[q] & 0xffffu), d[q]);
                put((int32_t)(w32[q] >> 16), d[q]);
            }
            combine_add(delta, t1, s1, lane);
            combine_add(delta, t2, s2, lane);
        }
    }
    wave_lds_sync();
    for (int32_t j = lane; j < a.n; j += 64) {
        float dv[GATHER_ROWS], tv[GATHER_ROWS];
#pragma unroll
        for (int r = 0; r < GATHER_ROWS; ++r) dv[r] = dl[(size_t)r * nmax + j];
#pragma unroll
        for (int r = 0; r < GATHER_ROWS; ++r)
            tv[r] = dv[r] != 0.0f ? a.tau_next[(size_t)(x0 + r) * a.P + j] : 0.0f;
#pragma unroll
        for (int r = 0; r < GATHER_ROWS; ++r)
            if (dv[r] != 0.0f) a.tau_next[(size_t)(x0 + r) * a.P + j] = tv[r] + dv[r];
    }
}

// Ant System batch deposit with fp32 atomics (no adjacency buffer): one
// block per ant.
__global__ __launch_bounds__(256) void kb_deposit(const BatchBuffers b) {
    const int32_t inst = b.ant_inst[blockIdx.x];
    const int32_t e = blockIdx.x - b.meta[inst].ant_off;
    const InstMeta im = b.meta[inst];
    deposit_tour(batch_deposit_args(b, inst, im), e, false);
}

// Rank-based batch deposit: block b takes global rows 64 b .. 64 b + 63,
// one band per instance those rows belong to, each deposited as the single
// colony's (deposit_rows_band): the same ordered fp32 sums per instance.
constexpr int32_t BATCH_DEP_ROWS = 64;
__global__ __launch_bounds__(1024) void kb_deposit_rows(const BatchBuffers b) {
    __shared__ DepositRowsShared<BATCH_DEP_ROWS> sh;
    const int32_t g0 = blockIdx.x * BATCH_DEP_ROWS, g1 = min(g0 + BATCH_DEP_ROWS, b.total_rows);
    for (int32_t g = g0; g < g1;) {
        const int32_t inst = b.row_inst[g];
        const InstMeta im = b.meta[inst];
        const int32_t x0 = g - im.city_off, rows = min(g1 - g, im.n - x0);
        deposit_rows_band<BATCH_DEP_ROWS>(batch_deposit_args(b, inst, im), x0, rows, false, sh);
        g += rows;
    }
}

// ---------------------------------------------------------------------------
// Launchers.
static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

__global__ __launch_bounds__(64) void k_stamp_end(uint64_t* stamp, int64_t* host) {
    if (threadIdx.x == 0) {
        stamp[4] = wall_clock64();
        volatile int64_t* h = host;
        for (int i = 0; i < 5; ++i) h[i] = (int64_t)stamp[i];
        __threadfence_system();
    }
}

hipError_t launch_stamp_end(uint64_t* stamp, int64_t* host, hipStream_t s) {
    hipLaunchKernelGGL(k_stamp_end, dim3(1), dim3(64), 0, s, stamp, host);
    return hipGetLastError();
}

hipError_t launch_dist_from_xy(const int32_t* xy, int32_t n, int32_t P, int32_t* dist, hipStream_t s) {
    hipLaunchKernelGGL(k_dist_from_xy, dim3(cdiv(P, 256), n), dim3(256), 0, s, xy, n, P, dist);
    return hipGetLastError();
}

hipError_t launch_eta_full(const int32_t* dist, const float* eta, int32_t eta_len, int32_t n, int32_t P, float* out,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_eta_full, dim3(cdiv(P, 256), n), dim3(256), 0, s, dist, eta, eta_len, n, P, out);
    return hipGetLastError();
}

hipError_t launch_fix_diagonal(int32_t* dist, int32_t n, int32_t P, hipStream_t s) {
    hipLaunchKernelGGL(k_fix_diagonal, dim3(cdiv(n, 256)), dim3(256), 0, s, dist, n, P);
    return hipGetLastError();
}

hipError_t launch_dist_max(const int32_t* dist, int32_t n, int32_t P, int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dist_max, dim3(n < 1024 ? n : 1024), dim3(256), 0, s, dist, n, P, out);
    return hipGetLastError();
}

hipError_t launch_nn_lists(const int32_t* dist, int32_t n, int32_t P, int32_t k, int32_t* nn, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_lists, dim3(cdiv(n, 4)), dim3(256), 0, s, dist, n, P, k, nn);
    return hipGetLastError();
}

hipError_t launch_fill_rows(float* a, int32_t n, int32_t P, float v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_rows, dim3(cdiv(P, 256), n), dim3(256), 0, s, a, n, P, v);
    return hipGetLastError();
}

hipError_t launch_greedy(const int32_t* dist, int32_t n, int32_t P, int32_t start, int64_t* out, const int32_t* nn,
                         int32_t k, hipStream_t s) {
    const int32_t words = (n + 31) / 32;
    const size_t lds = (size_t)((words + 1) & ~1) * 4 + 16 * sizeof(uint64_t);
    hipLaunchKernelGGL(k_greedy, dim3(1), dim3(1024), lds, s, dist, n, P, start, out, nn, k);
    return hipGetLastError();
}

// A launch that records the events it is given (start and/or stop) from
// its own dispatch; without events a plain launch.
template <typename F, typename... Args>
static inline void launch_k(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, hipEvent_t start,
                            hipEvent_t stop, Args... args) {
    if (start != nullptr || stop != nullptr)
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)lds, s, start, stop, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

template <int SRC, int PW>
static void launch_sweep_t(const SweepArgs& a, size_t lds, hipStream_t s, const LaunchTiming& t) {
    if (a.alpha == 1.0f)
        launch_k(k_sweep<SRC, PW, true>, dim3(cdiv(a.n, 4)), dim3(256), lds, s, t.start, t.stop, a);
    else
        launch_k(k_sweep<SRC, PW, false>, dim3(cdiv(a.n, 4)), dim3(256), lds, s, t.start, t.stop, a);
}

hipError_t launch_sweep(const SweepArgs& a, hipStream_t s, const LaunchTiming& t) {
    const size_t lds = a.top != nullptr ? (size_t)4 * TOP_BUF * sizeof(uint64_t) : 0;
    // heuristic source: eta_full row, else distances from the coordinates or
    // the matrix, as 1 / d^eta_pow or from the table
    if (a.eta_full != nullptr) {
        launch_sweep_t<0, 0>(a, lds, s, t);
    } else {
        const bool xy = a.xy != nullptr;
        switch (a.eta_pow) {
            case 1: xy ? launch_sweep_t<2, 1>(a, lds, s, t) : launch_sweep_t<1, 1>(a, lds, s, t); break;
            case 2: xy ? launch_sweep_t<2, 2>(a, lds, s, t) : launch_sweep_t<1, 2>(a, lds, s, t); break;
            case 3: xy ? launch_sweep_t<2, 3>(a, lds, s, t) : launch_sweep_t<1, 3>(a, lds, s, t); break;
            default: xy ? launch_sweep_t<2, 0>(a, lds, s, t) : launch_sweep_t<1, 0>(a, lds, s, t); break;
        }
    }
    return hipGetLastError();
}

hipError_t launch_eta_check(int32_t eta_pow, const float* eta, int32_t eta_len, int32_t* flag, hipStream_t s) {
    const dim3 g(cdiv(eta_len, 256));
    switch (eta_pow) {
        case 1: hipLaunchKernelGGL(k_eta_check<1>, g, dim3(256), 0, s, eta, eta_len, flag); break;
        case 2: hipLaunchKernelGGL(k_eta_check<2>, g, dim3(256), 0, s, eta, eta_len, flag); break;
        case 3: hipLaunchKernelGGL(k_eta_check<3>, g, dim3(256), 0, s, eta, eta_len, flag); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t staged_lds_bytes(int32_t n, int32_t k, int32_t words, int32_t ants_per_block) {
    const size_t nk = (size_t)n * k;
    return ((nk + 4) & ~(size_t)3) * 4 + ((nk + 8) & ~(size_t)7) * 2 + (size_t)ants_per_block * words * 4;
}

hipError_t launch_construct(const ConstructArgs& a, int32_t path, int32_t ants_per_block, hipStream_t s,
                            const LaunchTiming& t) {
    const int W = ants_per_block;
    if (path == 0) {
        const size_t lds = (size_t)W * a.words * sizeof(uint32_t);
        launch_k(k_construct_cand, dim3(cdiv(a.m, W)), dim3(64 * W), lds, s, t.start, t.stop, a);
    } else if (path == 2) {
        const size_t lds = staged_lds_bytes(a.n, a.k, a.words, W);
        launch_k(k_construct_staged, dim3(cdiv(a.m, W) + a.young_blocks), dim3(64 * W), lds, s, t.start, t.stop, a);
    } else {
        const size_t lds = (size_t)W * full_lds_words(a.n, a.words) * 4;
        launch_k(k_construct_full, dim3(cdiv(a.m, W)), dim3(64 * W), lds, s, t.start, t.stop, a);
    }
    return hipGetLastError();
}

hipError_t launch_nn_aux(const int32_t* nn, const int32_t* dist, const float* eta, int32_t eta_len, int32_t n,
                         int32_t P, int32_t k, uint16_t* nn16, float* eta_nn, hipStream_t s) {
    hipLaunchKernelGGL(k_nn_aux, dim3(cdiv((int64_t)n * k, 256)), dim3(256), 0, s, nn, dist, eta, eta_len, n, P, k,
                       nn16, eta_nn);
    return hipGetLastError();
}

hipError_t launch_stats(const StatsArgs& a, hipStream_t s, const LaunchTiming& t) {
    const size_t lds = (size_t)((a.m + 31) / 32) * sizeof(uint32_t);
    launch_k(k_stats, dim3(1), dim3(1024), lds, s, t.start, t.stop, a);
    return hipGetLastError();
}

hipError_t launch_deposit(const DepositArgs& a, int32_t grid, hipStream_t s, const LaunchTiming& t) {
    launch_k(k_deposit, dim3(grid), dim3(256), 0, s, t.start, t.stop, a);
    return hipGetLastError();
}

bool stats_deposit_fusable(int32_t variant, int32_t n, int32_t m, int32_t w) {
    return variant != 0 && n >= 2 && n <= 4096 && m >= 1 && m <= 1024 && w >= 1 && w <= DEP_PASS &&
           w - 1 <= RANK_MAX;
}

hipError_t launch_stats_deposit(const StatsArgs& s, const DepositArgs& d, hipStream_t st, const LaunchTiming& t) {
    if (!stats_deposit_fusable(s.variant, s.n, s.m, s.w) || d.dense) return hipErrorInvalidValue;
    launch_k(k_stats_deposit, dim3(cdiv(s.n, 16)), dim3(1024), 0, st, t.start, t.stop, s, d);
    return hipGetLastError();
}

hipError_t launch_deposit_rows(const DepositArgs& a, hipStream_t s, const LaunchTiming& t) {
    if (a.dense) return hipErrorInvalidValue;
    if (a.n <= 4096)
        launch_k(k_deposit_rows<16>, dim3(cdiv(a.n, 16)), dim3(1024), 0, s, t.start, t.stop, a);
    else
        launch_k(k_deposit_rows<256>, dim3(cdiv(a.n, 256)), dim3(1024), 0, s, t.start, t.stop, a);
    return hipGetLastError();
}

hipError_t launch_deposit_adj(const DepositArgs& a, hipStream_t s, const LaunchTiming& t) {
    if (!a.adj || !a.dense || a.n > ADJ_MAX_N) return hipErrorInvalidValue;
    const int32_t A = adj_stage_ants(a.n);
    if (A) {
        launch_k(k_adj_scatter_staged, dim3(cdiv(a.m, A)), dim3(1024), adj_stage_lds(a.n, A), s, t.start, nullptr, a,
                 A, a.n);
    } else {  // bands of ADJ_BAND_ROWS rows, ADJ_BAND_ANTS ants per block
        const int32_t bands = cdiv(a.n, ADJ_BAND_ROWS);
        launch_k(k_adj_scatter_staged, dim3(cdiv(a.m, ADJ_BAND_ANTS) * bands), dim3(1024),
                 adj_stage_lds(ADJ_BAND_ROWS, ADJ_BAND_ANTS), s, t.start, nullptr, a, (int32_t)ADJ_BAND_ANTS,
                 (int32_t)ADJ_BAND_ROWS);
    }
    launch_k(k_adj_gather, dim3(a.n), dim3(1024), (size_t)a.n * sizeof(float), s, nullptr, t.stop, a);
    return hipGetLastError();
}

size_t batch_construct_lds_bytes(int32_t nmax, int32_t k, int32_t ants_per_block, int32_t ants_per_wave) {
    // several ants per wave keep bitmaps in registers and tours in global
    // memory: only the (city, weight) candidate table
    if (ants_per_wave > 1) return 8 * ((size_t)nmax * k + 1);
    return staged_lds_bytes(nmax, k, (nmax + 31) / 32, ants_per_block);
}

int32_t batch_max_ants_per_wave(int32_t nmax, int32_t k) {
    if (k <= 16 && nmax <= 512) return 4;
    if (k <= 32 && nmax <= 1024) return 2;
    return 1;
}

int32_t batch_construct_blocks_per_cu(int32_t nmax, int32_t ants_per_block, int32_t ants_per_wave,
                                      size_t lds_bytes) {
    int nb = 0;
    const void* f = ants_per_wave == 4   ? reinterpret_cast<const void*>(kb_construct_group<4>)
                    : ants_per_wave == 2 ? reinterpret_cast<const void*>(kb_construct_group<2>)
                    : nmax <= 2048       ? reinterpret_cast<const void*>(kb_construct_vreg)
                                         : reinterpret_cast<const void*>(kb_construct_lds);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * ants_per_block, lds_bytes) != hipSuccess) nb = 1;
    return nb > 0 ? nb : 1;
}

// Initialisation: distances, candidate lists, heuristic on candidate edges,
// C_nn and tau0, pheromone fill.
hipError_t launch_batch_init(const BatchBuffers& b, int32_t pmax, int32_t nmax, hipStream_t s) {
    hipLaunchKernelGGL(kb_dist, dim3(cdiv(pmax, 256), b.total_rows), dim3(256), 0, s, b);
    hipLaunchKernelGGL(kb_nn, dim3(cdiv(b.total_rows, 4)), dim3(256), 0, s, b);
    hipLaunchKernelGGL(kb_nn_aux, dim3(cdiv(b.total_rows, 4)), dim3(256), 0, s, b);
    const int32_t words = (nmax + 31) / 32;
    const size_t glds = (size_t)((words + 1) & ~1) * 4 + 16 * sizeof(uint64_t);
    hipLaunchKernelGGL(kb_greedy, dim3(b.count), dim3(1024), glds, s, b);
    hipLaunchKernelGGL(kb_fill_tau, dim3(cdiv(pmax, 256), b.total_rows), dim3(256), 0, s, b);
    return hipGetLastError();
}

// which: 0 sweep, 1 construct (grid = persistent blocks), 2 stats, 3 deposit.
hipError_t launch_batch_iteration_kernel(const BatchBuffers& b, int32_t which, int32_t nmax, int32_t grid,
                                         hipStream_t s) {
    switch (which) {
        case 0:
            hipLaunchKernelGGL(kb_sweep, dim3(cdiv(b.total_rows, 4)), dim3(256), 0, s, b);
            break;
        case 1: {
            // several ants per wave, else one: register bitmap (n <= 2048)
            // or LDS bitmap; one instantiation each keeps the common
            // kernel's register count low
            const size_t lds = batch_construct_lds_bytes(nmax, b.k, b.ants_per_block, b.ants_per_wave);
            const dim3 blk(64 * b.ants_per_block);
            if (b.ants_per_wave == 4)
                hipLaunchKernelGGL(kb_construct_group<4>, dim3(grid), blk, lds, s, b);
            else if (b.ants_per_wave == 2)
                hipLaunchKernelGGL(kb_construct_group<2>, dim3(grid), blk, lds, s, b);
            else if (nmax <= 2048)
                hipLaunchKernelGGL(kb_construct_vreg, dim3(grid), blk, lds, s, b);
            else
                hipLaunchKernelGGL(kb_construct_lds, dim3(grid), blk, lds, s, b);
            break;
        }
        case 2:  // grid carries the RBAS ranking bitmap words, (max m + 31) / 32
            hipLaunchKernelGGL(kb_stats, dim3(b.count), dim3(256), (size_t)grid * sizeof(uint32_t), s, b);
            break;
        default:
            if (b.variant == 0 && b.adj) {
                hipLaunchKernelGGL(kb_adj_scatter, dim3(b.nadj_items), dim3(1024), adj_stage_lds(nmax, b.adj_ants), s,
                                   b);
                const int R = gather_rows(nmax);
                const size_t glds = 4 * (size_t)R * nmax * sizeof(float);
                if (R == 4)
                    hipLaunchKernelGGL(kb_adj_gather<4>, dim3(cdiv(b.total_rows, 16)), dim3(256), glds, s, b, nmax);
                else
                    hipLaunchKernelGGL(kb_adj_gather<1>, dim3(cdiv(b.total_rows, 4)), dim3(256), glds, s, b, nmax);
            } else if (b.variant == 0) {
                hipLaunchKernelGGL(kb_deposit, dim3(b.total_ants), dim3(256), 0, s, b);
            } else {
                hipLaunchKernelGGL(kb_deposit_rows, dim3(cdiv(b.total_rows, BATCH_DEP_ROWS)), dim3(1024), 0, s, b);
            }
    }
    return hipGetLastError();
}

hipError_t launch_batch_exchange(const BatchBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(kb_exchange, dim3(b.count), dim3(256), 0, s, b);
    return hipGetLastError();
}

// Latency floor of the staged construction step: one wave repeats the
// step's dependent chain and nothing else -- the LDS reads of the candidate
// entry and weight at the current city (24-bit multiply-add addresses, lanes
// >= k on a zero entry), the ds_bpermute visited test, the roulette scan,
// ballot and pick of roulette_pick, the lane read of the step's uniform --
// over a 64-city table.  out = {core clock cycles, wall-clock ticks} of
// `iters` steps; sink keeps the chain live.
template <bool K32>
__global__ __launch_bounds__(64) void k_probe_step(int32_t iters, int32_t k, unsigned long long* out,
                                                   int32_t* sink) {
    __shared__ uint16_t nl[64 * 64 + 1];
    __shared__ float wl[64 * 64 + 1];
    const int lane = threadIdx.x;
    for (int e = lane; e < 64 * 64 + 1; e += 64) {
        const int c = e / 64, r = e % 64;
        nl[e] = (uint16_t)((c + 1 + 7 * r) & 63);
        wl[e] = e == 64 * 64 ? 0.0f : 1.0f + (float)((c * 31 + r * 17) & 15);
    }
    __syncthreads();
    const uint32_t vreg = lane < 2 ? 0x5a5a5a5au : 0u;  // about half of the 64 cities visited
    const uint32_t kk = lane < k ? (uint32_t)k : 0u;
    const int32_t loff = lane < k ? lane : 64 * 64;
    const char* nlb = reinterpret_cast<const char*>(nl + loff);
    const char* wlb = reinterpret_cast<const char*>(wl + loff);
    const uint32_t kk2 = 2u * kk, kk4 = 4u * kk;
    const float uc = (float)((lane * 40503u) & 0xffffu) * (1.0f / 65536.0f);
    int32_t c = 0;
    int32_t j = *reinterpret_cast<const uint16_t*>(nlb);
    float cw = *reinterpret_cast<const float*>(wlb);
    const unsigned long long t0 = clock64(), w0 = wall_clock64();
    for (int32_t it = 0; it < iters; ++it) {
        const float u = readlane_f(uc, it & 63);
        __builtin_assume((uint32_t)j < 65536u);
        const int32_t word = __builtin_amdgcn_ds_bpermute(j >> 3, (int)vreg);
        const int32_t vmask = __builtin_amdgcn_sbfe(word, (uint32_t)j, 1u);
        const float w = __int_as_float(__float_as_int(cw) & ~vmask);
        const int32_t next = roulette_pick<K32>(w, j, u, [&]() { return (c + 1) & 63; });
        j = *reinterpret_cast<const uint16_t*>(nlb + __umul24((uint32_t)next, kk2));
        cw = *reinterpret_cast<const float*>(wlb + __umul24((uint32_t)next, kk4));
        c = next;
    }
    const unsigned long long t1 = clock64(), w1 = wall_clock64();
    if (lane == 0) {
        out[0] = t1 - t0;
        out[1] = w1 - w0;
        sink[0] = c;
    }
}

hipError_t launch_probe_step(int32_t iters, int32_t k, unsigned long long* out, int32_t* sink, hipStream_t s) {
    if (k <= 32)
        hipLaunchKernelGGL(k_probe_step<true>, dim3(1), dim3(64), 0, s, iters, k, out, sink);
    else
        hipLaunchKernelGGL(k_probe_step<false>, dim3(1), dim3(64), 0, s, iters, k, out, sink);
    return hipGetLastError();
}

hipError_t launch_copy(const float4* src, float4* dst, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, src, dst, count);
    return hipGetLastError();
}

}  // namespace aco
