// Device helpers shared by the gfx950 kernels: counter-based RNG, the wave-64
// DPP prefix sum, wave-level synchronisation.  Wave64 only (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define ACO_MAX_DISTANCE 0x7fffffff
#define ACO_MIN_PHEROMONE 1e-18f  // fp32 nearest of MIN_PHEROMONE (inc/default.h:17)

namespace aco {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter {step/4, ant, iteration,
// stream}, key {seed lo, seed hi}; one call yields the words of four
// consecutive steps.  Replaces cuRAND XORWOW (inc/Colony.cuh:58-72).
struct Philox4 {
    uint32_t x0, x1, x2, x3;
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                 uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return {c0, c1, c2, c3};
}

__device__ __forceinline__ uint32_t philox_word(const Philox4& p, uint32_t i) {
    return i == 0 ? p.x0 : (i == 1 ? p.x1 : (i == 2 ? p.x2 : p.x3));
}

// 24 random bits -> [0,1), exact in fp32.
__device__ __forceinline__ float word_to_uniform(uint32_t w) {
    return (float)(w >> 8) * 5.9604644775390625e-08f;
}

// floor(w * n / 2^32): always in [0, n).
__device__ __forceinline__ int32_t word_to_city(uint32_t w, int32_t n) {
    return (int32_t)__umulhi(w, (uint32_t)n);
}

// ---------------------------------------------------------------------------
// Wave-64 inclusive prefix sum: row_shr 1,2,4,8 inside each 16-lane row, then
// row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.  Lanes
// without a source receive +0 (old = 0), so every add is one IEEE fp32 add;
// the CPU oracle emulates exactly this order.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_src(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

__device__ __forceinline__ float wave_inclusive_scan(float v) {
    v = v + dpp_src<0x111, 0xf>(v);  // row_shr:1
    v = v + dpp_src<0x112, 0xf>(v);  // row_shr:2
    v = v + dpp_src<0x114, 0xf>(v);  // row_shr:4
    v = v + dpp_src<0x118, 0xf>(v);  // row_shr:8
    v = v + dpp_src<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3
    v = v + dpp_src<0x143, 0xc>(v);  // row_bcast:31 -> rows 2, 3
    return v;
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Orders this wave's LDS accesses (a wave writes its own visited bitmap and
// reads it back on the next step).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 64-bit wave minimum through DPP: the scan pattern of wave_inclusive_scan
// (row shifts 1, 2, 4, 8, then the two row broadcasts); sources outside a
// row keep the identity.  Lane 63 ends with the result, which every lane
// receives through v_readlane.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v, uint32_t identity) {
    const uint32_t lo =
        (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)(uint32_t)v, CTRL, ROW_MASK, 0xf, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)(uint32_t)(v >> 32), CTRL, ROW_MASK, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t lane63_u64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
    v = min(v, dpp_u64<0x111, 0xf>(v, ~0u));
    v = min(v, dpp_u64<0x112, 0xf>(v, ~0u));
    v = min(v, dpp_u64<0x114, 0xf>(v, ~0u));
    v = min(v, dpp_u64<0x118, 0xf>(v, ~0u));
    v = min(v, dpp_u64<0x142, 0xa>(v, ~0u));
    v = min(v, dpp_u64<0x143, 0xc>(v, ~0u));
    return lane63_u64(v);
}

// DPP source with the lane's own value where the pattern has no source lane
// (row shifts past the row start, rows outside ROW_MASK).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int32_t dpp_keep_i(int32_t v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROW_MASK, 0xf, false);
}

// Wave argmax of per-lane (best, index) pairs: the largest value, ties to the
// lowest index; best = -1 means "no candidate" (weights above -1 only; NaN
// never becomes a lane's best).  The maximum comes from a 32-bit float max
// over the scan pattern of wave_inclusive_scan (one v_max per step); the
// lanes holding it are one ballot, and when several do (equal weights) a
// minimum over their indices decides.  -0 and +0 compare equal, as in the
// sequential scan.  0x7fffffff if no lane has a candidate.
__device__ __forceinline__ int32_t wave_argmax_index(float best, int32_t idx) {
    float m = best;
    m = fmaxf(m, __int_as_float(dpp_keep_i<0x111, 0xf>(__float_as_int(m))));  // row_shr:1
    m = fmaxf(m, __int_as_float(dpp_keep_i<0x112, 0xf>(__float_as_int(m))));  // row_shr:2
    m = fmaxf(m, __int_as_float(dpp_keep_i<0x114, 0xf>(__float_as_int(m))));  // row_shr:4
    m = fmaxf(m, __int_as_float(dpp_keep_i<0x118, 0xf>(__float_as_int(m))));  // row_shr:8
    m = fmaxf(m, __int_as_float(dpp_keep_i<0x142, 0xa>(__float_as_int(m))));  // row_bcast:15
    m = fmaxf(m, __int_as_float(dpp_keep_i<0x143, 0xc>(__float_as_int(m))));  // row_bcast:31
    m = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 63));
    if (!(m > -1.0f)) return 0x7fffffff;
    const unsigned long long hit = __builtin_amdgcn_ballot_w64(best == m);
    if (__builtin_expect(__popcll(hit) == 1, 1)) return __builtin_amdgcn_readlane(idx, __ffsll((long long)hit) - 1);
    int32_t v = best == m ? idx : 0x7fffffff;
    v = min(v, dpp_keep_i<0x111, 0xf>(v));
    v = min(v, dpp_keep_i<0x112, 0xf>(v));
    v = min(v, dpp_keep_i<0x114, 0xf>(v));
    v = min(v, dpp_keep_i<0x118, 0xf>(v));
    v = min(v, dpp_keep_i<0x142, 0xa>(v));
    v = min(v, dpp_keep_i<0x143, 0xc>(v));
    return __builtin_amdgcn_readlane(v, 63);
}

// Heuristic lookup: table by integer distance, 0 outside (diagonal).
__device__ __forceinline__ float eta_of(const float* __restrict__ eta, int32_t eta_len, int32_t d) {
    return ((uint32_t)d < (uint32_t)eta_len) ? eta[d] : 0.0f;
}

}  // namespace aco
