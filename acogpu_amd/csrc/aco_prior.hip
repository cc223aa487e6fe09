// Learned edge prior: features on candidate edges, an 8 -> 64 -> 64 -> 1 MLP
// (GELU, bf16 weights and activations, fp32 accumulate) on gfx950 MFMA, its
// backward pass, Adam, and the heuristic it writes for the candidate edges.
// No counterpart in the reference; it replaces eta^beta on NN-list edges.
//
// MFMA v_mfma_f32_16x16x16_bf16 (wave64), lane l, q = 0..3:
//   A (16 x 16): A[l % 16][4 (l / 16) + q]
//   B (16 x 16): B[4 (l / 16) + q][l % 16]
//   D (16 x 16): D[4 (l / 16) + q][l % 16]
// Every layer is evaluated transposed, Z^T = W H^T with edges as columns:
// D of one layer (features 4(l/16)+q of edge l%16) is then exactly the B
// operand of the next, so the forward pass needs no data movement.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "aco_device.h"
#include "aco_prior.h"

namespace aco {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 bf16x2h __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x4h __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8h __attribute__((ext_vector_type(8)));

// Four fp32 values rounded to bf16 (nearest even) by gfx950's
// v_cvt_pk_bf16_f32, two per instruction.
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    const f32x4 v = {a, b, c, d};
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bf16x4h));
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// gfx950's 16x16x32 bf16 product with each lane's eight k values given as
// two 4-value halves.  k is summed over, so A and B only have to list the
// same k values in the same order in each lane: the halves are taken from
// two 16-row blocks of a D-layout tile (rows 4(l/16)+q of each), which is
// the layout the previous layer's product leaves.
__device__ __forceinline__ f32x4 mfma32(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c) {
    const bf16x8h a = __builtin_bit_cast(bf16x8h, __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7));
    const bf16x8h b = __builtin_bit_cast(bf16x8h, __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7));
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// GELU(z) = z Phi(z), Phi from the erf approximation 7.1.26 of Abramowitz
// and Stegun (|error of erf| <= 1.5e-7): with u = |z| / sqrt(2),
// erf(u) ~ 1 - P(t) e^{-u^2}, t = 1 / (1 + p u), P(t) = a1 t + ... + a5 t^5,
// so Phi = 1 - Q for z >= 0 and Q for z < 0, Q = P(t) e^{-z^2/2} / 2 (the
// coefficients below are the a_i / 2).  The derivative is GELU'(z) =
// Phi(z) + z phi(z) with the exact normal density phi(z) = e^{-z^2/2} /
// sqrt(2 pi), which reuses the exponential: only Phi is approximated, so it
// is within 7.5e-8 of the exact derivative.  One hardware exponential
// (v_exp_f32) and one reciprocal (v_rcp_f32) per value.
constexpr float GELU_H1 = 0.127414796f, GELU_H2 = -0.142248368f, GELU_H3 = 0.7107068705f;
constexpr float GELU_H4 = -0.7265760135f, GELU_H5 = 0.5307027145f;
constexpr float GELU_C = 0.231641888f;       // 0.3275911 / sqrt(2)
constexpr float GELU_HLOG2E = 0.72134752f;   // log2(e) / 2
constexpr float GELU_RSQRT2PI = 0.398942280f;  // 1 / sqrt(2 pi)

template <bool GRAD>
__device__ __forceinline__ float gelu_eval(float z, float& dg) {
    const float e = __builtin_amdgcn_exp2f((z * z) * -GELU_HLOG2E);  // e^{-z^2/2}
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(GELU_C, fabsf(z), 1.0f));
    const float poly = __builtin_fmaf(
        t, __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, GELU_H5, GELU_H4), GELU_H3), GELU_H2), GELU_H1);
    const float q = (t * poly) * e;
    const float phi = z >= 0.0f ? 1.0f - q : q;
    if (GRAD) dg = __builtin_fmaf(z * e, GELU_RSQRT2PI, phi);
    return z * phi;
}

__device__ __forceinline__ float gelu(float z) {
    float unused = 0.0f;
    return gelu_eval<false>(z, unused);
}

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

constexpr float LOG2E_F = 1.44269504f;  // log2(e)
constexpr float LN2_F = 0.693147181f;   // ln(2)

// ---------------------------------------------------------------------------
// Features of candidate edge (i, r), j = nn[i][r]; D_i^q = distance from i to
// its q-th nearest neighbour.  Integer ratios are rounded once:
//   f0 d_ij / (D_i^{k-1} + 1)      f1 d_ij / (D_j^{k-1} + 1)
//   f2 r / k                        f3 rank of i in NN(j) / k (1 if absent)
//   f4 d_ij n / (sum_i D_i^0 + n)   (fp64, rounded once)
//   f5 D_i^0 / (D_i^{k-1} + 1)      f6 D_j^0 / (D_j^{k-1} + 1)
//   f7 (d_ij - D_i^0) / (D_i^{k-1} - D_i^0 + 1)
__device__ __forceinline__ int32_t nn_dist(const PriorInstance& in, int32_t i, int32_t q) {
    return in.dist[(size_t)i * in.P + in.nn[(size_t)i * in.k + q]];
}

// Sum over the cities of the nearest-neighbour distance, by one block (an
// integer sum: exact in any order).  The result is written, not accumulated,
// so no buffer has to be cleared first.
__global__ __launch_bounds__(1024) void k_nn0_sum(const PriorInstance in, unsigned long long* sum) {
    __shared__ unsigned long long part[16];
    unsigned long long v = 0ull;
    for (int32_t i = threadIdx.x; i < in.n; i += 1024) v += (unsigned long long)nn_dist(in, i, 0);
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0ull;
        for (int w = 0; w < 16; ++w) t += part[w];
        *sum = t;
    }
}

__global__ __launch_bounds__(256) void k_features(const PriorInstance in, const unsigned long long* sum,
                                                  float* __restrict__ feat) {
    const int32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= in.n * in.k) return;
    const int32_t i = e / in.k, r = e % in.k;
    const int32_t j = in.nn[e];
    const int32_t d = in.dist[(size_t)i * in.P + j];
    const int32_t di0 = nn_dist(in, i, 0), dik = nn_dist(in, i, in.k - 1);
    const int32_t dj0 = nn_dist(in, j, 0), djk = nn_dist(in, j, in.k - 1);
    int32_t rank = in.k;
    for (int32_t q = 0; q < in.k; ++q)
        if (in.nn[(size_t)j * in.k + q] == i) {
            rank = q;
            break;
        }
    float* f = feat + (size_t)e * 8;
    f[0] = (float)d / (float)(dik + 1);
    f[1] = (float)d / (float)(djk + 1);
    f[2] = (float)r / (float)in.k;
    f[3] = (float)rank / (float)in.k;
    f[4] = (float)((double)d * (double)in.n / (double)(*sum + (unsigned long long)in.n));
    f[5] = (float)di0 / (float)(dik + 1);
    f[6] = (float)dj0 / (float)(djk + 1);
    f[7] = (float)(d - di0) / (float)(dik - di0 + 1);
}

// Position of every city on a tour, then the label of each candidate edge:
// 1 when j follows or precedes i on the tour.  A colony that has not built a
// tour yet holds no permutation (its best tour is zeroed), so positions
// start at -1, entries outside [0, n) are skipped, and a city without a
// position labels its edges 0.
__global__ __launch_bounds__(256) void k_pos_clear(int32_t n, int32_t* __restrict__ pos) {
    const int32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) pos[t] = -1;
}

__global__ __launch_bounds__(256) void k_tour_pos(const int32_t* __restrict__ tour, int32_t n,
                                                  int32_t* __restrict__ pos) {
    const int32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) {
        const int32_t c = tour[t];
        if ((uint32_t)c < (uint32_t)n) pos[c] = t;
    }
}

__global__ __launch_bounds__(256) void k_labels(const PriorInstance in, const int32_t* __restrict__ tour,
                                                const int32_t* __restrict__ pos, float* __restrict__ y) {
    const int32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= in.n * in.k) return;
    const int32_t i = e / in.k, j = in.nn[e];
    const int32_t t = pos[i];
    if ((uint32_t)t >= (uint32_t)in.n) {
        y[e] = 0.0f;
        return;
    }
    const int32_t nx = tour[t + 1 < in.n ? t + 1 : 0], pv = tour[t > 0 ? t - 1 : in.n - 1];
    y[e] = (nx == j || pv == j) ? 1.0f : 0.0f;
}

// ---------------------------------------------------------------------------
// MLP tiles.  One wave handles 16 edges per tile.  Weights (bf16): W1 is
// 64 x 16 (inputs 8..15 zero), W2 64 x 64, row = output feature.
struct MlpRegs {
    bf16x4 w1a[4];  // layer 1 A operands, output block ob
};

// Layer-2 weights as MFMA A operands (16 bytes per lane per fragment, read
// from LDS with one ds_read_b128), written by k_pack_bf16 after every update
// of the parameters and copied to LDS by every block that uses them:
//   fwd [ob][c][lane]  W2[16 ob + l%16][16 (2c) + 4(l/16) + 0..3] then
//                      W2[16 ob + l%16][16 (2c+1) + 4(l/16) + 0..3]
//                      (z2 = W2 h1, K pairs of h1's 16-row blocks)
//   bwd [ib][c][lane]  W2[16 (2c) + 4(l/16) + 0..3][16 ib + l%16] then
//                      W2[16 (2c+1) + 4(l/16) + 0..3][16 ib + l%16]
//                      (dh1 = W2^T dz2, K pairs of dz2's blocks)
struct W2Frags {
    uint4 fwd[4 * 2 * 64];
    uint4 bwd[4 * 2 * 64];
};

__device__ __forceinline__ void stage_w2(const MlpParams& p, W2Frags* F, bool backward) {
    uint4* dst = F->fwd;  // bwd follows fwd
    const int n = backward ? MLP_W2_FRAGS : MLP_W2_FRAGS / 2;
    for (int f = threadIdx.x; f < n; f += blockDim.x) dst[f] = p.w2_frag[f];
}

// The two 4-value halves of a staged fragment.
__device__ __forceinline__ void frag_halves(const uint4& v, bf16x4& a0, bf16x4& a1) {
    a0 = bf16x4{(short)(v.x & 0xffffu), (short)(v.x >> 16), (short)(v.y & 0xffffu), (short)(v.y >> 16)};
    a1 = bf16x4{(short)(v.z & 0xffffu), (short)(v.z >> 16), (short)(v.w & 0xffffu), (short)(v.w >> 16)};
}

// b1, b2, w3 (64 each) and b3, staged in LDS by every MLP kernel.
enum : int { VEC_B1 = 0, VEC_B2 = 64, VEC_W3 = 128, VEC_B3 = 192, VEC_LEN = 196 };

__device__ __forceinline__ void stage_vectors(const MlpParams& p, float* vec) {
    for (int i = threadIdx.x; i < VEC_LEN; i += blockDim.x) {
        float v = 0.0f;
        if (i < 64)
            v = p.master[MLP_B1 + i];
        else if (i < 128)
            v = p.master[MLP_B2 + i - 64];
        else if (i < 192)
            v = p.master[MLP_W3 + i - 128];
        else if (i == 192)
            v = p.master[MLP_B3];
        vec[i] = v;
    }
    __syncthreads();
}

// Four consecutive vector entries of output features 16 ob + 4 (l/16) + q.
__device__ __forceinline__ float4 vec4(const float* vec, int base, int ob, int grp) {
    return *reinterpret_cast<const float4*>(vec + base + 16 * ob + 4 * grp);
}

__device__ __forceinline__ void load_mlp(const MlpParams& p, MlpRegs& R, int lane) {
    for (int ob = 0; ob < 4; ++ob) R.w1a[ob] = __builtin_bit_cast(bf16x4, p.w1_frag[ob * 64 + lane]);
}

// Input tile as the layer-1 B operand: features 4 (l/16) + q of edge l % 16.
__device__ __forceinline__ bf16x4 load_x(const float* __restrict__ x, int64_t e0, int64_t count, int lane) {
    const int64_t e = e0 + (lane & 15);
    const int grp = lane >> 4;
    if (grp >= 2 || e >= count) return bf16x4{0, 0, 0, 0};
    const float4 v = *reinterpret_cast<const float4*>(x + (size_t)e * 8 + 4 * grp);
    return pack4(v.x, v.y, v.z, v.w);
}

// Forward of one tile; with GRAD it keeps each unit's GELU derivative for
// the backward pass.
struct Activations {
    f32x4 d1[4], d2[4];  // GELU'(z1), GELU'(z2) (GRAD only)
    bf16x4 h1[4], h2[4];
    float out;  // logit of edge l % 16 (valid in every lane group)
};

template <bool GRAD>
__device__ __forceinline__ bf16x4 gelu4(const f32x4& z, f32x4& d) {
    float g[4];
    for (int q = 0; q < 4; ++q) {
        float dq = 0.0f;
        g[q] = gelu_eval<GRAD>(z[q], dq);
        d[q] = dq;
    }
    return pack4(g[0], g[1], g[2], g[3]);
}

template <bool GRAD = false>
__device__ __forceinline__ void forward_tile(const MlpRegs& R, const W2Frags* F, const float* vec, bf16x4 xb,
                                             Activations& A, int lane) {
    const int grp = lane >> 4;
    // the biases are the products' initial accumulator values
    for (int ob = 0; ob < 4; ++ob) {
        const float4 b = vec4(vec, VEC_B1, ob, grp);
        const f32x4 z = mfma16(R.w1a[ob], xb, f32x4{b.x, b.y, b.z, b.w});
        A.h1[ob] = gelu4<GRAD>(z, A.d1[ob]);
    }
    float partial = 0.0f;
    for (int ob = 0; ob < 4; ++ob) {
        const float4 b = vec4(vec, VEC_B2, ob, grp);
        f32x4 z = {b.x, b.y, b.z, b.w};
        for (int c = 0; c < 2; ++c) {
            bf16x4 a0, a1;
            frag_halves(F->fwd[(ob * 2 + c) * 64 + lane], a0, a1);
            z = mfma32(a0, a1, A.h1[2 * c], A.h1[2 * c + 1], z);
        }
        A.h2[ob] = gelu4<GRAD>(z, A.d2[ob]);
        const float4 w = vec4(vec, VEC_W3, ob, grp);
        partial = __builtin_fmaf(w.x, bf2f((uint16_t)A.h2[ob][0]), partial);
        partial = __builtin_fmaf(w.y, bf2f((uint16_t)A.h2[ob][1]), partial);
        partial = __builtin_fmaf(w.z, bf2f((uint16_t)A.h2[ob][2]), partial);
        partial = __builtin_fmaf(w.w, bf2f((uint16_t)A.h2[ob][3]), partial);
    }
    // the four lane groups hold disjoint features of the same edge
    partial = partial + __shfl_xor(partial, 16);
    partial = partial + __shfl_xor(partial, 32);
    A.out = partial + vec[VEC_B3];
}

__global__ __launch_bounds__(256) void k_mlp_forward(const MlpParams p, const float* __restrict__ x, int64_t count,
                                                     float* __restrict__ out) {
    __shared__ __align__(16) float vec[VEC_LEN];
    __shared__ __align__(16) uint4 w2f[4 * 2 * 64];
    W2Frags* F = reinterpret_cast<W2Frags*>(w2f);  // forward fragments only
    const int lane = threadIdx.x & 63;
    stage_w2(p, F, false);
    stage_vectors(p, vec);
    MlpRegs R;
    load_mlp(p, R, lane);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t * 16 < count; t += waves) {
        const int64_t e0 = t * 16;
        Activations A;
        forward_tile(R, F, vec, load_x(x, e0, count, lane), A, lane);
        if (lane < 16 && e0 + lane < count) out[e0 + lane] = A.out;
    }
}

// Transposes of D-layout 16x16 bf16 tiles (feature rows 4(l/16)+q, edge
// column l%16) into the A/B layout with edges along K: lane l gets edge
// 4(l/16)+q of feature l%16.  All T tiles go through the wave's LDS scratch
// (T x 256 entries) with one ordering point between the stores and loads.
template <int T>
__device__ __forceinline__ void transpose16_batch(const bf16x4 (&v)[T], bf16x4 (&r)[T], uint16_t* scratch,
                                                  int lane) {
    const int col = lane & 15, grp = lane >> 4;
    for (int t = 0; t < T; ++t)
        for (int q = 0; q < 4; ++q) scratch[t * 256 + (4 * grp + q) * 16 + col] = (uint16_t)v[t][q];
    wave_lds_sync();
    for (int t = 0; t < T; ++t) {
        const uint2 w = *reinterpret_cast<const uint2*>(scratch + t * 256 + col * 16 + 4 * grp);
        r[t] = bf16x4{(short)(w.x & 0xffffu), (short)(w.x >> 16), (short)(w.y & 0xffffu), (short)(w.y >> 16)};
    }
    wave_lds_sync();
}

// Sum over the 16 lanes of each row of the wave (the 16 edges of a tile):
// row_ror 8, 4, 2, 1, after which every lane of the row holds the total
// (each lane in its own fixed order; callers use lane 0 of the row).
__device__ __forceinline__ float row16_sum(float v) {
    v = v + dpp_src<0x128, 0xf>(v);  // row_ror:8
    v = v + dpp_src<0x124, 0xf>(v);  // row_ror:4
    v = v + dpp_src<0x122, 0xf>(v);  // row_ror:2
    v = v + dpp_src<0x121, 0xf>(v);  // row_ror:1
    return v;
}

// Input tile of the backward pass as loaded from global memory (features
// 4 (l/16) + 0..3 of edge l % 16 in lanes of groups 0 and 1; its label in
// every lane), converted after the load has landed.
struct TileIn {
    float4 x;
    float y;
};

__device__ __forceinline__ TileIn load_tile(const float* __restrict__ x, const float* __restrict__ y, int64_t e0,
                                            int64_t count, int lane) {
    const int64_t e = e0 + (lane & 15);
    TileIn r{make_float4(0.f, 0.f, 0.f, 0.f), 0.0f};
    if (e < count) {
        if ((lane >> 4) < 2) r.x = *reinterpret_cast<const float4*>(x + (size_t)e * 8 + 4 * (lane >> 4));
        r.y = y[e];
    }
    return r;
}

// Backward: binary cross-entropy with logits, positive weight `pos_weight`,
// mean over `count` edges.  Every wave accumulates its tiles' gradients --
// W1 and W2 (and b1, through a constant-one input column of the W1 product)
// in MFMA accumulators, b2 and w3 in LDS slots owned by one lane each (read,
// added to and written back once per tile) --
// while the next tile's inputs load.  After the loop the block's waves
// (MLP_BWD_WAVES, two per SIMD: at most 256 registers each) combine their
// gradients in LDS -- the upper half of the waves store rows in accumulator
// order with 16-byte stores, the lower half add theirs to them (wave w to
// row w) -- and the block writes the sum of those rows, taken in row order,
// as one partial row (layout MLP_ROW_*).
__global__ __launch_bounds__(64 * MLP_BWD_WAVES) void k_mlp_backward(const MlpParams p, const float* __restrict__ x,
                                                      const float* __restrict__ y, int64_t count, float pos_weight,
                                                      float* __restrict__ partial) {
    // The loop's LDS (W2 fragments, per-wave transpose scratch and b2 / w3
    // sums) is reused for the gradient rows afterwards.
    struct LoopShared {
        W2Frags F;
        uint16_t tr[MLP_BWD_WAVES][8 * 256];
        float4 bias[MLP_BWD_WAVES][4][2][64];  // wave, ob, (b2, w3), lane: q = 0..3
    };
    union BlockShared {
        LoopShared loop;
        float4 rows[MLP_BWD_WAVES / 2][MLP_ROW / 4];
    };
    __shared__ __align__(16) BlockShared sh;
    __shared__ __align__(16) float vec[VEC_LEN];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane >> 4, col = lane & 15;
    W2Frags* F = &sh.loop.F;
    float4(*bias)[2][64] = sh.loop.bias[wave];
    for (int ob = 0; ob < 4; ++ob) {
        bias[ob][0][lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        bias[ob][1][lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    stage_w2(p, F, true);
    stage_vectors(p, vec);
    MlpRegs R;
    load_mlp(p, R, lane);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gW2[4][4], gW1[4];
    for (int a = 0; a < 4; ++a) {
        gW1[a] = zero;
        for (int b = 0; b < 4; ++b) gW2[a][b] = zero;
    }
    float gb3 = 0.0f, lsum = 0.0f;
    const float inv = 1.0f / (float)count;
    uint16_t* sc = sh.loop.tr[wave];
    // input 8 of every edge is 1 (W1's padding columns are zero, so the
    // forward pass does not see it): column 8 of the W1 gradient is b1's
    const bf16x4 one8 = pack4(1.0f, 0.0f, 0.0f, 0.0f);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    TileIn next = load_tile(x, y, t * 16, count, lane);
    for (; t * 16 < count; t += waves) {
        const int64_t e0 = t * 16;
        const bool valid = e0 + col < count;
        const TileIn cur = next;
        next = load_tile(x, y, (t + waves) * 16, count, lane);
        const bf16x4 xb = grp == 2 ? one8 : pack4(cur.x.x, cur.x.y, cur.x.z, cur.x.w);
        const float yy = cur.y;
        Activations A;
        forward_tile<true>(R, F, vec, xb, A, lane);
        // s = sigmoid(out) and the loss terms -log s = max(-out, 0) + L,
        // -log(1 - s) = max(out, 0) + L with u = e^{-|out|}, L = log(1 + u),
        // s = 1 / (1 + u) for out >= 0 and u / (1 + u) otherwise: one
        // hardware exponential, reciprocal and logarithm
        const float u = __builtin_amdgcn_exp2f(fabsf(A.out) * -LOG2E_F);
        const float r = __builtin_amdgcn_rcpf(1.0f + u);
        const float s = A.out >= 0.0f ? r : u * r;
        // d loss / d logit for -w y log s - (1-y) log(1-s)
        const float g = valid ? (s * (1.0f - yy + pos_weight * yy) - pos_weight * yy) * inv : 0.0f;
        if (grp == 0 && valid) {
            const float L = __builtin_amdgcn_logf(1.0f + u) * LN2_F;
            const float sp_pos = fmaxf(-A.out, 0.0f) + L;  // -log s
            const float sp_neg = fmaxf(A.out, 0.0f) + L;   // -log(1-s)
            lsum += (pos_weight * yy * sp_pos + (1.0f - yy) * sp_neg) * inv;
        }
        if (grp == 0) gb3 += g;
        // layer 3 and dz2 (D layout: features 16ob + 4grp + q of edge col)
        bf16x4 dz2b[4];
        for (int ob = 0; ob < 4; ++ob) {
            float dz[4];
            const float4 w = vec4(vec, VEC_W3, ob, grp);
            const float w3[4] = {w.x, w.y, w.z, w.w};
            for (int q = 0; q < 4; ++q) dz[q] = g * w3[q] * A.d2[ob][q];
            float4 sb = bias[ob][0][lane], sw = bias[ob][1][lane];  // this lane's running sums
            sb.x += dz[0];
            sb.y += dz[1];
            sb.z += dz[2];
            sb.w += dz[3];
            sw.x += g * bf2f((uint16_t)A.h2[ob][0]);
            sw.y += g * bf2f((uint16_t)A.h2[ob][1]);
            sw.z += g * bf2f((uint16_t)A.h2[ob][2]);
            sw.w += g * bf2f((uint16_t)A.h2[ob][3]);
            bias[ob][0][lane] = sb;
            bias[ob][1][lane] = sw;
            dz2b[ob] = pack4(dz[0], dz[1], dz[2], dz[3]);
        }
        // dW2[o][i] = sum_e dz2[o][e] h1[i][e]: A = dz2 (o x e), B = h1^T
        // (e x i); both transposed so that edges run along K: lane (o or i =
        // l%16, e = 4grp+q)
        const bf16x4 tin[8] = {dz2b[0], dz2b[1], dz2b[2], dz2b[3], A.h1[0], A.h1[1], A.h1[2], A.h1[3]};
        bf16x4 tout[8];
        transpose16_batch<8>(tin, tout, sc, lane);
        for (int ob = 0; ob < 4; ++ob)
            for (int ib = 0; ib < 4; ++ib) gW2[ob][ib] = mfma16(tout[ob], tout[4 + ib], gW2[ob][ib]);
        // dh1 = W2^T dz2 (D layout; K = the 64 outputs, in pairs of dz2's
        // blocks), dz1 = dh1 * gelu'(z1)
        bf16x4 t2in[5], t2out[5];
        for (int ib = 0; ib < 4; ++ib) {
            f32x4 dh = zero;
            for (int c = 0; c < 2; ++c) {
                bf16x4 a0, a1;
                frag_halves(F->bwd[(ib * 2 + c) * 64 + lane], a0, a1);
                dh = mfma32(a0, a1, dz2b[2 * c], dz2b[2 * c + 1], dh);
            }
            t2in[ib] = pack4(dh[0] * A.d1[ib][0], dh[1] * A.d1[ib][1], dh[2] * A.d1[ib][2], dh[3] * A.d1[ib][3]);
        }
        // dW1[o][i] = sum_e dz1[o][e] x[i][e]: B = x^T (e x i), i in 0..15
        t2in[4] = xb;
        transpose16_batch<5>(t2in, t2out, sc, lane);
        for (int ob = 0; ob < 4; ++ob) gW1[ob] = mfma16(t2out[ob], t2out[4], gW1[ob]);
    }
    // b2 / w3 gradients: reduce each lane's LDS sums over the 16 edge lanes
    float gb2[4][4], gw3[4][4];
    for (int ob = 0; ob < 4; ++ob) {
        const float4 sb = bias[ob][0][lane], sw = bias[ob][1][lane];
        gb2[ob][0] = sb.x;
        gb2[ob][1] = sb.y;
        gb2[ob][2] = sb.z;
        gb2[ob][3] = sb.w;
        gw3[ob][0] = sw.x;
        gw3[ob][1] = sw.y;
        gw3[ob][2] = sw.z;
        gw3[ob][3] = sw.w;
    }
    for (int ob = 0; ob < 4; ++ob)
        for (int q = 0; q < 4; ++q) {
            gb2[ob][q] = row16_sum(gb2[ob][q]);
            gw3[ob][q] = row16_sum(gw3[ob][q]);
        }
    for (int off = 32; off >= 1; off >>= 1) {
        gb3 += __shfl_xor(gb3, off);
        lsum += __shfl_xor(lsum, off);
    }
    __syncthreads();  // every wave is past its last use of the loop's LDS
    constexpr int HALF = MLP_BWD_WAVES / 2;
    float4* row = sh.rows[wave % HALF];
    // one wave's gradients into `row`: stored (upper half) or added (lower)
    auto put = [&](int idx, float4 v, bool add) {
        if (add) {
            const float4 r = row[idx];
            v = make_float4(r.x + v.x, r.y + v.y, r.z + v.z, r.w + v.w);
        }
        row[idx] = v;
    };
    auto put_all = [&](bool add) {
        for (int ob = 0; ob < 4; ++ob)
            for (int ib = 0; ib < 4; ++ib) {
                const f32x4 g4 = gW2[ob][ib];
                put(MLP_ROW_W2 / 4 + (ob * 4 + ib) * 64 + lane, make_float4(g4[0], g4[1], g4[2], g4[3]), add);
            }
        for (int ob = 0; ob < 4; ++ob) {
            const f32x4 g4 = gW1[ob];
            put(MLP_ROW_W1 / 4 + ob * 64 + lane, make_float4(g4[0], g4[1], g4[2], g4[3]), add);
        }
        if (col == 0)  // lane 16 grp holds its group's row totals
            for (int ob = 0; ob < 4; ++ob) {
                put(MLP_ROW_B2 / 4 + grp * 4 + ob, make_float4(gb2[ob][0], gb2[ob][1], gb2[ob][2], gb2[ob][3]), add);
                put(MLP_ROW_W3 / 4 + grp * 4 + ob, make_float4(gw3[ob][0], gw3[ob][1], gw3[ob][2], gw3[ob][3]), add);
            }
        if (lane == 0) put(MLP_ROW_B3 / 4, make_float4(gb3, lsum, 0.0f, 0.0f), add);
    };
    if (wave >= HALF) put_all(false);
    __syncthreads();
    if (wave < HALF) put_all(true);
    __syncthreads();
    float4* out = reinterpret_cast<float4*>(partial + (size_t)blockIdx.x * MLP_ROW);
    for (int i = threadIdx.x; i < MLP_ROW / 4; i += blockDim.x) {
        float4 sum = sh.rows[0][i];
        for (int w = 1; w < HALF; ++w) {
            const float4 r = sh.rows[w][i];
            sum = make_float4(sum.x + r.x, sum.y + r.y, sum.z + r.z, sum.w + r.w);
        }
        out[i] = sum;
    }
}

// Parameter index of slot s of a partial row (layout MLP_ROW_*); MLP_PARAMS
// for the loss, -1 for a padding slot.  Accumulator slots: lane l, value q
// of block (ob, ib) is row o = 16 ob + 4 (l / 16) + q, column 16 ib + l % 16.
__device__ __forceinline__ int32_t row_slot_param(int32_t s) {
    if (s < MLP_ROW_W1) {
        const int32_t lane = (s >> 2) & 63, q = s & 3, blk = s >> 8;
        const int32_t o = 16 * (blk >> 2) + 4 * (lane >> 4) + q, i = 16 * (blk & 3) + (lane & 15);
        return MLP_W2 + o * 64 + i;
    }
    if (s < MLP_ROW_B2) {
        const int32_t t = s - MLP_ROW_W1, lane = (t >> 2) & 63, q = t & 3;
        const int32_t o = 16 * (t >> 8) + 4 * (lane >> 4) + q, c = lane & 15;
        return c < 8 ? MLP_W1 + o * 8 + c : (c == 8 ? MLP_B1 + o : -1);
    }
    if (s < MLP_ROW_B3) {
        const int32_t t = (s - MLP_ROW_B2) & 63, g = t >> 4, ob = (t >> 2) & 3, q = t & 3;
        return (s < MLP_ROW_W3 ? MLP_B2 : MLP_W3) + 16 * ob + 4 * g + q;
    }
    if (s == MLP_ROW_B3) return MLP_B3;
    if (s == MLP_ROW_LOSS) return MLP_PARAMS;
    return -1;
}

// Sum of the partial rows in a fixed order: block b takes slots 32 b ..
// 32 b + 31; its thread group g = thread / 32 (of 32) adds rows g, g + 32,
// g + 64, ... in order (loads issued together, at most 8 per thread), and
// the group sums are added in order g = 0..31.  grad gets the parameter
// slots, loss accumulates the loss slot.
enum : int {
    REDUCE_COLS = 32,
    REDUCE_GROUPS = 32,
    REDUCE_ROWS_PER_GROUP = MLP_MAX_ROWS / REDUCE_GROUPS,
};
static_assert(REDUCE_GROUPS * REDUCE_ROWS_PER_GROUP == MLP_MAX_ROWS, "every partial row is summed");
static_assert(REDUCE_COLS * REDUCE_GROUPS == 1024, "one 1024-thread block");

__global__ __launch_bounds__(1024) void k_reduce_grad(const float* __restrict__ partial, int32_t rows,
                                                      float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float part[REDUCE_GROUPS][REDUCE_COLS];
    const int32_t c = threadIdx.x % REDUCE_COLS, g = threadIdx.x / REDUCE_COLS;
    const int32_t slot = blockIdx.x * REDUCE_COLS + c;
    float v[REDUCE_ROWS_PER_GROUP];
#pragma unroll
    for (int k = 0; k < REDUCE_ROWS_PER_GROUP; ++k) {
        const int32_t r = g + REDUCE_GROUPS * k;
        v[k] = (slot < MLP_ROW && r < rows) ? partial[(size_t)r * MLP_ROW + slot] : 0.0f;
    }
    float acc = v[0];
#pragma unroll
    for (int k = 1; k < REDUCE_ROWS_PER_GROUP; ++k) acc += v[k];
    part[g][c] = acc;
    __syncthreads();
    if (g == 0 && slot < MLP_ROW) {
        float sum = part[0][c];
        for (int k = 1; k < REDUCE_GROUPS; ++k) sum += part[k][c];
        const int32_t j = row_slot_param(slot);
        if (j >= 0 && j < MLP_PARAMS)
            grad[j] = sum;
        else if (j == MLP_PARAMS)
            *loss += sum;
    }
}

// Adam on the fp32 master parameters; bc1 = 1 - beta1^t and bc2 = 1 - beta2^t
// come from the host.  The gradient is cleared for the next step.
__global__ __launch_bounds__(256) void k_adam(const MlpParams p, float* __restrict__ grad, float* __restrict__ m,
                                              float* __restrict__ v, float lr, float beta1, float beta2, float eps,
                                              float bc1, float bc2) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < MLP_PARAMS) {
        const float g = grad[i];
        const float mi = beta1 * m[i] + (1.0f - beta1) * g;
        const float vi = beta2 * v[i] + (1.0f - beta2) * (g * g);
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / bc1;
        const float vh = vi / bc2;
        p.master[i] = p.master[i] - lr * mh / (sqrtf(vh) + eps);
        grad[i] = 0.0f;
    }
}

__device__ __forceinline__ uint32_t bf_pair(float lo, float hi) {
    return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}

// bf16 operand fragments of W1 (64 x 8, padded to 16 inputs: row 16 ob +
// l%16, inputs 4(l/16) + 0..3) and W2 (layout above) from the fp32 master.
__global__ __launch_bounds__(256) void k_pack_bf16(MlpParams p) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    const float* m = p.master;
    if (i < MLP_W1_FRAGS) {
        const int ob = i >> 6, l = i & 63, o = 16 * ob + (l & 15), c = 4 * (l >> 4);
        const float* w = m + MLP_W1 + o * 8;
        p.w1_frag[i] = c < 8 ? make_uint2(bf_pair(w[c], w[c + 1]), bf_pair(w[c + 2], w[c + 3])) : make_uint2(0u, 0u);
    }
    if (i < MLP_W2_FRAGS) {
        const int f = i & 511, blk = f >> 7, c = (f >> 6) & 1, l = f & 63, r = l & 15, g = l >> 4;
        const int k0 = 16 * (2 * c) + 4 * g, k1 = 16 * (2 * c + 1) + 4 * g;
        const float* w = m + MLP_W2;  // 64 x 64, row = output feature
        if (i < 512) {
            const int o = 16 * blk + r;  // rows o, columns k0.. and k1..
            p.w2_frag[i] = make_uint4(bf_pair(w[o * 64 + k0], w[o * 64 + k0 + 1]),
                                      bf_pair(w[o * 64 + k0 + 2], w[o * 64 + k0 + 3]),
                                      bf_pair(w[o * 64 + k1], w[o * 64 + k1 + 1]),
                                      bf_pair(w[o * 64 + k1 + 2], w[o * 64 + k1 + 3]));
        } else {
            const int in = 16 * blk + r;  // column in, rows k0.. and k1..
            p.w2_frag[i] = make_uint4(bf_pair(w[k0 * 64 + in], w[(k0 + 1) * 64 + in]),
                                      bf_pair(w[(k0 + 2) * 64 + in], w[(k0 + 3) * 64 + in]),
                                      bf_pair(w[k1 * 64 + in], w[(k1 + 1) * 64 + in]),
                                      bf_pair(w[(k1 + 2) * 64 + in], w[(k1 + 3) * 64 + in]));
        }
    }
}

// Heuristic on candidate edges from the score l: transform 0 gives
// sigmoid(l)^gamma (probability), transform 1 gives exp(gamma min(l, 80 /
// gamma)) (odds^gamma, capped); both floored at 1e-12 so that no candidate
// weight is exactly zero.
__global__ __launch_bounds__(256) void k_logit_to_heuristic(const float* __restrict__ logit, int64_t count,
                                                            float gamma, int32_t transform, float* __restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) {
        const float l = logit[i];
        const float v = transform == 0 ? powf(sigmoidf(l), gamma) : expf(fminf(gamma * l, 80.0f));
        h[i] = fmaxf(v, 1e-12f);
    }
}

static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_features(const PriorInstance& in, unsigned long long* sum, float* feat, hipStream_t s) {
    hipLaunchKernelGGL(k_nn0_sum, dim3(1), dim3(1024), 0, s, in, sum);
    hipLaunchKernelGGL(k_features, dim3(cdiv((int64_t)in.n * in.k, 256)), dim3(256), 0, s, in, sum, feat);
    return hipGetLastError();
}

hipError_t launch_labels(const PriorInstance& in, const int32_t* tour, int32_t* pos, float* y, hipStream_t s) {
    hipLaunchKernelGGL(k_pos_clear, dim3(cdiv(in.n, 256)), dim3(256), 0, s, in.n, pos);
    hipLaunchKernelGGL(k_tour_pos, dim3(cdiv(in.n, 256)), dim3(256), 0, s, tour, in.n, pos);
    hipLaunchKernelGGL(k_labels, dim3(cdiv((int64_t)in.n * in.k, 256)), dim3(256), 0, s, in, tour, pos, y);
    return hipGetLastError();
}

static unsigned mlp_grid(int64_t count) {
    const int64_t tiles = (count + 15) / 16;
    const int64_t blocks = (tiles + 3) / 4;
    return (unsigned)(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
}

hipError_t launch_mlp_forward(const MlpParams& p, const float* x, int64_t count, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_grid(count)), dim3(256), 0, s, p, x, count, out);
    return hipGetLastError();
}

static unsigned backward_blocks(int64_t count) {
    const int64_t tiles = (count + 15) / 16;
    const int64_t waves = tiles < MLP_MAX_WAVES ? tiles : MLP_MAX_WAVES;
    const int64_t blocks = (waves + MLP_BWD_WAVES - 1) / MLP_BWD_WAVES;
    return (unsigned)(blocks > 0 ? blocks : 1);
}

hipError_t launch_mlp_backward(const MlpParams& p, const float* x, const float* y, int64_t count, float pos_weight,
                               float* partial, float* grad, float* loss, hipStream_t s) {
    const unsigned blocks = backward_blocks(count);  // one partial row per block
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(64 * MLP_BWD_WAVES), 0, s, p, x, y, count, pos_weight,
                       partial);
    hipLaunchKernelGGL(k_reduce_grad, dim3(cdiv(MLP_ROW, REDUCE_COLS)), dim3(1024), 0, s, partial, (int32_t)blocks,
                       grad, loss);
    return hipGetLastError();
}

hipError_t launch_adam(const MlpParams& p, float* grad, float* m, float* v, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(cdiv(MLP_PARAMS, 256)), dim3(256), 0, s, p, grad, m, v, lr, beta1, beta2, eps,
                       bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_pack_bf16(const MlpParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_bf16, dim3(cdiv(MLP_W2_FRAGS, 256)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_logit_to_heuristic(const float* logit, int64_t count, float gamma, int32_t transform, float* h,
                                     hipStream_t s) {
    hipLaunchKernelGGL(k_logit_to_heuristic, dim3(cdiv(count, 256)), dim3(256), 0, s, logit, count, gamma, transform,
                       h);
    return hipGetLastError();
}

}  // namespace aco

// ---------------------------------------------------------------------------
// C ABI (include/acogpu.h).

using aco::InstanceView;
using aco::MlpParams;

struct aco_prior {
    aco_prior_config cfg;
    hipStream_t stream = nullptr;
    float* d_master = nullptr;
    uint2* d_w1 = nullptr;  // bf16 operand fragments of W1 and W2
    uint4* d_w2 = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_grad = nullptr;
    float* d_loss = nullptr;
    float* d_partial = nullptr;  // backward partial-gradient rows
    unsigned long long* d_sum = nullptr;
    int32_t step = 0;
    // training set: features (count x 8) and labels (count)
    float* d_x = nullptr;
    float* d_y = nullptr;
    int64_t count = 0, cap = 0;
    // per-instance scratch
    float* d_feat = nullptr;
    float* d_logit = nullptr;
    int32_t* d_pos = nullptr;
    int64_t scratch_edges = 0, scratch_cities = 0;
    char err[512] = {0};

    MlpParams params() const { return MlpParams{d_master, d_w1, d_w2}; }
};

static thread_local char g_prior_err[512] = "no error";

static int pfail(aco_prior* pr, int code, const char* fmt, ...) {
    char* out = pr ? pr->err : g_prior_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(out, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define PHIP(pr, expr)                                                                                    \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return pfail((pr), ACO_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static int palloc(aco_prior* pr, T** out, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return pfail(pr, ACO_E_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    *out = static_cast<T*>(q);
    return ACO_OK;
}

template <typename T>
static void pfree(T*& q) {
    if (q) (void)hipFree(q);
    q = nullptr;
}

extern "C" void aco_prior_config_init(aco_prior_config* cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->seed = 12345;
    cfg->lr = 2e-3f;
    cfg->beta1 = 0.9f;
    cfg->beta2 = 0.999f;
    cfg->eps = 1e-8f;
    cfg->pos_weight = 1.0f;
    cfg->gamma = 3.0f;
    cfg->transform = 1;
    cfg->minibatch = 65536;
}

extern "C" const char* aco_prior_last_error(const aco_prior* pr) { return pr ? pr->err : g_prior_err; }

// splitmix64: deterministic weight initialisation on the host
static uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static void init_params(uint64_t seed, float* w) {
    uint64_t s = seed;
    auto uni = [&](float a) { return (float)(((double)(splitmix64(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * a); };
    memset(w, 0, sizeof(float) * ACO_PRIOR_PARAMS);
    const float a1 = sqrtf(6.0f / (8 + 64)), a2 = sqrtf(6.0f / (64 + 64)), a3 = sqrtf(6.0f / (64 + 1));
    for (int i = 0; i < 512; ++i) w[aco::MLP_W1 + i] = uni(a1);
    for (int i = 0; i < 4096; ++i) w[aco::MLP_W2 + i] = uni(a2);
    for (int i = 0; i < 64; ++i) w[aco::MLP_W3 + i] = uni(a3);
}

static int upload_params(aco_prior* pr, const float* w) {
    PHIP(pr, hipMemcpyAsync(pr->d_master, w, sizeof(float) * ACO_PRIOR_PARAMS, hipMemcpyHostToDevice, pr->stream));
    PHIP(pr, hipMemsetAsync(pr->d_m, 0, sizeof(float) * ACO_PRIOR_PARAMS, pr->stream));
    PHIP(pr, hipMemsetAsync(pr->d_v, 0, sizeof(float) * ACO_PRIOR_PARAMS, pr->stream));
    PHIP(pr, hipMemsetAsync(pr->d_grad, 0, sizeof(float) * ACO_PRIOR_PARAMS, pr->stream));
    PHIP(pr, aco::launch_pack_bf16(pr->params(), pr->stream));
    PHIP(pr, hipStreamSynchronize(pr->stream));
    pr->step = 0;
    return ACO_OK;
}

extern "C" int aco_prior_destroy(aco_prior* pr) {
    if (!pr) return ACO_OK;
    if (pr->stream) (void)hipStreamSynchronize(pr->stream);
    pfree(pr->d_master);
    pfree(pr->d_w1);
    pfree(pr->d_w2);
    pfree(pr->d_m);
    pfree(pr->d_v);
    pfree(pr->d_grad);
    pfree(pr->d_loss);
    pfree(pr->d_partial);
    pfree(pr->d_sum);
    pfree(pr->d_x);
    pfree(pr->d_y);
    pfree(pr->d_feat);
    pfree(pr->d_logit);
    pfree(pr->d_pos);
    if (pr->stream) (void)hipStreamDestroy(pr->stream);
    delete pr;
    return ACO_OK;
}

extern "C" int aco_prior_create(const aco_prior_config* cfg, aco_prior** out) {
    if (!out) return pfail(nullptr, ACO_E_ARG, "out is NULL");
    *out = nullptr;
    aco_prior_config c;
    aco_prior_config_init(&c);
    if (cfg) c = *cfg;
    if (!(c.lr > 0.0f) || !(c.beta1 >= 0.0f && c.beta1 < 1.0f) || !(c.beta2 >= 0.0f && c.beta2 < 1.0f) ||
        !(c.eps > 0.0f) || !(c.pos_weight > 0.0f) || !(c.gamma > 0.0f) || c.minibatch < 16 ||
        (c.transform != 0 && c.transform != 1))
        return pfail(nullptr, ACO_E_ARG, "invalid prior configuration");
    aco_prior* pr = new aco_prior();
    pr->cfg = c;
    int rc = ACO_OK;
    hipError_t e = hipStreamCreateWithFlags(&pr->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        pfail(nullptr, ACO_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        delete pr;
        return ACO_E_HIP;
    }
    if ((rc = palloc(pr, &pr->d_master, ACO_PRIOR_PARAMS)) || (rc = palloc(pr, &pr->d_w1, aco::MLP_W1_FRAGS)) ||
        (rc = palloc(pr, &pr->d_w2, aco::MLP_W2_FRAGS)) || (rc = palloc(pr, &pr->d_m, ACO_PRIOR_PARAMS)) ||
        (rc = palloc(pr, &pr->d_v, ACO_PRIOR_PARAMS)) || (rc = palloc(pr, &pr->d_grad, ACO_PRIOR_PARAMS)) ||
        (rc = palloc(pr, &pr->d_loss, 1)) || (rc = palloc(pr, &pr->d_sum, 1)) ||
        (rc = palloc(pr, &pr->d_partial, (size_t)aco::MLP_MAX_ROWS * aco::MLP_ROW))) {
        snprintf(g_prior_err, sizeof(g_prior_err), "%s", pr->err);
        aco_prior_destroy(pr);
        return rc;
    }
    float w[ACO_PRIOR_PARAMS];
    init_params(c.seed, w);
    if ((rc = upload_params(pr, w))) {
        snprintf(g_prior_err, sizeof(g_prior_err), "%s", pr->err);
        aco_prior_destroy(pr);
        return rc;
    }
    *out = pr;
    return ACO_OK;
}

extern "C" int aco_prior_get_params(aco_prior* pr, float* params) {
    if (!pr || !params) return ACO_E_ARG;
    PHIP(pr, hipMemcpyAsync(params, pr->d_master, sizeof(float) * ACO_PRIOR_PARAMS, hipMemcpyDeviceToHost,
                            pr->stream));
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_set_params(aco_prior* pr, const float* params) {
    if (!pr || !params) return ACO_E_ARG;
    return upload_params(pr, params);
}

extern "C" int64_t aco_prior_size(const aco_prior* pr) { return pr ? pr->count : 0; }

extern "C" int aco_prior_clear(aco_prior* pr) {
    if (!pr) return ACO_E_ARG;
    pr->count = 0;
    return ACO_OK;
}

static int ensure_scratch(aco_prior* pr, int64_t edges, int64_t cities) {
    int rc;
    if (edges > pr->scratch_edges) {
        PHIP(pr, hipStreamSynchronize(pr->stream));
        pfree(pr->d_feat);
        pfree(pr->d_logit);
        if ((rc = palloc(pr, &pr->d_feat, (size_t)edges * 8)) || (rc = palloc(pr, &pr->d_logit, (size_t)edges)))
            return rc;
        pr->scratch_edges = edges;
    }
    if (cities > pr->scratch_cities) {
        PHIP(pr, hipStreamSynchronize(pr->stream));
        pfree(pr->d_pos);
        if ((rc = palloc(pr, &pr->d_pos, (size_t)cities))) return rc;
        pr->scratch_cities = cities;
    }
    return ACO_OK;
}

static int ensure_capacity(aco_prior* pr, int64_t need) {
    if (need <= pr->cap) return ACO_OK;
    int64_t cap = pr->cap ? pr->cap : (int64_t)1 << 20;
    while (cap < need) cap *= 2;
    float *x = nullptr, *y = nullptr;
    int rc;
    if ((rc = palloc(pr, &x, (size_t)cap * 8)) || (rc = palloc(pr, &y, (size_t)cap))) {
        pfree(x);
        return rc;
    }
    if (pr->count) {
        PHIP(pr, hipMemcpyAsync(x, pr->d_x, sizeof(float) * 8 * pr->count, hipMemcpyDeviceToDevice, pr->stream));
        PHIP(pr, hipMemcpyAsync(y, pr->d_y, sizeof(float) * pr->count, hipMemcpyDeviceToDevice, pr->stream));
    }
    PHIP(pr, hipStreamSynchronize(pr->stream));
    pfree(pr->d_x);
    pfree(pr->d_y);
    pr->d_x = x;
    pr->d_y = y;
    pr->cap = cap;
    return ACO_OK;
}

// Features + labels of one instance appended to the training set.
static int collect_view(aco_prior* pr, const InstanceView& v) {
    const int64_t edges = (int64_t)v.in.n * v.in.k;
    int rc;
    if ((rc = ensure_capacity(pr, pr->count + edges)) || (rc = ensure_scratch(pr, 0, v.in.n))) return rc;
    PHIP(pr, hipStreamSynchronize(v.stream));
    PHIP(pr, aco::launch_features(v.in, pr->d_sum, pr->d_x + pr->count * 8, pr->stream));
    PHIP(pr, aco::launch_labels(v.in, v.best_tour, pr->d_pos, pr->d_y + pr->count, pr->stream));
    pr->count += edges;
    return ACO_OK;
}

extern "C" int aco_prior_collect(aco_prior* pr, aco_ctx* c) {
    if (!pr) return ACO_E_ARG;
    InstanceView v;
    int rc = aco::colony_view(c, &v, pr->err, sizeof(pr->err));
    if (rc) return rc;
    if ((rc = collect_view(pr, v))) return rc;
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_collect_batch(aco_prior* pr, aco_batch* b) {
    if (!pr) return ACO_E_ARG;
    const int count = aco::batch_count(b);
    if (count <= 0) return pfail(pr, ACO_E_ARG, "empty or NULL batch");
    for (int i = 0; i < count; ++i) {
        InstanceView v;
        int rc = aco::batch_view(b, i, &v, pr->err, sizeof(pr->err));
        if (rc || (rc = collect_view(pr, v))) return rc;
    }
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

// One Adam step on the gradient accumulated in d_grad.
static int adam_step(aco_prior* pr) {
    pr->step += 1;
    const double bc1 = 1.0 - pow((double)pr->cfg.beta1, (double)pr->step);
    const double bc2 = 1.0 - pow((double)pr->cfg.beta2, (double)pr->step);
    PHIP(pr, aco::launch_adam(pr->params(), pr->d_grad, pr->d_m, pr->d_v, pr->cfg.lr, pr->cfg.beta1, pr->cfg.beta2,
                              pr->cfg.eps, (float)bc1, (float)bc2, pr->stream));
    PHIP(pr, aco::launch_pack_bf16(pr->params(), pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_train(aco_prior* pr, int32_t epochs, float* loss) {
    if (!pr || epochs < 0) return ACO_E_ARG;
    if (pr->count == 0) return pfail(pr, ACO_E_STATE, "empty training set");
    const int64_t mb = pr->cfg.minibatch;
    for (int32_t ep = 0; ep < epochs; ++ep) {
        PHIP(pr, hipMemsetAsync(pr->d_loss, 0, sizeof(float), pr->stream));
        int64_t steps = 0;
        for (int64_t off = 0; off < pr->count; off += mb, ++steps) {
            const int64_t cnt = pr->count - off < mb ? pr->count - off : mb;
            PHIP(pr, aco::launch_mlp_backward(pr->params(), pr->d_x + off * 8, pr->d_y + off, cnt,
                                              pr->cfg.pos_weight, pr->d_partial, pr->d_grad, pr->d_loss,
                                              pr->stream));
            int rc = adam_step(pr);
            if (rc) return rc;
        }
        float l = 0.0f;
        PHIP(pr, hipMemcpyAsync(&l, pr->d_loss, sizeof(float), hipMemcpyDeviceToHost, pr->stream));
        PHIP(pr, hipStreamSynchronize(pr->stream));
        if (loss) loss[ep] = l / (float)steps;
    }
    return ACO_OK;
}

static int apply_view(aco_prior* pr, const InstanceView& v) {
    const int64_t edges = (int64_t)v.in.n * v.in.k;
    int rc = ensure_scratch(pr, edges, 0);
    if (rc) return rc;
    PHIP(pr, hipStreamSynchronize(v.stream));
    PHIP(pr, aco::launch_features(v.in, pr->d_sum, pr->d_feat, pr->stream));
    PHIP(pr, aco::launch_mlp_forward(pr->params(), pr->d_feat, edges, pr->d_logit, pr->stream));
    PHIP(pr, aco::launch_logit_to_heuristic(pr->d_logit, edges, pr->cfg.gamma, pr->cfg.transform, v.eta_nn,
                                                pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_apply(aco_prior* pr, aco_ctx* c) {
    if (!pr) return ACO_E_ARG;
    InstanceView v;
    int rc = aco::colony_view(c, &v, pr->err, sizeof(pr->err));
    if (rc || (rc = apply_view(pr, v))) return rc;
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free apply.  Its features and scores live in `scratch`, a
// file-scope device array of fixed size (NOALLOC_MAX_EDGES candidate edges:
// 256 bytes for the nearest-neighbour distance sum, then 8 features and one
// score per edge, 9.0 MiB in all), which every GPU the library runs on holds
// until the process ends.  Calls that overlap on one GPU share it and race.
// k_nn0_sum writes the sum (nothing needs clearing) and the features and
// scores are written before they are read, so the previous call's contents
// never matter.
constexpr int64_t NOALLOC_MAX_EDGES = int64_t(1) << 18;
constexpr size_t NOALLOC_HEAD = 64;  // floats: the 64-bit sum, padded to 256 bytes
static __device__ __align__(256) float scratch[NOALLOC_HEAD + 9 * NOALLOC_MAX_EDGES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (the array's name is only a host-side stand-in).
static int noalloc_scratch(aco_prior* pr, float** out) {
    int dev = 0;
    PHIP(pr, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return pfail(pr, ACO_E_RANGE, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            return pfail(pr, ACO_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return ACO_OK;
}

extern "C" int aco_prior_apply_noalloc(aco_prior* pr, aco_ctx* c) {
    if (!pr) return ACO_E_ARG;
    InstanceView v;
    int rc = aco::colony_view(c, &v, pr->err, sizeof(pr->err));
    if (rc) return rc;
    const int64_t edges = (int64_t)v.in.n * v.in.k;
    if (edges > NOALLOC_MAX_EDGES)
        return pfail(pr, ACO_E_RANGE,
                     "n*k = %lld candidate edges exceeds the allocation-free limit of %lld (use aco_prior_apply)",
                     (long long)edges, (long long)NOALLOC_MAX_EDGES);
    float* base = nullptr;
    if ((rc = noalloc_scratch(pr, &base))) return rc;
    unsigned long long* sum = reinterpret_cast<unsigned long long*>(base);
    float* feat = base + NOALLOC_HEAD;
    float* logit = feat + 8 * NOALLOC_MAX_EDGES;
    PHIP(pr, hipStreamSynchronize(v.stream));
    PHIP(pr, aco::launch_features(v.in, sum, feat, pr->stream));
    PHIP(pr, aco::launch_mlp_forward(pr->params(), feat, edges, logit, pr->stream));
    PHIP(pr, aco::launch_logit_to_heuristic(logit, edges, pr->cfg.gamma, pr->cfg.transform, v.eta_nn, pr->stream));
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_apply_batch(aco_prior* pr, aco_batch* b) {
    if (!pr) return ACO_E_ARG;
    const int count = aco::batch_count(b);
    if (count <= 0) return pfail(pr, ACO_E_ARG, "empty or NULL batch");
    for (int i = 0; i < count; ++i) {
        InstanceView v;
        int rc = aco::batch_view(b, i, &v, pr->err, sizeof(pr->err));
        if (rc || (rc = apply_view(pr, v))) return rc;
    }
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_features(aco_prior* pr, aco_ctx* c, float* features, float* labels) {
    if (!pr || (!features && !labels)) return ACO_E_ARG;
    InstanceView v;
    int rc = aco::colony_view(c, &v, pr->err, sizeof(pr->err));
    if (rc) return rc;
    const int64_t edges = (int64_t)v.in.n * v.in.k;
    if ((rc = ensure_scratch(pr, edges, v.in.n))) return rc;
    PHIP(pr, hipStreamSynchronize(v.stream));
    PHIP(pr, aco::launch_features(v.in, pr->d_sum, pr->d_feat, pr->stream));
    PHIP(pr, aco::launch_labels(v.in, v.best_tour, pr->d_pos, pr->d_logit, pr->stream));
    if (features)
        PHIP(pr, hipMemcpyAsync(features, pr->d_feat, sizeof(float) * 8 * edges, hipMemcpyDeviceToHost, pr->stream));
    if (labels)
        PHIP(pr, hipMemcpyAsync(labels, pr->d_logit, sizeof(float) * edges, hipMemcpyDeviceToHost, pr->stream));
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

// Host feature rows into device scratch (x at d_feat, y at d_logit).
static int upload_rows(aco_prior* pr, const float* x, const float* y, int64_t count) {
    int rc = ensure_scratch(pr, count, 0);
    if (rc) return rc;
    PHIP(pr, hipMemcpyAsync(pr->d_feat, x, sizeof(float) * 8 * count, hipMemcpyHostToDevice, pr->stream));
    if (y) PHIP(pr, hipMemcpyAsync(pr->d_logit, y, sizeof(float) * count, hipMemcpyHostToDevice, pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_forward(aco_prior* pr, const float* x, int64_t count, float* scores) {
    if (!pr || !x || !scores || count <= 0) return pr ? pfail(pr, ACO_E_ARG, "bad forward arguments") : ACO_E_ARG;
    int rc = upload_rows(pr, x, nullptr, count);
    if (rc) return rc;
    PHIP(pr, aco::launch_mlp_forward(pr->params(), pr->d_feat, count, pr->d_logit, pr->stream));
    PHIP(pr, hipMemcpyAsync(scores, pr->d_logit, sizeof(float) * count, hipMemcpyDeviceToHost, pr->stream));
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_gradient(aco_prior* pr, const float* x, const float* y, int64_t count, float* grad,
                                  float* loss) {
    if (!pr || !x || !y || !grad || count <= 0)
        return pr ? pfail(pr, ACO_E_ARG, "bad gradient arguments") : ACO_E_ARG;
    int rc = upload_rows(pr, x, y, count);
    if (rc) return rc;
    // y lives in d_logit; the gradient in a separate buffer so that a pending
    // training gradient is not disturbed (d_grad is zero between steps).
    float* g = nullptr;
    if ((rc = palloc(pr, &g, ACO_PRIOR_PARAMS + 1))) return rc;
    hipError_t e = hipMemsetAsync(g, 0, sizeof(float) * (ACO_PRIOR_PARAMS + 1), pr->stream);
    if (e == hipSuccess)
        e = aco::launch_mlp_backward(pr->params(), pr->d_feat, pr->d_logit, count, pr->cfg.pos_weight,
                                     pr->d_partial, g, g + ACO_PRIOR_PARAMS, pr->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(grad, g, sizeof(float) * ACO_PRIOR_PARAMS, hipMemcpyDeviceToHost, pr->stream);
    float l = 0.0f;
    if (e == hipSuccess) e = hipMemcpyAsync(&l, g + ACO_PRIOR_PARAMS, sizeof(float), hipMemcpyDeviceToHost, pr->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(pr->stream);
    (void)hipFree(g);
    if (e != hipSuccess) return pfail(pr, ACO_E_HIP, "gradient failed: %s", hipGetErrorString(e));
    if (loss) *loss = l;
    return ACO_OK;
}

extern "C" int aco_prior_adam_step(aco_prior* pr, const float* grad) {
    if (!pr || !grad) return ACO_E_ARG;
    PHIP(pr, hipMemcpyAsync(pr->d_grad, grad, sizeof(float) * ACO_PRIOR_PARAMS, hipMemcpyHostToDevice, pr->stream));
    int rc = adam_step(pr);
    if (rc) return rc;
    PHIP(pr, hipStreamSynchronize(pr->stream));
    return ACO_OK;
}

extern "C" int aco_prior_profile(aco_prior* pr, int32_t reps, float* ms) {
    if (!pr || !ms || reps <= 0) return ACO_E_ARG;
    if (pr->count == 0) return pfail(pr, ACO_E_STATE, "empty training set");
    const int64_t cnt = pr->count < pr->cfg.minibatch ? pr->count : pr->cfg.minibatch;
    int rc = ensure_scratch(pr, cnt, 0);
    if (rc) return rc;
    float* g = nullptr;
    if ((rc = palloc(pr, &g, ACO_PRIOR_PARAMS + 1))) return rc;
    hipEvent_t ev[3] = {};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], pr->stream);
    for (int32_t r = 0; r < reps && e == hipSuccess; ++r)
        e = aco::launch_mlp_forward(pr->params(), pr->d_x, cnt, pr->d_logit, pr->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[1], pr->stream);
    for (int32_t r = 0; r < reps && e == hipSuccess; ++r)
        e = aco::launch_mlp_backward(pr->params(), pr->d_x, pr->d_y, cnt, pr->cfg.pos_weight, pr->d_partial, g,
                                     g + ACO_PRIOR_PARAMS, pr->stream);
    if (e == hipSuccess) e = hipEventRecord(ev[2], pr->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
    float a = 0.0f, b = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&a, ev[0], ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[1], ev[2]);
    for (int i = 0; i < 3; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    (void)hipFree(g);
    if (e != hipSuccess) return pfail(pr, ACO_E_HIP, "profile failed: %s", hipGetErrorString(e));
    ms[0] = a / (float)reps;
    ms[1] = b / (float)reps;
    ms[2] = (float)cnt;
    return ACO_OK;
}
