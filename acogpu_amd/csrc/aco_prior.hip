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
typedef __bf16 bf16x8h __attribute__((ext_vector_type(8)));

// Four fp32 values rounded to bf16 (nearest even) by gfx950's
// v_cvt_pk_bf16_f32, two per instruction.
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x2h lo, hi;
    lo[0] = (__bf16)a;
    lo[1] = (__bf16)b;
    hi[0] = (__bf16)c;
    hi[1] = (__bf16)d;
    const uint32_t l = __builtin_bit_cast(uint32_t, lo), h = __builtin_bit_cast(uint32_t, hi);
    bf16x4 v;
    v[0] = (short)(l & 0xffffu);
    v[1] = (short)(l >> 16);
    v[2] = (short)(h & 0xffffu);
    v[3] = (short)(h >> 16);
    return v;
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
// so Phi = 1 - Q for z >= 0 and Q for z < 0, Q = P(t) e^{-z^2/2} / 2.  The
// derivative is that of this function: Phi + z Phi', Phi' = e^{-z^2/2}
// (c t^2 P'(t) + |z| P(t)) / 2 with c = p / sqrt(2).  One hardware
// exponential (v_exp_f32) and one reciprocal (v_rcp_f32) per value.
constexpr float GELU_A1 = 0.254829592f, GELU_A2 = -0.284496736f, GELU_A3 = 1.421413741f;
constexpr float GELU_A4 = -1.453152027f, GELU_A5 = 1.061405429f;
constexpr float GELU_C = 0.231641888f;       // 0.3275911 / sqrt(2)
constexpr float GELU_HLOG2E = 0.72134752f;   // log2(e) / 2

template <bool GRAD>
__device__ __forceinline__ float gelu_eval(float z, float& dg) {
    const float a = fabsf(z);
    const float e = __builtin_amdgcn_exp2f(-GELU_HLOG2E * (z * z));  // e^{-z^2/2}
    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(GELU_C, a, 1.0f));
    const float poly = __builtin_fmaf(
        t, __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, GELU_A5, GELU_A4), GELU_A3), GELU_A2), GELU_A1);
    const float P = t * poly;
    const float q = 0.5f * P * e;
    const float phi = z >= 0.0f ? 1.0f - q : q;
    if (GRAD) {
        const float dP = __builtin_fmaf(
            t, __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, 5.0f * GELU_A5, 4.0f * GELU_A4), 3.0f * GELU_A3),
                              2.0f * GELU_A2),
            GELU_A1);
        const float dphi = 0.5f * e * __builtin_fmaf(GELU_C * t * t, dP, a * P);
        dg = __builtin_fmaf(z, dphi, phi);
    }
    return z * phi;
}

__device__ __forceinline__ float gelu(float z) {
    float unused = 0.0f;
    return gelu_eval<false>(z, unused);
}

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

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

// Layer-2 weights as MFMA A operands, staged in LDS once per block (16
// bytes per lane per fragment, read with one ds_read_b128):
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

__device__ __forceinline__ uint32_t bf_pair(const uint16_t* w, int i0, int i1) {
    return (uint32_t)w[i0] | ((uint32_t)w[i1] << 16);
}

__device__ __forceinline__ void stage_w2(const MlpParams& p, W2Frags* F, bool backward) {
    const uint16_t* w = p.w2_bf16;  // 64 x 64, row = output feature
    for (int f = threadIdx.x; f < 4 * 2 * 64; f += blockDim.x) {
        const int blk = f >> 7, c = (f >> 6) & 1, l = f & 63, r = l & 15, g = l >> 4;
        const int k0 = 16 * (2 * c) + 4 * g, k1 = 16 * (2 * c + 1) + 4 * g;
        const int o = 16 * blk + r;
        F->fwd[f] = make_uint4(bf_pair(w, o * 64 + k0, o * 64 + k0 + 1), bf_pair(w, o * 64 + k0 + 2, o * 64 + k0 + 3),
                               bf_pair(w, o * 64 + k1, o * 64 + k1 + 1), bf_pair(w, o * 64 + k1 + 2, o * 64 + k1 + 3));
        if (backward) {
            const int i = 16 * blk + r;
            F->bwd[f] = make_uint4(bf_pair(w, k0 * 64 + i, (k0 + 1) * 64 + i),
                                   bf_pair(w, (k0 + 2) * 64 + i, (k0 + 3) * 64 + i),
                                   bf_pair(w, k1 * 64 + i, (k1 + 1) * 64 + i),
                                   bf_pair(w, (k1 + 2) * 64 + i, (k1 + 3) * 64 + i));
        }
    }
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
    const int row = lane & 15, grp = lane >> 4;
    for (int ob = 0; ob < 4; ++ob) {
        const uint16_t* w1 = p.w1_bf16 + (size_t)(16 * ob + row) * 16 + 4 * grp;
        R.w1a[ob] = bf16x4{(short)w1[0], (short)w1[1], (short)w1[2], (short)w1[3]};
    }
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
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    for (int ob = 0; ob < 4; ++ob) {
        f32x4 z = mfma16(R.w1a[ob], xb, zero);
        const float4 b = vec4(vec, VEC_B1, ob, grp);
        z[0] = z[0] + b.x;
        z[1] = z[1] + b.y;
        z[2] = z[2] + b.z;
        z[3] = z[3] + b.w;
        A.h1[ob] = gelu4<GRAD>(z, A.d1[ob]);
    }
    float partial = 0.0f;
    for (int ob = 0; ob < 4; ++ob) {
        f32x4 z = zero;
        for (int c = 0; c < 2; ++c) {
            bf16x4 a0, a1;
            frag_halves(F->fwd[(ob * 2 + c) * 64 + lane], a0, a1);
            z = mfma32(a0, a1, A.h1[2 * c], A.h1[2 * c + 1], z);
        }
        const float4 b = vec4(vec, VEC_B2, ob, grp);
        z[0] = z[0] + b.x;
        z[1] = z[1] + b.y;
        z[2] = z[2] + b.z;
        z[3] = z[3] + b.w;
        A.h2[ob] = gelu4<GRAD>(z, A.d2[ob]);
        const float4 w = vec4(vec, VEC_W3, ob, grp);
        partial = partial + w.x * bf2f((uint16_t)A.h2[ob][0]);
        partial = partial + w.y * bf2f((uint16_t)A.h2[ob][1]);
        partial = partial + w.z * bf2f((uint16_t)A.h2[ob][2]);
        partial = partial + w.w * bf2f((uint16_t)A.h2[ob][3]);
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

// Backward: binary cross-entropy with logits, positive weight `pos_weight`,
// mean over `count` edges.  Every wave accumulates its tiles' gradients in
// fp32 registers; the block's waves then add their rows in LDS in wave order
// and the block stores one partial row (parameters, then the loss).
__global__ __launch_bounds__(256) void k_mlp_backward(const MlpParams p, const float* __restrict__ x,
                                                      const float* __restrict__ y, int64_t count, float pos_weight,
                                                      float* __restrict__ partial) {
    __shared__ __align__(16) uint16_t tr[4][8 * 256];
    __shared__ __align__(16) float vec[VEC_LEN];
    __shared__ __align__(16) W2Frags F;
    __shared__ __align__(16) float acc[MLP_STRIDE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, grp = lane >> 4, col = lane & 15;
    stage_w2(p, &F, true);
    stage_vectors(p, vec);
    MlpRegs R;
    load_mlp(p, R, lane);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gW2[4][4], gW1[4];
    for (int a = 0; a < 4; ++a) {
        gW1[a] = zero;
        for (int b = 0; b < 4; ++b) gW2[a][b] = zero;
    }
    float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.0f, lsum = 0.0f;
    for (int a = 0; a < 4; ++a)
        for (int q = 0; q < 4; ++q) gb1[a][q] = gb2[a][q] = gw3[a][q] = 0.0f;
    const float inv = 1.0f / (float)count;
    uint16_t* sc = tr[wave];
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave; t * 16 < count; t += waves) {
        const int64_t e0 = t * 16;
        const bool valid = e0 + col < count;
        const bf16x4 xb = load_x(x, e0, count, lane);
        const float yy = valid ? y[e0 + col] : 0.0f;
        Activations A;
        forward_tile<true>(R, &F, vec, xb, A, lane);
        const float s = sigmoidf(A.out);
        // d loss / d logit for -w y log s - (1-y) log(1-s)
        const float g = valid ? (s * (1.0f - yy + pos_weight * yy) - pos_weight * yy) * inv : 0.0f;
        if (grp == 0 && valid) {
            const float sp_pos = fmaxf(-A.out, 0.0f) + log1pf(expf(-fabsf(A.out)));  // -log s
            const float sp_neg = fmaxf(A.out, 0.0f) + log1pf(expf(-fabsf(A.out)));   // -log(1-s)
            lsum += (pos_weight * yy * sp_pos + (1.0f - yy) * sp_neg) * inv;
        }
        if (grp == 0) gb3 += g;
        // layer 3 and dz2 (D layout: features 16ob + 4grp + q of edge col)
        bf16x4 dz2b[4];
        for (int ob = 0; ob < 4; ++ob) {
            float dz[4];
            const float4 w = vec4(vec, VEC_W3, ob, grp);
            const float w3[4] = {w.x, w.y, w.z, w.w};
            for (int q = 0; q < 4; ++q) {
                gw3[ob][q] += g * bf2f((uint16_t)A.h2[ob][q]);
                dz[q] = g * w3[q] * A.d2[ob][q];
                gb2[ob][q] += dz[q];
            }
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
                frag_halves(F.bwd[(ib * 2 + c) * 64 + lane], a0, a1);
                dh = mfma32(a0, a1, dz2b[2 * c], dz2b[2 * c + 1], dh);
            }
            float dz[4];
            for (int q = 0; q < 4; ++q) {
                dz[q] = dh[q] * A.d1[ib][q];
                gb1[ib][q] += dz[q];
            }
            t2in[ib] = pack4(dz[0], dz[1], dz[2], dz[3]);
        }
        // dW1[o][i] = sum_e dz1[o][e] x[i][e]: B = x^T (e x i), i in 0..15
        t2in[4] = xb;
        transpose16_batch<5>(t2in, t2out, sc, lane);
        for (int ob = 0; ob < 4; ++ob) gW1[ob] = mfma16(t2out[ob], t2out[4], gW1[ob]);
    }
    // bias / w3 gradients: reduce over the 16 edge lanes of each group
    for (int ob = 0; ob < 4; ++ob)
        for (int q = 0; q < 4; ++q)
            for (int off = 8; off >= 1; off >>= 1) {
                gb1[ob][q] += __shfl_xor(gb1[ob][q], off);
                gb2[ob][q] += __shfl_xor(gb2[ob][q], off);
                gw3[ob][q] += __shfl_xor(gw3[ob][q], off);
            }
    for (int off = 32; off >= 1; off >>= 1) {
        gb3 += __shfl_xor(gb3, off);
        lsum += __shfl_xor(lsum, off);
    }
    // the block's row: waves add theirs in LDS in wave order (every entry of
    // a row has exactly one writing lane), then one coalesced store
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) {
        if (wave == wv) {
            const bool first = wv == 0;
            auto put = [&](int idx, float v) { acc[idx] = first ? v : acc[idx] + v; };
            // D layout of a gradient tile: row 4grp+q, column col
            for (int ob = 0; ob < 4; ++ob)
                for (int q = 0; q < 4; ++q) {
                    const int o = 16 * ob + 4 * grp + q;
                    if (col < 8) put(MLP_W1 + o * 8 + col, gW1[ob][q]);
                    for (int ib = 0; ib < 4; ++ib) put(MLP_W2 + o * 64 + 16 * ib + col, gW2[ob][ib][q]);
                    if (col == 0) {
                        put(MLP_B1 + o, gb1[ob][q]);
                        put(MLP_B2 + o, gb2[ob][q]);
                        put(MLP_W3 + o, gw3[ob][q]);
                    }
                }
            if (lane == 0) {
                put(MLP_B3, gb3);
                put(MLP_PARAMS, lsum);
            }
        }
        __syncthreads();
    }
    float* row = partial + (size_t)blockIdx.x * MLP_STRIDE;
    for (int i = threadIdx.x; i < MLP_STRIDE; i += blockDim.x) row[i] = acc[i];
}

// Sum of the per-wave partial rows in a fixed order, in two stages.  Stage 1:
// block (x, y) sums rows 64 y + 16 g .. +15 (g = thread / 64) of 64 columns
// into row 4 y + g of `stage`.  Stage 2: grad = sum of the stage rows in
// order, loss += its last column.
__global__ __launch_bounds__(256) void k_reduce_grad1(const float* __restrict__ partial, int32_t rows,
                                                      float* __restrict__ stage) {
    const int32_t col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int32_t g = threadIdx.x >> 6;
    if (col >= MLP_STRIDE) return;
    const int32_t r0 = blockIdx.y * 64 + g * 16;
    float acc = 0.0f;
    for (int32_t i = 0; i < 16; ++i)
        if (r0 + i < rows) acc += partial[(size_t)(r0 + i) * MLP_STRIDE + col];
    stage[(size_t)(blockIdx.y * 4 + g) * MLP_STRIDE + col] = acc;
}

__global__ __launch_bounds__(256) void k_reduce_grad2(const float* __restrict__ stage, int32_t rows,
                                                      float* __restrict__ grad, float* __restrict__ loss) {
    const int32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= MLP_STRIDE) return;
    float acc = 0.0f;
    for (int32_t r = 0; r < rows; ++r) acc += stage[(size_t)r * MLP_STRIDE + j];
    if (j < MLP_PARAMS)
        grad[j] = acc;
    else
        *loss += acc;
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

__global__ __launch_bounds__(256) void k_pack_bf16(MlpParams p) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < 64 * 16) {
        const int o = i / 16, c = i % 16;
        p.w1_bf16[i] = c < 8 ? f2bf(p.master[MLP_W1 + o * 8 + c]) : (uint16_t)0;
    }
    if (i < 64 * 64) p.w2_bf16[i] = f2bf(p.master[MLP_W2 + i]);
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
    return (unsigned)((waves + 3) / 4 > 0 ? (waves + 3) / 4 : 1);
}

hipError_t launch_mlp_backward(const MlpParams& p, const float* x, const float* y, int64_t count, float pos_weight,
                               float* partial, float* grad, float* loss, hipStream_t s) {
    const unsigned blocks = backward_blocks(count);  // one partial row per block
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(256), 0, s, p, x, y, count, pos_weight, partial);
    const int32_t rows = (int32_t)blocks;
    const int32_t chunks = (rows + 63) / 64;
    float* stage = partial + (size_t)MLP_MAX_ROWS * MLP_STRIDE;
    hipLaunchKernelGGL(k_reduce_grad1, dim3(cdiv(MLP_STRIDE, 64), chunks), dim3(256), 0, s, partial, rows, stage);
    hipLaunchKernelGGL(k_reduce_grad2, dim3(cdiv(MLP_STRIDE, 256)), dim3(256), 0, s, stage, chunks * 4, grad, loss);
    return hipGetLastError();
}

hipError_t launch_adam(const MlpParams& p, float* grad, float* m, float* v, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(cdiv(MLP_PARAMS, 256)), dim3(256), 0, s, p, grad, m, v, lr, beta1, beta2, eps,
                       bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_pack_bf16(const MlpParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_bf16, dim3(cdiv(64 * 64, 256)), dim3(256), 0, s, p);
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
    uint16_t* d_w1 = nullptr;
    uint16_t* d_w2 = nullptr;
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
    if ((rc = palloc(pr, &pr->d_master, ACO_PRIOR_PARAMS)) || (rc = palloc(pr, &pr->d_w1, 64 * 16)) ||
        (rc = palloc(pr, &pr->d_w2, 64 * 64)) || (rc = palloc(pr, &pr->d_m, ACO_PRIOR_PARAMS)) ||
        (rc = palloc(pr, &pr->d_v, ACO_PRIOR_PARAMS)) || (rc = palloc(pr, &pr->d_grad, ACO_PRIOR_PARAMS)) ||
        (rc = palloc(pr, &pr->d_loss, 1)) || (rc = palloc(pr, &pr->d_sum, 1)) ||
        (rc = palloc(pr, &pr->d_partial, (size_t)(aco::MLP_MAX_ROWS + aco::MLP_MAX_ROWS / 16) * aco::MLP_STRIDE))) {
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
