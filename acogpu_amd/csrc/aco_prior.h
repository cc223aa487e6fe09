// Internal interface of the learned edge prior (kernels in aco_prior.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/acogpu.h"

namespace aco {

// One instance's inputs on the device: pitched distance matrix and NN lists.
struct PriorInstance {
    const int32_t* dist;  // n x P
    const int32_t* nn;    // n x k
    int32_t n, P, k;
};

// MLP 8 -> 64 -> 64 -> 1.  fp32 master parameters at the offsets below; bf16
// copies of the two matrix layers, already in the per-lane order of the MFMA
// operands (W1 padded to 16 inputs), feed the matrix cores.
enum : int32_t {
    MLP_W1 = 0,      // 64 x 8, row = output feature
    MLP_B1 = 512,    // 64
    MLP_W2 = 576,    // 64 x 64, row = output feature
    MLP_B2 = 4672,   // 64
    MLP_W3 = 4736,   // 64
    MLP_B3 = 4800,   // 1
    MLP_PARAMS = 4801,
    // Backward partial-gradient row, in the per-lane order of the MFMA
    // accumulators (W2 then W1 with b1 in its input column 8; b2 and w3 in
    // the order of their lane groups), then b3 and the loss; padded to a
    // multiple of four floats.
    MLP_ROW_W2 = 0,         // 4 x 4 blocks (ob, ib) x 64 lanes x 4 (q)
    MLP_ROW_W1 = 4096,      // 4 blocks (ob) x 64 lanes x 4 (q)
    MLP_ROW_B2 = 5120,      // 4 groups x 4 blocks (ob) x 4 (q)
    MLP_ROW_W3 = 5184,      // same order
    MLP_ROW_B3 = 5248,
    MLP_ROW_LOSS = 5249,
    MLP_ROW = 5252,
    MLP_BWD_WAVES = 8,      // backward waves per block (two per SIMD)
    MLP_MAX_ROWS = 256,     // backward blocks per launch, one partial row each
    MLP_MAX_WAVES = MLP_MAX_ROWS * MLP_BWD_WAVES,
};
static_assert(MLP_PARAMS == ACO_PRIOR_PARAMS, "parameter count");

enum : int32_t {
    MLP_W1_FRAGS = 4 * 64,      // uint2 per lane and 16-row block of W1
    MLP_W2_FRAGS = 2 * 4 * 2 * 64,  // uint4: forward then backward fragments of W2
};

struct MlpParams {
    float* master;
    uint2* w1_frag;  // MLP_W1_FRAGS
    uint4* w2_frag;  // MLP_W2_FRAGS
};

// A colony's or batch instance's device arrays the prior reads and writes.
struct InstanceView {
    PriorInstance in;
    const int32_t* best_tour;  // n
    float* eta_nn;             // n x k heuristic on candidate edges
    hipStream_t stream;
};

int colony_view(aco_ctx* c, InstanceView* v, char* err, size_t errlen);
int batch_count(aco_batch* b);
int batch_view(aco_batch* b, int32_t i, InstanceView* v, char* err, size_t errlen);

hipError_t launch_features(const PriorInstance& in, unsigned long long* sum, float* feat, hipStream_t s);
hipError_t launch_labels(const PriorInstance& in, const int32_t* tour, int32_t* pos, float* y, hipStream_t s);
hipError_t launch_mlp_forward(const MlpParams& p, const float* x, int64_t count, float* out, hipStream_t s);
// grad (MLP_PARAMS) is overwritten, *loss accumulates; partial holds
// MLP_MAX_ROWS x MLP_ROW floats of scratch.
hipError_t launch_mlp_backward(const MlpParams& p, const float* x, const float* y, int64_t count, float pos_weight,
                               float* partial, float* grad, float* loss, hipStream_t s);
hipError_t launch_adam(const MlpParams& p, float* grad, float* m, float* v, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, hipStream_t s);
hipError_t launch_pack_bf16(const MlpParams& p, hipStream_t s);
hipError_t launch_logit_to_heuristic(const float* logit, int64_t count, float gamma, int32_t transform, float* h,
                                     hipStream_t s);

}  // namespace aco
