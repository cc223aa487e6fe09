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
// copies of the two matrix layers feed MFMA (W1 padded to 16 inputs).
enum : int32_t {
    MLP_W1 = 0,      // 64 x 8, row = output feature
    MLP_B1 = 512,    // 64
    MLP_W2 = 576,    // 64 x 64, row = output feature
    MLP_B2 = 4672,   // 64
    MLP_W3 = 4736,   // 64
    MLP_B3 = 4800,   // 1
    MLP_PARAMS = 4801,
    MLP_STRIDE = 4802,      // partial-gradient row: parameters + loss
    MLP_MAX_WAVES = 1024,   // backward waves per launch (four per block)
    MLP_MAX_ROWS = 256,     // backward partial rows per launch (one per block)
};
static_assert(MLP_PARAMS == ACO_PRIOR_PARAMS, "parameter count");

struct MlpParams {
    float* master;
    uint16_t* w1_bf16;  // 64 x 16
    uint16_t* w2_bf16;  // 64 x 64
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
// (MLP_MAX_ROWS + MLP_MAX_ROWS / 16) x MLP_STRIDE floats of scratch.
hipError_t launch_mlp_backward(const MlpParams& p, const float* x, const float* y, int64_t count, float pos_weight,
                               float* partial, float* grad, float* loss, hipStream_t s);
hipError_t launch_adam(const MlpParams& p, float* grad, float* m, float* v, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, hipStream_t s);
hipError_t launch_pack_bf16(const MlpParams& p, hipStream_t s);
hipError_t launch_logit_to_heuristic(const float* logit, int64_t count, float gamma, int32_t transform, float* h,
                                     hipStream_t s);

}  // namespace aco
