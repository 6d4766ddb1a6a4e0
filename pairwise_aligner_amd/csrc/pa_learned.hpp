// pa_learned.hpp -- launch interface of the learned-dispatch kernels (pa_learned.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_device.h"

namespace pa {

inline constexpr int kModelFeatures = PA_FEATURES;
inline constexpr int kModelHidden = PA_MODEL_HIDDEN;
inline constexpr int kModelOutputs = PA_MODEL_OUTPUTS;
inline constexpr int kModelParams = kModelFeatures * kModelHidden + kModelHidden + kModelHidden * kModelHidden +
                                    kModelHidden + kModelHidden * kModelOutputs + kModelOutputs;

struct FeatureArgs {
    const uint8_t *seq1;
    const uint64_t *off1;       // n_pairs + 1 offsets, or with one_query the query's two
    int32_t one_query;          // every pair's first sequence is seq1[off1[0], off1[1])
    const uint8_t *seq2;
    const uint64_t *off2;
    uint64_t n_pairs;
    float *out;          // n_pairs x kModelFeatures
    int32_t k;           // k-mer length
    int32_t bits;        // bits per residue code in a k-mer
    float smax;          // largest substitution score
    uint8_t code[256];   // residue -> code
};

hipError_t launch_features(const FeatureArgs &a, hipStream_t st);
hipError_t launch_forward(const float *params, const float *x, uint64_t n, float *y, hipStream_t st);
hipError_t launch_loss_grad(const float *params, const float *x, const float *target, const float *weight,
                            uint64_t n, float *grad, float *loss, hipStream_t st);
hipError_t launch_adam(float *params, const float *grad, float *m1, float *m2, float lr, int step, hipStream_t st);

} // namespace pa
