// Internal declarations of the tail-model kernels (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfcehip {

constexpr int kMlpIn = 32;
constexpr int kMlpH = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpOut * kMlpH + kMlpOut;
constexpr int kMlpBatch = 128;
constexpr int kMlpThreads = 256;

struct MlpTrainArgs {
  float* params;          // [kMlpParams], updated in place
  float* adam_m;          // [kMlpParams]
  float* adam_v;          // [kMlpParams]
  const float* X;         // [n][32]
  const float* Y;         // [n][2]
  const int32_t* batch_idx;  // [steps][batch] example indices
  int32_t n, steps, batch, t0;
  float lr, beta1, beta2, eps, weight_decay;
  float* loss;            // [steps] minibatch loss before the update, or null
  float* grad0;           // [kMlpParams] gradient of the first step, or null
};

hipError_t launch_mlp_forward(const float* params, const float* X, int n, float* Y, hipStream_t s);
hipError_t launch_mlp_train(const MlpTrainArgs& a, hipStream_t s);

}  // namespace tfcehip
