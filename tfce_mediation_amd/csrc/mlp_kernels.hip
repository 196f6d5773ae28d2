// Tail model of the learned early stop: a 32 -> 64 -> 64 -> 2 MLP (tanh
// hidden units, linear output, FP32) that maps 32 standardised quantiles of
// the first null values to the standardised 95% / 99% FWER critical values.
// There is no counterpart in the reference; its nearest analogues are the
// parametric null fits of bin/tm_maths:532-552.
//
// Parameter layout (float, kMlpParams): W1[64][32] b1[64] W2[64][64] b2[64]
// W3[2][64] b3[2].
//
// mlp_forward_kernel: one thread per example.
// mlp_train_kernel:   one workgroup runs `steps` Adam steps on minibatches of
//   up to kMlpBatch examples.  Parameters live in LDS; per-example activations
//   and deltas are staged in LDS, then every parameter gradient is a
//   deterministic reduction over the batch (no atomics), so a run is
//   reproducible bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_internal.h"

namespace tfcehip {

namespace {

constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;
static_assert(kOffB3 + kMlpOut == kMlpParams, "parameter layout");

__device__ __forceinline__ void forward_one(const float* P, const float* x, float* h1, float* h2, float* y) {
  for (int i = 0; i < kMlpH; ++i) {
    float s = P[kOffB1 + i];
    for (int j = 0; j < kMlpIn; ++j) s = fmaf(P[kOffW1 + i * kMlpIn + j], x[j], s);
    h1[i] = tanhf(s);
  }
  for (int i = 0; i < kMlpH; ++i) {
    float s = P[kOffB2 + i];
    for (int j = 0; j < kMlpH; ++j) s = fmaf(P[kOffW2 + i * kMlpH + j], h1[j], s);
    h2[i] = tanhf(s);
  }
  for (int i = 0; i < kMlpOut; ++i) {
    float s = P[kOffB3 + i];
    for (int j = 0; j < kMlpH; ++j) s = fmaf(P[kOffW3 + i * kMlpH + j], h2[j], s);
    y[i] = s;
  }
}

__global__ void mlp_forward_kernel(const float* params, const float* X, int n, float* Y) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float h1[kMlpH], h2[kMlpH], y[kMlpOut];
  forward_one(params, X + (int64_t)e * kMlpIn, h1, h2, y);
  for (int i = 0; i < kMlpOut; ++i) Y[(int64_t)e * kMlpOut + i] = y[i];
}

__global__ __launch_bounds__(kMlpThreads) void mlp_train_kernel(MlpTrainArgs a) {
  __shared__ float sP[kMlpParams];
  __shared__ float sX[kMlpBatch * kMlpIn];
  __shared__ float sA[kMlpBatch * kMlpH];  // h1, then delta of layer 1
  __shared__ float sB[kMlpBatch * kMlpH];  // h2, then delta of layer 2
  __shared__ float sD[kMlpBatch * kMlpOut];
  __shared__ float sG[kMlpParams];         // gradient of the current step
  __shared__ float sLoss[kMlpThreads];
  __shared__ int sBad;
  const int tid = threadIdx.x;
  const int B = a.batch;
  for (int i = tid; i < kMlpParams; i += kMlpThreads) sP[i] = a.params[i];
  __syncthreads();
  for (int step = 0; step < a.steps; ++step) {
    const int32_t* idx = a.batch_idx + (int64_t)step * B;
    if (tid == 0) sBad = 0;
    __syncthreads();
    if (tid < B && (idx[tid] < 0 || idx[tid] >= a.n)) sBad = 1;
    __syncthreads();
    const int bad = sBad;
    __syncthreads();  // every thread has read the flag before it is cleared again
    if (bad) {  // uniform across the workgroup: skip this minibatch
      if (tid == 0 && a.loss) a.loss[step] = __builtin_nanf("");
      continue;
    }
    // forward, one example per thread
    float loss = 0.0f;
    if (tid < B) {
      const int e = idx[tid];
      float* x = sX + tid * kMlpIn;
      for (int j = 0; j < kMlpIn; ++j) x[j] = a.X[(int64_t)e * kMlpIn + j];
      float* h1 = sA + tid * kMlpH;
      float* h2 = sB + tid * kMlpH;
      float y[kMlpOut];
      forward_one(sP, x, h1, h2, y);
      for (int i = 0; i < kMlpOut; ++i) {
        const float r = y[i] - a.Y[(int64_t)e * kMlpOut + i];
        loss += r * r;
        // d(mean squared error over batch and outputs)/dy
        sD[tid * kMlpOut + i] = 2.0f * r / (float)(B * kMlpOut);
      }
    }
    sLoss[tid] = loss;
    __syncthreads();
    if (tid == 0 && a.loss) {
      float s = 0.0f;
      for (int t = 0; t < B; ++t) s += sLoss[t];
      a.loss[step] = s / (float)(B * kMlpOut);
    }
    // layer 3 gradients: dW3[i][j] = sum_e d[e][i] h2[e][j], db3[i] = sum_e d[e][i]
    for (int p = tid; p < kMlpOut * kMlpH + kMlpOut; p += kMlpThreads) {
      float s = 0.0f;
      if (p < kMlpOut * kMlpH) {
        const int i = p / kMlpH, j = p % kMlpH;
        for (int e = 0; e < B; ++e) s = fmaf(sD[e * kMlpOut + i], sB[e * kMlpH + j], s);
        sG[kOffW3 + p] = s;
      } else {
        const int i = p - kMlpOut * kMlpH;
        for (int e = 0; e < B; ++e) s += sD[e * kMlpOut + i];
        sG[kOffB3 + i] = s;
      }
    }
    __syncthreads();
    // delta of layer 2 replaces h2
    if (tid < B) {
      float* h2 = sB + tid * kMlpH;
      for (int j = 0; j < kMlpH; ++j) {
        float s = 0.0f;
        for (int i = 0; i < kMlpOut; ++i) s = fmaf(sP[kOffW3 + i * kMlpH + j], sD[tid * kMlpOut + i], s);
        h2[j] = s * (1.0f - h2[j] * h2[j]);
      }
    }
    __syncthreads();
    // layer 2 gradients: dW2[i][j] = sum_e d2[e][i] h1[e][j]
    for (int p = tid; p < kMlpH * kMlpH + kMlpH; p += kMlpThreads) {
      float s = 0.0f;
      if (p < kMlpH * kMlpH) {
        const int i = p / kMlpH, j = p % kMlpH;
        for (int e = 0; e < B; ++e) s = fmaf(sB[e * kMlpH + i], sA[e * kMlpH + j], s);
        sG[kOffW2 + p] = s;
      } else {
        const int i = p - kMlpH * kMlpH;
        for (int e = 0; e < B; ++e) s += sB[e * kMlpH + i];
        sG[kOffB2 + i] = s;
      }
    }
    __syncthreads();
    // delta of layer 1 replaces h1
    if (tid < B) {
      float* h1 = sA + tid * kMlpH;
      const float* d2 = sB + tid * kMlpH;
      for (int j = 0; j < kMlpH; ++j) {
        float s = 0.0f;
        for (int i = 0; i < kMlpH; ++i) s = fmaf(sP[kOffW2 + i * kMlpH + j], d2[i], s);
        h1[j] = s * (1.0f - h1[j] * h1[j]);  // reads only its own h1[j]
      }
    }
    __syncthreads();
    // layer 1 gradients: dW1[i][j] = sum_e d1[e][i] x[e][j]
    for (int p = tid; p < kMlpH * kMlpIn + kMlpH; p += kMlpThreads) {
      float s = 0.0f;
      if (p < kMlpH * kMlpIn) {
        const int i = p / kMlpIn, j = p % kMlpIn;
        for (int e = 0; e < B; ++e) s = fmaf(sA[e * kMlpH + i], sX[e * kMlpIn + j], s);
        sG[kOffW1 + p] = s;
      } else {
        const int i = p - kMlpH * kMlpIn;
        for (int e = 0; e < B; ++e) s += sA[e * kMlpH + i];
        sG[kOffB1 + i] = s;
      }
    }
    __syncthreads();
    if (step == 0 && a.grad0)
      for (int i = tid; i < kMlpParams; i += kMlpThreads) a.grad0[i] = sG[i];
    // Adam with bias correction (t counts from a.t0 + 1)
    const float t = (float)(a.t0 + step + 1);
    const float c1 = 1.0f - powf(a.beta1, t), c2 = 1.0f - powf(a.beta2, t);
    for (int i = tid; i < kMlpParams; i += kMlpThreads) {
      const float g = sG[i] + a.weight_decay * sP[i];
      const float m = a.beta1 * a.adam_m[i] + (1.0f - a.beta1) * g;
      const float v = a.beta2 * a.adam_v[i] + (1.0f - a.beta2) * g * g;
      a.adam_m[i] = m;
      a.adam_v[i] = v;
      sP[i] -= a.lr * (m / c1) / (sqrtf(v / c2) + a.eps);
    }
    __syncthreads();
  }
  for (int i = tid; i < kMlpParams; i += kMlpThreads) a.params[i] = sP[i];
}

}  // namespace

hipError_t launch_mlp_forward(const float* params, const float* X, int n, float* Y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, params, X, n, Y);
  return hipGetLastError();
}

hipError_t launch_mlp_train(const MlpTrainArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(kMlpThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace tfcehip
