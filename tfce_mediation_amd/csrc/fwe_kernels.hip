// Family-wise-error lookups of element statistics in a sorted max-TFCE null
// (FP64, as the reference's numpy code).
//   mode 0 "left" - calculate_fweP_vertex.py:37-42 / calculate_fweP_voxel.py:23-28
//     for the elements with value > 0 (the scripts' mask data > 0, :54-55):
//     idx = #(null < value); p = idx == 0 ? 0 : (idx - 1) / N; 0 elsewhere.
//   mode 1 "nearest" - tm_func.py:566-573 / pyfunc.py:254-261 for every element:
//     idx as above; idx == N -> p[N-1]; otherwise the nearer of null[idx-1]
//     (index -1 wraps to the last entry, as Python indexing does) and
//     null[idx], ties to idx; p[j] = j / N.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfcehip {

namespace {

__device__ __forceinline__ int lower_bound(const double* s, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <class T>
__global__ void fwe_lookup_kernel(const double* sorted, int n, const T* vals, int64_t V, int mode,
                                  double* out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const double x = (double)vals[v];
  const double N = (double)n;
  if (mode == 0) {
    if (!(x > 0.0)) { out[v] = 0.0; return; }
    const int idx = lower_bound(sorted, n, x);
    out[v] = idx == 0 ? 0.0 : (double)(idx - 1) / N;
    return;
  }
  const int idx = (x != x) ? n : lower_bound(sorted, n, x);  // NaN sorts last in numpy
  if (idx == n) { out[v] = (double)(n - 1) / N; return; }
  const int before = idx == 0 ? n - 1 : idx - 1;
  out[v] = (fabs(x - sorted[before]) < fabs(x - sorted[idx])) ? (double)before / N : (double)idx / N;
}

}  // namespace

hipError_t launch_fwe_lookup(const double* sorted, int n, const float* vals, int64_t V, int mode, double* out,
                             hipStream_t s) {
  if (V <= 0) return hipSuccess;
  hipLaunchKernelGGL(fwe_lookup_kernel<float>, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, sorted, n, vals,
                     V, mode, out);
  return hipGetLastError();
}

hipError_t launch_fwe_lookup64(const double* sorted, int n, const double* vals, int64_t V, int mode, double* out,
                               hipStream_t s) {
  if (V <= 0) return hipSuccess;
  hipLaunchKernelGGL(fwe_lookup_kernel<double>, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, sorted, n,
                     vals, V, mode, out);
  return hipGetLastError();
}

}  // namespace tfcehip
