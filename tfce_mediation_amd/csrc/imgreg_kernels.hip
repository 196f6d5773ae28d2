// Per-element regression with an element-wise regressor, and the variance
// inflation image of the step-1 scripts (pyfunc.py:1024-1114), FP64 per
// element.  One thread per element; the shared design lives in LDS, the data
// are read in [subject][element] order (coalesced across a wave).
#include <hip/hip_runtime.h>
#include <math.h>

#include "ols_internal.h"

namespace tfcehip {
namespace {

// image_regression (pyfunc.py:1024-1086) for element v: the design is the
// shared columns R (intercept first) and the element's own regressor z.  With
// A = R'R (its inverse Ainv from the host), b = R'z, c = z'z, d = R'y, e = z'y
// and u = Ainv b, the block inverse of X'X = [[A, b], [b', c]] has Schur
// complement s = c - b'u:
//   coefficients  a_z = (e - u'd) / s,  a_R = Ainv d - u a_z;
//   diag(inv(X'X))  Ainv_jj + u_j^2 / s  and  1 / s.
// sigma2 = sum (y - X a)^2 / (n - k) from the residuals themselves; t = a /
// sqrt(sigma2 * diag), all in double, stored as float.  Elements whose
// regressor holds a NaN or has a standard deviation (population, as numpy's
// std) below std_min get t = 0 in every column (pyfunc.py:1069-1070).
__global__ __launch_bounds__(256) void image_regression_kernel(ImgRegArgs a) {
  extern __shared__ double sR[];  // [n][r]
  for (int i = threadIdx.x; i < a.n * a.r; i += blockDim.x) sR[i] = a.R[i];
  __syncthreads();
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.V) return;
  const int n = a.n, r = a.r, k = r + 1;
  constexpr int RM = kOlsMaxK - 1;
  const float* y = a.Y + v;
  const float* z = a.Z + v;
  // regressor mean (and NaN check)
  double sz = 0.0;
  bool nan = false;
  for (int i = 0; i < n; ++i) {
    const double zi = (double)z[(int64_t)i * a.ldz];
    nan |= isnan(zi);
    sz += zi;
  }
  const double mz = sz / (double)n;
  double d[RM], b[RM];
#pragma unroll
  for (int j = 0; j < RM; ++j) d[j] = b[j] = 0.0;
  double ssz = 0.0, c = 0.0, e = 0.0;
  for (int i = 0; i < n; ++i) {
    const double zi = (double)z[(int64_t)i * a.ldz];
    const double yi = (double)y[(int64_t)i * a.ldy];
    const double dz = zi - mz;
    ssz = fma(dz, dz, ssz);
    c = fma(zi, zi, c);
    e = fma(zi, yi, e);
#pragma unroll
    for (int j = 0; j < RM; ++j)
      if (j < r) {
        d[j] = fma(sR[i * r + j], yi, d[j]);
        b[j] = fma(sR[i * r + j], zi, b[j]);
      }
  }
  if (nan || !(sqrt(ssz / (double)n) >= a.std_min)) {
    for (int j = 0; j < k; ++j) a.T[(int64_t)j * a.ldt + v] = 0.0f;
    return;
  }
  double u[RM], g[RM];
#pragma unroll
  for (int j = 0; j < RM; ++j) {
    double su = 0.0, sg = 0.0;
    if (j < r)
      for (int q = 0; q < r; ++q) {
        su = fma(a.Ainv[j * r + q], b[q], su);
        sg = fma(a.Ainv[j * r + q], d[q], sg);
      }
    u[j] = su;
    g[j] = sg;
  }
  double bu = 0.0, ud = 0.0;
#pragma unroll
  for (int j = 0; j < RM; ++j)
    if (j < r) {
      bu = fma(b[j], u[j], bu);
      ud = fma(u[j], d[j], ud);
    }
  const double s = c - bu;
  const double az = (e - ud) / s;
  double coef[RM];
#pragma unroll
  for (int j = 0; j < RM; ++j) coef[j] = j < r ? g[j] - u[j] * az : 0.0;
  double sse = 0.0;
  for (int i = 0; i < n; ++i) {
    const double zi = (double)z[(int64_t)i * a.ldz];
    const double yi = (double)y[(int64_t)i * a.ldy];
    double fit = zi * az;
#pragma unroll
    for (int j = 0; j < RM; ++j)
      if (j < r) fit = fma(sR[i * r + j], coef[j], fit);
    const double res = yi - fit;
    sse = fma(res, res, sse);
  }
  const double s2 = sse / (double)(n - k);
#pragma unroll
  for (int j = 0; j < RM; ++j)
    if (j < r) {
      const double diag = a.Ainv[j * r + j] + u[j] * u[j] / s;
      a.T[(int64_t)j * a.ldt + v] = (float)(coef[j] / sqrt(s2 * diag));
    }
  a.T[(int64_t)r * a.ldt + v] = (float)(az / sqrt(s2 / s));
}

// image_reg_VIF (pyfunc.py:1088-1114) for element v: R2 = 1 - RSS / TSS and
// VIF = 1 / (1 - R2), NaN set to 0.  RSS from the FP64 least-squares fit of
// y on X (intercept first, cy_lin_lstsqr_mat); TSS = sum((y - mean(y))^2) in
// the data's own type T, summed subject by subject as numpy reduces over the
// first axis of an [n][V] array.
template <class T>
__global__ __launch_bounds__(256) void image_vif_kernel(ImgVifArgs a) {
  extern __shared__ double sX[];  // [n][k]
  for (int i = threadIdx.x; i < a.n * a.k; i += blockDim.x) sX[i] = a.X[i];
  __syncthreads();
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.V) return;
  const int n = a.n, k = a.k;
  const T* y = reinterpret_cast<const T*>(a.Y) + v;
  T sum = T(0);
  double xty[kOlsMaxK];
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j) xty[j] = 0.0;
  for (int i = 0; i < n; ++i) {
    const T yi = y[(int64_t)i * a.ldy];
    sum = sum + yi;
#pragma unroll
    for (int j = 0; j < kOlsMaxK; ++j)
      if (j < k) xty[j] = fma(sX[i * k + j], (double)yi, xty[j]);
  }
  // numpy divides the sum by the element count in double and stores the
  // mean in the array's type
  const T mean = (T)((double)sum / (double)n);
  double coef[kOlsMaxK];
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j) {
    double s = 0.0;
    if (j < k)
      for (int q = 0; q < k; ++q) s = fma(a.Ainv[j * k + q], xty[q], s);
    coef[j] = s;
  }
  T tss = T(0);
  double rss = 0.0;
  for (int i = 0; i < n; ++i) {
    const T yi = y[(int64_t)i * a.ldy];
    const T dl = yi - mean;
    tss = tss + dl * dl;
    double fit = 0.0;
#pragma unroll
    for (int j = 0; j < kOlsMaxK; ++j)
      if (j < k) fit = fma(sX[i * k + j], coef[j], fit);
    const double res = (double)yi - fit;
    rss = fma(res, res, rss);
  }
  const double r2 = 1.0 - rss / (double)tss;
  const double vif = 1.0 / (1.0 - r2);
  a.vif[v] = isnan(vif) ? 0.0 : vif;
}

// Designs above 64 KiB of LDS (up to 2,048 subjects x 8 columns = 128 KiB)
// need the kernel's dynamic-LDS limit raised first.
template <class K>
hipError_t allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

}  // namespace

hipError_t launch_image_regression(const ImgRegArgs& a, hipStream_t s) {
  const int64_t blocks = (a.V + 255) / 256;
  const size_t lds = sizeof(double) * (size_t)a.n * a.r;
  hipError_t e = allow_lds(image_regression_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(image_regression_kernel, dim3((unsigned)blocks), dim3(256), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_image_vif(const ImgVifArgs& a, bool f64, hipStream_t s) {
  const int64_t blocks = (a.V + 255) / 256;
  const size_t lds = sizeof(double) * (size_t)a.n * a.k;
  hipError_t e = f64 ? allow_lds(image_vif_kernel<double>, lds) : allow_lds(image_vif_kernel<float>, lds);
  if (e != hipSuccess) return e;
  if (f64)
    hipLaunchKernelGGL(image_vif_kernel<double>, dim3((unsigned)blocks), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL(image_vif_kernel<float>, dim3((unsigned)blocks), dim3(256), lds, s, a);
  return hipGetLastError();
}

}  // namespace tfcehip
