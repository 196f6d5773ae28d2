// One-call forms of the batched statistics with the argument lists of the
// drop-in boundary: each builds the prepared-data and design handles, runs one
// batch of permutations and releases them (blocking).  Callers that run many
// batches keep the handles instead (tfce_ols_data_create, tfce_ols_model_create).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "tfcehip.h"
#include "capi_common.h"

using namespace tfcehip;

namespace {

struct Handles {
  tfce_ols_data_t* data = nullptr;
  tfce_ols_model_t* a = nullptr;
  tfce_ols_model_t* b = nullptr;
  float* tmp = nullptr;
  ~Handles() {
    if (a) tfce_ols_model_destroy(a);
    if (b) tfce_ols_model_destroy(b);
    if (data) tfce_ols_data_destroy(data);
    if (tmp) (void)hipFree(tmp);
  }
};

bool constant_column(const double* X, int32_t n, int32_t k, int32_t j) {
  for (int32_t i = 1; i < n; ++i)
    if (X[(size_t)i * k + j] != X[j]) return false;
  return X[j] != 0.0;
}

// t of column 1 of [1, x] against y taken as float32, with se rounded to
// float32 as tval_int stores it (cynumstats.pyx:47-64); the scalar A path of
// type-Y mediation.
double scalar_t(const double* x, const double* y32, int32_t n) {
  std::vector<double> y(y32, y32 + n);
  for (double& v : y) v = (double)(float)v;
  double sx = 0, sy = 0, sxx = 0, sxy = 0;
  for (int32_t i = 0; i < n; ++i) {
    sx += x[i];
    sy += y[i];
    sxx += x[i] * x[i];
    sxy += x[i] * y[i];
  }
  const double det = n * sxx - sx * sx;
  const double i00 = sxx / det, i01 = -sx / det, i11 = n / det;
  const double a0 = i00 * sy + i01 * sxy, a1 = i01 * sy + i11 * sxy;
  double sse = 0;
  for (int32_t i = 0; i < n; ++i) {
    const double r = y[i] - a0 - a1 * x[i];
    sse += r * r;
  }
  const float se = (float)std::sqrt(sse / (n - 2) * i11);
  return a1 / (double)se;
}

}  // namespace

int ols_perm_t(const double* X, const float* d_Y, const int32_t* d_perm, int32_t n, int32_t k, int64_t V,
               int32_t contrast, int32_t precision, int32_t P, float* d_t, hipStream_t stream) {
  if (!X || !d_Y || !d_perm || !d_t) return set_error(TFCE_EINVAL, "NULL argument");
  if (k < 1 || k > 8 || contrast < 0 || contrast >= k) return set_error(TFCE_EINVAL, "need 1 <= k <= 8, 0 <= contrast < k");
  if (P < 1) return set_error(TFCE_EINVAL, "P must be >= 1");
  uint32_t mask = 0;
  bool has_const = false;
  for (int32_t j = 0; j < k; ++j) {
    if (constant_column(X, n, k, j))
      has_const = true;
    else
      mask |= 1u << j;
  }
  Handles h;
  int rc = tfce_ols_data_create(d_Y, n, V, V, precision, has_const ? 1 : 0, &h.data, stream);
  if (rc) return rc;
  rc = tfce_ols_model_create(h.data, X, k, mask, &contrast, 1, &h.a, stream);
  if (rc) return rc;
  rc = tfce_ols_perm_t(h.a, d_perm, P, d_t, V, stream);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return set_error(-(int)e, "ols_perm_t");
  return 0;
}

int sobel_perm_z(const float* d_M, int32_t n, int64_t V, const double* x, const double* y, int32_t medtype,
                 int32_t alg, const int32_t* d_perm, int32_t P, int32_t precision, float* d_z, hipStream_t stream) {
  if (!d_M || !x || !y || !d_perm || !d_z) return set_error(TFCE_EINVAL, "NULL argument");
  if (medtype < 0 || medtype > 2) return set_error(TFCE_EINVAL, "medtype must be 0 (M), 1 (I) or 2 (Y)");
  if (P < 1 || n < 4) return set_error(TFCE_EINVAL, "need P >= 1 and n >= 4");
  Handles h;
  int rc = tfce_ols_data_create(d_M, n, V, V, precision, 1, &h.data, stream);
  if (rc) return rc;
  std::vector<double> XA((size_t)n * 2), XB((size_t)n * 3);
  for (int32_t i = 0; i < n; ++i) {
    XA[(size_t)i * 2] = 1.0;
    XA[(size_t)i * 2 + 1] = x[i];
    XB[(size_t)i * 3] = 1.0;
    // type M: [1, y, x], t of y; type I: [1, x, y], t of x; type Y: [1, y, x], t of y
    XB[(size_t)i * 3 + 1] = medtype == 1 ? x[i] : y[i];
    XB[(size_t)i * 3 + 2] = medtype == 1 ? y[i] : x[i];
  }
  const int32_t con = 1;
  // types M and I permute x only; type Y permutes x and y together
  // (vertex_tfce_mediation_randomise.py:82-90)
  const uint32_t mask_b = medtype == 0 ? (1u << 2) : (medtype == 1 ? (1u << 1) : ((1u << 1) | (1u << 2)));
  rc = tfce_ols_model_create(h.data, XB.data(), 3, mask_b, &con, 1, &h.b, stream);
  if (rc) return rc;
  hipError_t e;
  if (medtype == 2) {
    // type Y: path A is the scalar x -> y regression
    if ((e = hipMalloc(&h.tmp, sizeof(float) * (size_t)P * (size_t)V)) != hipSuccess)
      return set_error(-(int)e, "hipMalloc(t)");
    rc = tfce_ols_perm_t(h.b, d_perm, P, h.tmp, V, stream);
    if (rc) return rc;
    rc = tfce_sobel_z(nullptr, h.tmp, (int64_t)P * V, alg, (float)scalar_t(x, y, n), d_z, stream);
    if (rc) return rc;
  } else {
    // types M and I: both paths permute x only, one contraction serves both
    rc = tfce_ols_model_create(h.data, XA.data(), 2, 1u << 1, &con, 1, &h.a, stream);
    if (rc) return rc;
    rc = tfce_ols_perm_sobel(h.a, h.b, d_perm, P, alg, d_z, V, stream);
    if (rc) return rc;
  }
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return set_error(-(int)e, "sobel_perm_z");
  return 0;
}
