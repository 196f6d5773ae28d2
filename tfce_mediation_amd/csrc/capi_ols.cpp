// C-ABI host layer of the permuted-OLS path: prepared data, design models,
// batched permuted t-values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "tfcehip.h"
#include "capi_common.h"
#include "mlp_internal.h"
#include "ols_internal.h"

using namespace tfcehip;

struct tfce_ols_data {
  int32_t n = 0, Kp = 0, precision = 0, pieces = 0, center = 0, device = 0;
  int64_t V = 0, Npad = 0;
  double* mu = nullptr;
  double* yy = nullptr;
  float* yinv = nullptr;  // fp16 mode: per-element inverse scale
  float* plane32 = nullptr;
  uint16_t* bf[3] = {nullptr, nullptr, nullptr};
};

struct tfce_ols_model {
  const tfce_ols_data* data = nullptr;
  int32_t k = 0, f = 0, m = 0, n_con = 0;
  int32_t con[kOlsMaxContrasts] = {0};
  int32_t con_is_intercept[kOlsMaxContrasts] = {0};
  bool shared = false;
  double* F = nullptr;   // [n][f]
  double* Z = nullptr;   // [n][m]
  double* FF = nullptr;  // [f][f]
  double* ZZ = nullptr;  // [m][m]
  float* Z32 = nullptr;  // [n][m]
  double* colsumZ = nullptr;
  float* gF = nullptr;
  float* hF = nullptr;
  double* r0 = nullptr;
  float* Linv1 = nullptr;  // shared factors
  double* dc1 = nullptr;
  float zscale[kOlsMaxK] = {1, 1, 1, 1, 1, 1, 1, 1};  // fp16 mode: per permuted column
  std::vector<double> hZ;  // [n][m] permuted columns (host copy: pairing checks)
  int32_t internal[kOlsMaxK] = {0};  // caller column -> internal column
  bool constant[kOlsMaxK] = {false};  // caller column is constant
};

namespace {

std::mutex g_ols_mutex;
struct OlsWs {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool captured = false;  // a graph captured on this stream holds ptr
};
std::map<std::pair<int, hipStream_t>, OlsWs> g_ols_ws;

// Grow-only scratch of the contraction per (device, stream); the capture rule
// of workspace_get (capi_tfce.cpp) applies.
int ols_ws(hipStream_t stream, size_t bytes, void** out) {
  std::lock_guard<std::mutex> g(g_ols_mutex);
  OlsWs& w = g_ols_ws[std::make_pair(current_device(), stream)];
  if (stream_capturing(stream)) {
    if (w.bytes < bytes)
      return set_error(TFCE_EINVAL, "contraction scratch of this stream is too small to capture this call: "
                                    "run it once eagerly on the capture stream first");
    w.captured = true;
    *out = w.ptr;
    return 0;
  }
  if (w.bytes < bytes) {
    if (w.ptr) {
      scratch_release(w.ptr, w.captured, stream);
      w.ptr = nullptr;
      w.bytes = 0;
      w.captured = false;
    }
    hipError_t e = hipMalloc(&w.ptr, bytes);
    if (e != hipSuccess) return set_error(-(int)e, std::string("hipMalloc(ols scratch): ") + hipGetErrorString(e));
    w.bytes = bytes;
  }
  *out = w.ptr;
  return 0;
}

template <class T>
int dalloc(T** p, size_t count, const char* what) {
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (count ? count : 1));
  if (e != hipSuccess) return set_error(-(int)e, std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
  return 0;
}

template <class T>
int upload(T* d, const T* h, size_t count, const char* what) {
  if (!count) return 0;
  hipError_t e = hipMemcpy(d, h, sizeof(T) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return set_error(-(int)e, std::string("upload ") + what + ": " + hipGetErrorString(e));
  return 0;
}

// Lower Cholesky factor and its inverse (row-major k x k); false if not SPD,
// or if a pivot is below 1e-13 of its diagonal entry (a column that is a
// combination of the ones before it, up to rounding).
bool chol_inv(const std::vector<double>& G, int k, std::vector<double>& Li) {
  std::vector<double> L(k * k, 0.0);
  Li.assign(k * k, 0.0);
  for (int j = 0; j < k; ++j) {
    double d = G[j * k + j];
    for (int q = 0; q < j; ++q) d -= L[j * k + q] * L[j * k + q];
    if (!(d > 1e-13 * G[j * k + j])) return false;
    d = std::sqrt(d);
    L[j * k + j] = d;
    for (int i = j + 1; i < k; ++i) {
      double s = G[i * k + j];
      for (int q = 0; q < j; ++q) s -= L[i * k + q] * L[j * k + q];
      L[i * k + j] = s / d;
    }
  }
  for (int j = 0; j < k; ++j) {
    Li[j * k + j] = 1.0 / L[j * k + j];
    for (int i = j + 1; i < k; ++i) {
      double s = 0.0;
      for (int q = j; q < i; ++q) s += L[i * k + q] * Li[q * k + j];
      Li[i * k + j] = -s / L[i * k + i];
    }
  }
  return true;
}

void free_model(tfce_ols_model* m) {
  if (!m) return;
  for (void* p : {(void*)m->F, (void*)m->Z, (void*)m->FF, (void*)m->ZZ, (void*)m->Z32, (void*)m->colsumZ,
                  (void*)m->gF, (void*)m->hF, (void*)m->r0, (void*)m->Linv1, (void*)m->dc1})
    if (p) (void)hipFree(p);
  delete m;
}

}  // namespace

namespace tfcehip {
hipError_t launch_fwe_lookup(const double* sorted, int n, const float* vals, int64_t V, int mode, double* out,
                             hipStream_t s);
hipError_t launch_fwe_lookup64(const double* sorted, int n, const double* vals, int64_t V, int mode, double* out,
                               hipStream_t s);
}

extern "C" {

int tfce_fwe_p(const double* d_sorted_null, int32_t n_null, const float* d_values, int64_t V, int32_t mode,
               double* d_out, hipStream_t stream) {
  if (!d_sorted_null || !d_values || !d_out) return set_error(TFCE_EINVAL, "NULL argument");
  if (n_null < 1) return set_error(TFCE_EINVAL, "the null needs at least one value");
  if (mode != 0 && mode != 1) return set_error(TFCE_EINVAL, "mode must be 0 (left) or 1 (nearest)");
  hipError_t e = launch_fwe_lookup(d_sorted_null, n_null, d_values, V, mode, d_out, stream);
  if (e != hipSuccess) return set_error(-(int)e, "fwe lookup launch");
  return 0;
}

int tfce_fwe_p64(const double* d_sorted_null, int32_t n_null, const double* d_values, int64_t V, int32_t mode,
                 double* d_out, hipStream_t stream) {
  if (!d_sorted_null || !d_values || !d_out) return set_error(TFCE_EINVAL, "NULL argument");
  if (n_null < 1) return set_error(TFCE_EINVAL, "the null needs at least one value");
  if (mode != 0 && mode != 1) return set_error(TFCE_EINVAL, "mode must be 0 (left) or 1 (nearest)");
  hipError_t e = launch_fwe_lookup64(d_sorted_null, n_null, d_values, V, mode, d_out, stream);
  if (e != hipSuccess) return set_error(-(int)e, "fwe lookup launch");
  return 0;
}

int tfce_ols_data_create(const float* d_Y, int32_t n, int64_t V, int64_t ldy, int32_t precision,
                         int32_t center, tfce_ols_data_t** out, hipStream_t stream) {
  if (!out || !d_Y) return set_error(TFCE_EINVAL, "NULL argument");
  *out = nullptr;
  if (n < 2 || V < 1 || ldy < V) return set_error(TFCE_EINVAL, "need n >= 2, V >= 1, ldy >= V");
  if (precision < kPrecFP32 || precision > kPrecFP16x3) return set_error(TFCE_EINVAL, "unknown precision");
  tfce_ols_data* d = new tfce_ols_data();
  d->n = n;
  d->Kp = (n + 15) / 16 * 16;
  d->V = V;
  d->Npad = (V + kColAlign - 1) / kColAlign * kColAlign;
  d->precision = precision;
  d->pieces = precision == kPrecFP32 ? 0 : ols_pieces(precision);
  d->center = center ? 1 : 0;
  d->device = current_device();
  const size_t plane = (size_t)d->Kp * d->Npad;
  int rc = 0;
  if ((rc = dalloc(&d->mu, V, "mu")) || (rc = dalloc(&d->yy, V, "yy")) ||
      (rc = dalloc(&d->plane32, plane, "plane32")) ||
      (precision == kPrecFP16x3 && (rc = dalloc(&d->yinv, V, "yinv")))) {
    tfce_ols_data_destroy(d);
    return rc;
  }
  for (int p = 0; p < d->pieces; ++p)
    if ((rc = dalloc(&d->bf[p], plane, "bf16 plane"))) {
      tfce_ols_data_destroy(d);
      return rc;
    }
  OlsPrepArgs a{};
  a.Y = d_Y;
  a.n = n;
  a.Kp = d->Kp;
  a.V = V;
  a.ldy = ldy;
  a.Npad = d->Npad;
  a.center = d->center;
  a.pieces = d->pieces;
  a.f16 = precision == kPrecFP16x3 ? 1 : 0;
  a.yinv = d->yinv;
  a.mu = d->mu;
  a.yy = d->yy;
  a.plane32 = d->plane32;
  for (int p = 0; p < 3; ++p) a.bf[p] = d->bf[p];
  hipError_t e = launch_ols_prep(a, stream);
  if (e != hipSuccess) {
    tfce_ols_data_destroy(d);
    return set_error(-(int)e, "ols prep launch");
  }
  *out = d;
  return 0;
}

int tfce_ols_data_destroy(tfce_ols_data_t* d) {
  if (!d) return 0;
  for (void* p : {(void*)d->mu, (void*)d->yy, (void*)d->yinv, (void*)d->plane32, (void*)d->bf[0], (void*)d->bf[1],
                  (void*)d->bf[2]})
    if (p) (void)hipFree(p);
  delete d;
  return 0;
}

int tfce_ols_model_create(const tfce_ols_data_t* data, const double* X, int32_t k, uint32_t perm_mask,
                          const int32_t* contrasts, int32_t n_contrasts, tfce_ols_model_t** out,
                          hipStream_t stream) {
  if (!out || !data || !X || !contrasts) return set_error(TFCE_EINVAL, "NULL argument");
  *out = nullptr;
  const int n = data->n;
  if (k < 1 || k > kOlsMaxK) return set_error(TFCE_EINVAL, "design must have 1..8 columns");
  if (n <= k) return set_error(TFCE_EINVAL, "need more subjects than design columns");
  if (n_contrasts < 1 || n_contrasts > kOlsMaxContrasts) return set_error(TFCE_EINVAL, "1..8 contrasts");
  if (perm_mask >= (1u << k)) return set_error(TFCE_EINVAL, "perm_mask names a column beyond k");
  std::vector<int> order, is_fixed(k);
  for (int j = 0; j < k; ++j)
    if (!((perm_mask >> j) & 1u)) order.push_back(j);
  const int f = (int)order.size();
  for (int j = 0; j < k; ++j)
    if ((perm_mask >> j) & 1u) order.push_back(j);
  const int m = k - f;
  if (m < 1) return set_error(TFCE_EINVAL, "at least one column must be permuted");
  std::vector<int> internal(k);
  for (int i = 0; i < k; ++i) internal[order[i]] = i;
  // constant columns: the intercept the centred data relies on
  std::vector<double> constval(k, 0.0);
  bool has_const = false;
  for (int j = 0; j < k; ++j) {
    bool c = X[j] != 0.0;
    for (int i = 1; i < n && c; ++i) c = X[i * k + j] == X[j];
    if (c) {
      constval[j] = X[j];
      has_const = true;
    }
  }
  if (data->center && !has_const)
    return set_error(TFCE_EINVAL, "data was centred but the design has no intercept column");
  tfce_ols_model* md = new tfce_ols_model();
  md->data = data;
  md->k = k;
  md->f = f;
  md->m = m;
  md->n_con = n_contrasts;
  for (int c = 0; c < n_contrasts; ++c) {
    if (contrasts[c] < 0 || contrasts[c] >= k) {
      free_model(md);
      return set_error(TFCE_EINVAL, "contrast index out of range");
    }
    md->con[c] = internal[contrasts[c]];
    md->con_is_intercept[c] = (data->center && constval[contrasts[c]] != 0.0) ? 1 : 0;
    if (md->con_is_intercept[c] && constval[contrasts[c]] != 1.0) {
      free_model(md);
      return set_error(TFCE_ENOTSUP, "a contrast on a constant column must use the value 1");
    }
  }
  std::vector<double> F(n * std::max(f, 1)), Z(n * m), FF(std::max(f * f, 1)), ZZ(m * m), colsum(m, 0.0);
  std::vector<float> Z32(n * m);
  for (int i = 0; i < n; ++i) {
    for (int q = 0; q < f; ++q) F[i * f + q] = X[i * k + order[q]];
    for (int q = 0; q < m; ++q) {
      Z[i * m + q] = X[i * k + order[f + q]];
      Z32[i * m + q] = (float)Z[i * m + q];
      colsum[q] += (double)Z32[i * m + q];
    }
  }
  // fp16 mode: power of two per permuted column with max|z| * s in [2^13, 2^14)
  for (int q = 0; q < m; ++q) {
    float mx = 0.0f;
    for (int i = 0; i < n; ++i) mx = std::max(mx, std::fabs(Z32[i * m + q]));
    md->zscale[q] = (mx > 0.0f && std::isfinite(mx)) ? std::ldexp(1.0f, 13 - std::ilogb(mx)) : 1.0f;
  }
  for (int a = 0; a < f; ++a)
    for (int b = 0; b < f; ++b) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += F[i * f + a] * F[i * f + b];
      FF[a * f + b] = s;
    }
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += Z[i * m + a] * Z[i * m + b];
      ZZ[a * m + b] = s;
    }
  std::vector<double> LFF;
  if (f > 0 && !chol_inv(FF, f, LFF)) {
    free_model(md);
    return set_error(TFCE_EINVAL, "fixed design columns are rank deficient");
  }
  // the Gram matrix of X[perm] is the same for every permutation when every
  // fixed column is constant (F'PZ = F'Z)
  bool shared = true;
  for (int q = 0; q < f; ++q) shared = shared && constval[order[q]] != 0.0;
  md->shared = shared;
  md->hZ = Z;
  for (int j = 0; j < k; ++j) {
    md->internal[j] = internal[j];
    md->constant[j] = constval[j] != 0.0;
  }
  int rc = 0;
  const int64_t V = data->V;
  if ((rc = dalloc(&md->F, F.size(), "F")) || (rc = dalloc(&md->Z, Z.size(), "Z")) ||
      (rc = dalloc(&md->FF, FF.size(), "FF")) || (rc = dalloc(&md->ZZ, ZZ.size(), "ZZ")) ||
      (rc = dalloc(&md->Z32, Z32.size(), "Z32")) || (rc = dalloc(&md->colsumZ, m, "colsum")) ||
      (rc = dalloc(&md->gF, (size_t)std::max(f, 1) * V, "gF")) ||
      (rc = dalloc(&md->hF, (size_t)std::max(f, 1) * V, "hF")) || (rc = dalloc(&md->r0, V, "r0")) ||
      (rc = upload(md->F, F.data(), F.size(), "F")) || (rc = upload(md->Z, Z.data(), Z.size(), "Z")) ||
      (rc = upload(md->FF, FF.data(), FF.size(), "FF")) || (rc = upload(md->ZZ, ZZ.data(), ZZ.size(), "ZZ")) ||
      (rc = upload(md->Z32, Z32.data(), Z32.size(), "Z32")) ||
      (rc = upload(md->colsumZ, colsum.data(), (size_t)m, "colsum"))) {
    free_model(md);
    return rc;
  }
  double* dLFF = nullptr;
  if ((rc = dalloc(&dLFF, std::max<size_t>(LFF.size(), 1), "LFF")) ||
      (rc = upload(dLFF, LFF.data(), LFF.size(), "LFF"))) {
    if (dLFF) (void)hipFree(dLFF);
    free_model(md);
    return rc;
  }
  OlsFixedArgs fa{};
  fa.plane32 = data->plane32;
  fa.Kp = data->Kp;
  fa.n = n;
  fa.V = V;
  fa.Npad = data->Npad;
  fa.yy = data->yy;
  fa.F = md->F;
  fa.f = f;
  fa.LFF = dLFF;
  fa.gF = md->gF;
  fa.hF = md->hF;
  fa.r0 = md->r0;
  hipError_t e = launch_ols_fixed(fa, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(dLFF);
  if (e != hipSuccess) {
    free_model(md);
    return set_error(-(int)e, "ols fixed-column kernel");
  }
  if (shared) {
    if ((rc = dalloc(&md->Linv1, 64, "Linv")) || (rc = dalloc(&md->dc1, kOlsMaxContrasts, "dc"))) {
      free_model(md);
      return rc;
    }
    OlsFactorArgs ka{};
    ka.F = md->F;
    ka.Z = md->Z;
    ka.FF = md->FF;
    ka.ZZ = md->ZZ;
    ka.perm = nullptr;
    ka.n = n;
    ka.f = f;
    ka.m = m;
    ka.P = 1;
    ka.n_con = md->n_con;
    for (int c = 0; c < kOlsMaxContrasts; ++c) ka.con[c] = md->con[c];
    ka.Linv = md->Linv1;
    ka.dc = md->dc1;
    e = launch_ols_factors(ka, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
      free_model(md);
      return set_error(-(int)e, "ols factor kernel");
    }
  }
  *out = md;
  return 0;
}

int tfce_ols_model_destroy(tfce_ols_model_t* m) {
  free_model(m);
  return 0;
}

static int g_gemm_tile = 0;

int tfce_ols_set_tile(int32_t tile) {
  if (tile < 0 || tile >= kGemmTiles) return set_error(TFCE_EINVAL, "tile must be 0, 1, 2 or 3");
  g_gemm_tile = tile;
  return 0;
}

int tfce_ols_tile(void) { return g_gemm_tile; }

// t-values (tss == null) or sqrt(max(F, 0)) for P permutations
// (n_blk > 0: partial F of the column blocks blk_mask, internal column sets)
static int perm_run(const tfce_ols_model_t* md, const int32_t* d_perm, int32_t P, float* d_t, int64_t ldt,
                    const double* tss, hipStream_t stream, const tfce_ols_model_t* md2 = nullptr, int alg = 0,
                    int n_blk = 0, const uint32_t* blk_mask = nullptr) {
  if (!md || !d_t) return set_error(TFCE_EINVAL, "NULL argument");
  const tfce_ols_data* d = md->data;
  if (!d_perm) P = 1;
  if (P < 1) return set_error(TFCE_EINVAL, "P must be >= 1");
  if (ldt < d->V) return set_error(TFCE_EINVAL, "ldt must be >= V");
  // rows per permutation = permuted columns: a row tile holds floor(rows / m)
  // whole permutations (the epilogue reads a permutation's rows from LDS, so
  // they need not align with the 32-row MFMA tiles)
  const int rpp = md->m > 0 ? md->m : 1;
  const int tile = (d->precision == kPrecFP32 || n_blk > 0) ? 0 : g_gemm_tile;
  const int tile_rows = d->precision == kPrecFP32 ? kGemmRows : gemm_tile_rows(tile);
  const int tile_cols = d->precision == kPrecFP32 ? kGemmCols : gemm_tile_cols(tile);
  const int per_tile = tile_rows / rpp;
  const int64_t n_mtiles = ((int64_t)P + per_tile - 1) / per_tile;
  const int64_t Mpad = n_mtiles * tile_rows;
  const int pieces = d->pieces;
  const size_t abytes = (size_t)d->Kp * Mpad * (pieces ? 2 * pieces : 4);
  const size_t fbytes1 = (size_t)P * (64 * sizeof(float) + kOlsMaxContrasts * sizeof(double));
  // the block F's factors: per permutation, or once (P = 1) for a shared Gram matrix
  const int P_blk = n_blk > 0 ? (md->shared ? 1 : P) : 0;
  const size_t fbytes_blk =
      n_blk > 0 ? (size_t)P_blk * (64 * sizeof(float) + kOlsMaxContrasts * sizeof(double)) + 256 +
                      (size_t)P_blk * n_blk * 64 * sizeof(float) + 256
                : 0;
  const size_t fbytes = (md->shared || n_blk > 0 ? 0 : fbytes1) + ((md2 && !md2->shared) ? fbytes1 + 256 : 0) +
                        fbytes_blk;
  void* ws = nullptr;
  int rc = ols_ws(stream, abytes + fbytes + 512, &ws);
  if (rc) return rc;
  char* base = reinterpret_cast<char*>(ws);
  OlsBuildAArgs ba{};
  ba.Z32 = md->Z32;
  ba.perm = d_perm;
  ba.n = d->n;
  ba.m = md->m;
  ba.P = P;
  ba.Kp = d->Kp;
  ba.rpp = rpp;
  ba.tile_rows = tile_rows;
  ba.Mpad = Mpad;
  ba.pieces = pieces;
  ba.f16 = d->precision == kPrecFP16x3 ? 1 : 0;
  for (int q = 0; q < kOlsMaxK; ++q) ba.zscale[q] = md->zscale[q];
  const size_t plane_a = (size_t)d->Kp * Mpad;
  if (pieces == 0) {
    ba.a32 = reinterpret_cast<float*>(base);
  } else {
    for (int p = 0; p < pieces; ++p) ba.abf[p] = reinterpret_cast<uint16_t*>(base) + p * plane_a;
  }
  hipError_t e = launch_ols_build_a(ba, stream);
  if (e != hipSuccess) return set_error(-(int)e, "ols build-A launch");
  // per-permutation Gram factors of a model whose fixed columns are not all
  // constant, in the workspace after A
  char* fb = base + ((abytes + 255) / 256 * 256);
  auto factors = [&](const tfce_ols_model_t* mm, const float** Linv, const double** dc, int nb = 0,
                     const float** Rb = nullptr) -> int {
    *Linv = mm->Linv1;
    *dc = mm->dc1;
    if (mm->shared && nb == 0) return 0;
    const int Pf = mm->shared ? 1 : P;
    float* Lw = reinterpret_cast<float*>(fb);
    double* dw = reinterpret_cast<double*>(fb + (size_t)Pf * 64 * sizeof(float));
    fb += ((size_t)Pf * (64 * sizeof(float) + kOlsMaxContrasts * sizeof(double)) + 255) / 256 * 256;
    float* Rw = nullptr;
    if (nb > 0) {
      Rw = reinterpret_cast<float*>(fb);
      fb += ((size_t)Pf * nb * 64 * sizeof(float) + 255) / 256 * 256;
    }
    OlsFactorArgs ka{};
    ka.F = mm->F;
    ka.Z = mm->Z;
    ka.FF = mm->FF;
    ka.ZZ = mm->ZZ;
    ka.perm = mm->shared ? nullptr : d_perm;
    ka.n = d->n;
    ka.f = mm->f;
    ka.m = mm->m;
    ka.P = Pf;
    ka.n_con = mm->n_con;
    for (int c = 0; c < kOlsMaxContrasts; ++c) ka.con[c] = mm->con[c];
    ka.Linv = Lw;
    ka.dc = dw;
    ka.n_blk = nb;
    for (int b = 0; b < nb; ++b) ka.blk_mask[b] = blk_mask[b];
    ka.Rb = Rw;
    hipError_t ef = launch_ols_factors(ka, stream);
    if (ef != hipSuccess) return set_error(-(int)ef, "ols factor launch");
    *Linv = Lw;
    *dc = dw;
    if (Rb) *Rb = Rw;
    return 0;
  };
  const float* Linv = nullptr;
  const double* dc = nullptr;
  const float* Rb = nullptr;
  if ((rc = factors(md, &Linv, &dc, n_blk, &Rb))) return rc;
  OlsGemmArgs ga{};
  ga.a32 = ba.a32;
  for (int p = 0; p < 3; ++p) {
    ga.abf[p] = ba.abf[p];
    ga.bbf[p] = d->bf[p];
  }
  ga.b32 = d->plane32;
  ga.Kp = d->Kp;
  ga.Mpad = Mpad;
  ga.Npad = d->Npad;
  ga.V = d->V;
  ga.rpp = rpp;
  ga.P = P;
  ga.n = d->n;
  ga.k = md->k;
  ga.f = md->f;
  ga.m = md->m;
  ga.inv_dof = 1.0 / (double)(d->n - md->k);
  ga.n_con = md->n_con;
  for (int c = 0; c < kOlsMaxContrasts; ++c) {
    ga.con[c] = md->con[c];
    ga.con_is_intercept[c] = md->con_is_intercept[c];
  }
  ga.shared_factor = md->shared ? 1 : 0;
  ga.Linv = Linv;
  ga.dc = dc;
  ga.gF = md->gF;
  ga.hF = md->hF;
  ga.r0 = md->r0;
  ga.mu = d->mu;
  ga.centered = d->center;
  ga.colsumZ = md->colsumZ;
  ga.t = d_t;
  ga.ldt = ldt;
  ga.yinv = d->yinv;
  for (int q = 0; q < kOlsMaxK; ++q) ga.zinv[q] = 1.0f / md->zscale[q];
  ga.n_mtiles = (int32_t)n_mtiles;
  ga.n_ntiles = (int32_t)(d->Npad / tile_cols);
  ga.tss = tss;
  if (n_blk > 0) {
    ga.n_blk = n_blk;
    ga.blk_shared = md->shared ? 1 : 0;
    for (int b = 0; b < n_blk; ++b) ga.blk_w[b] = __builtin_popcount(blk_mask[b]);
    ga.Rb = Rb;
  }
  if (md2) {
    const float* Linv2 = nullptr;
    const double* dc2 = nullptr;
    if ((rc = factors(md2, &Linv2, &dc2))) return rc;
    ga.pair = 1;
    ga.alg = alg;
    ga.f2 = md2->f;
    ga.k2 = md2->k;
    ga.con2 = md2->con[0];
    ga.con2_is_intercept = md2->con_is_intercept[0];
    ga.shared2 = md2->shared ? 1 : 0;
    ga.Linv2 = Linv2;
    ga.dc2 = dc2;
    ga.gF2 = md2->gF;
    ga.hF2 = md2->hF;
    ga.r02 = md2->r0;
  }
  e = launch_ols_gemm(ga, d->precision, tile, stream);
  if (e != hipSuccess) return set_error(-(int)e, "ols gemm launch");
  return 0;
}

int tfce_ols_perm_t(const tfce_ols_model_t* md, const int32_t* d_perm, int32_t P, float* d_t, int64_t ldt,
                    hipStream_t stream) {
  return perm_run(md, d_perm, P, d_t, ldt, nullptr, stream);
}

int tfce_ols_perm_sobel(const tfce_ols_model_t* path_a, const tfce_ols_model_t* path_b, const int32_t* d_perm,
                        int32_t P, int32_t alg, float* d_z, int64_t ldz, hipStream_t stream) {
  if (!path_a || !path_b || !d_z) return set_error(TFCE_EINVAL, "NULL argument");
  if (alg < 0 || alg > 2) return set_error(TFCE_EINVAL, "alg must be 0 (aroian), 1 (sobel) or 2 (goodman)");
  if (path_a->data != path_b->data) return set_error(TFCE_EINVAL, "the two models must share one data handle");
  if (path_a->m != 1 || path_b->m != 1)
    return set_error(TFCE_EINVAL, "each model must permute exactly one column");
  if (path_a->n_con != 1 || path_b->n_con != 1) return set_error(TFCE_EINVAL, "each model must have one contrast");
  if (path_a->hZ != path_b->hZ)
    return set_error(TFCE_EINVAL, "the two models must permute the same column values");
  return perm_run(path_a, d_perm, P, d_z, ldz, nullptr, stream, path_b, alg);
}

int tfce_ols_perm_partial_f(const tfce_ols_model_t* md, const int32_t* d_perm, int32_t P,
                            const int32_t* block_start, const int32_t* block_width, int32_t n_blocks, float* d_f,
                            int64_t ldf, hipStream_t stream) {
  if (!md || !block_start || !block_width || !d_f) return set_error(TFCE_EINVAL, "NULL argument");
  if (n_blocks < 1 || n_blocks > kOlsMaxK) return set_error(TFCE_EINVAL, "1..8 column blocks");
  uint32_t masks[kOlsMaxK] = {0};
  uint32_t used = 0;
  for (int b = 0; b < n_blocks; ++b) {
    const int s0 = block_start[b], w = block_width[b];
    if (w < 1 || s0 < 0 || s0 + w > md->k) return set_error(TFCE_EINVAL, "column block out of range");
    if (w >= md->k) return set_error(TFCE_EINVAL, "a column block must leave at least one column");
    for (int j = s0; j < s0 + w; ++j) {
      if (md->constant[j] && md->data->center)
        return set_error(TFCE_ENOTSUP, "a column block holds the constant column of centred data");
      masks[b] |= 1u << md->internal[j];
    }
    if (used & masks[b]) return set_error(TFCE_EINVAL, "column blocks overlap");
    used |= masks[b];
  }
  return perm_run(md, d_perm, P, d_f, ldf, nullptr, stream, nullptr, 0, n_blocks, masks);
}

int tfce_ols_perm_f(const tfce_ols_model_t* md, const int32_t* d_perm, int32_t P, const double* d_tss,
                    float* d_f, int64_t ldf, hipStream_t stream) {
  if (!md || !d_tss) return set_error(TFCE_EINVAL, "NULL argument");
  if (md->k < 2) return set_error(TFCE_EINVAL, "the F statistic needs k >= 2 design columns");
  return perm_run(md, d_perm, P, d_f, ldf, d_tss, stream);
}

int tfce_ols_direct(const float* d_Y, int32_t n, int64_t V, int64_t ldy, const double* X, int32_t k,
                    const double* invXX_se, double* d_coef, float* d_se, double* d_t, double* d_resid,
                    double* d_sse, double* d_sumy, double* d_sumyy, hipStream_t stream) {
  if (!d_Y || !X) return set_error(TFCE_EINVAL, "NULL argument");
  if (k < 1 || k > kOlsMaxK || n <= k || n > kOlsDirectMaxN || V < 1 || ldy < V)
    return set_error(TFCE_EINVAL, "need 1 <= k <= 8, k < n <= 2048, V >= 1, ldy >= V");
  // inv(X'X) for the coefficients, as cy_lin_lstsqr_mat does (cynumstats.pyx:28-29)
  std::vector<double> G(k * k, 0.0), Li, Ainv(k * k, 0.0);
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += X[i * k + a] * X[i * k + b];
      G[a * k + b] = s;
    }
  if (!chol_inv(G, k, Li)) return set_error(TFCE_EINVAL, "X'X is singular");
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) {
      double s = 0.0;
      for (int q = std::max(a, b); q < k; ++q) s += Li[q * k + a] * Li[q * k + b];
      Ainv[a * k + b] = s;
    }
  double *dX = nullptr, *dA = nullptr, *dS = nullptr;
  int rc = 0;
  if ((rc = dalloc(&dX, (size_t)n * k, "X")) || (rc = dalloc(&dA, (size_t)k * k, "Ainv")) ||
      (rc = dalloc(&dS, (size_t)k * k, "Sinv")) || (rc = upload(dX, X, (size_t)n * k, "X")) ||
      (rc = upload(dA, Ainv.data(), (size_t)k * k, "Ainv")) ||
      (rc = upload(dS, invXX_se ? invXX_se : Ainv.data(), (size_t)k * k, "Sinv"))) {
    for (void* p : {(void*)dX, (void*)dA, (void*)dS}) if (p) (void)hipFree(p);
    return rc;
  }
  OlsDirectArgs a{};
  a.Y = d_Y;
  a.n = n;
  a.k = k;
  a.V = V;
  a.ldy = ldy;
  a.X = dX;
  a.Ainv = dA;
  a.Sinv = dS;
  a.coef = d_coef;
  a.se = d_se;
  a.t = d_t;
  a.resid = d_resid;
  a.sse = d_sse;
  a.sumy = d_sumy;
  a.sumyy = d_sumyy;
  hipError_t e = launch_ols_direct(a, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  for (void* p : {(void*)dX, (void*)dA, (void*)dS}) (void)hipFree(p);
  if (e != hipSuccess) return set_error(-(int)e, "ols direct kernel");
  return 0;
}

// Host FP64 inverse of D'D for a row-major n x m design D; false if singular.
static bool gram_inverse(const double* D, int n, int m, std::vector<double>& inv) {
  std::vector<double> G(m * m, 0.0), Li;
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += D[i * m + a] * D[i * m + b];
      G[a * m + b] = s;
    }
  if (!chol_inv(G, m, Li)) return false;
  inv.assign(m * m, 0.0);
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      double s = 0.0;
      for (int q = std::max(a, b); q < m; ++q) s += Li[q * m + a] * Li[q * m + b];
      inv[a * m + b] = s;
    }
  return true;
}

int tfce_image_regression(const float* d_Y, int64_t ldy, const float* d_Z, int64_t ldz, const double* R,
                          int32_t n, int32_t r, int64_t V, double std_min, float* d_t, int64_t ldt,
                          hipStream_t stream) {
  if (!d_Y || !d_Z || !R || !d_t) return set_error(TFCE_EINVAL, "NULL argument");
  if (r < 1 || r + 1 > kOlsMaxK || n <= r + 1 || n > kOlsDirectMaxN || V < 1 || ldy < V || ldz < V || ldt < V)
    return set_error(TFCE_EINVAL, "need 1 <= r <= 7, r + 1 < n <= 2048, V >= 1 and strides >= V");
  std::vector<double> Ainv;
  if (!gram_inverse(R, n, r, Ainv)) return set_error(TFCE_EINVAL, "R'R is singular");
  double *dR = nullptr, *dA = nullptr;
  int rc = 0;
  if ((rc = dalloc(&dR, (size_t)n * r, "R")) || (rc = dalloc(&dA, (size_t)r * r, "Ainv")) ||
      (rc = upload(dR, R, (size_t)n * r, "R")) || (rc = upload(dA, Ainv.data(), (size_t)r * r, "Ainv"))) {
    for (void* p : {(void*)dR, (void*)dA}) if (p) (void)hipFree(p);
    return rc;
  }
  ImgRegArgs a{};
  a.Y = d_Y;
  a.Z = d_Z;
  a.ldy = ldy;
  a.ldz = ldz;
  a.n = n;
  a.r = r;
  a.V = V;
  a.R = dR;
  a.Ainv = dA;
  a.std_min = std_min;
  a.T = d_t;
  a.ldt = ldt;
  hipError_t e = launch_image_regression(a, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  for (void* p : {(void*)dR, (void*)dA}) (void)hipFree(p);
  if (e != hipSuccess) return set_error(-(int)e, "image regression kernel");
  return 0;
}

int tfce_image_vif(const void* d_Y, int32_t y_f64, int64_t ldy, const double* X, int32_t n, int32_t k, int64_t V,
                   double* d_vif, hipStream_t stream) {
  if (!d_Y || !X || !d_vif) return set_error(TFCE_EINVAL, "NULL argument");
  if (k < 1 || k > kOlsMaxK || n <= k || n > kOlsDirectMaxN || V < 1 || ldy < V)
    return set_error(TFCE_EINVAL, "need 1 <= k <= 8, k < n <= 2048, V >= 1 and ldy >= V");
  std::vector<double> Ainv;
  if (!gram_inverse(X, n, k, Ainv)) return set_error(TFCE_EINVAL, "X'X is singular");
  double *dX = nullptr, *dA = nullptr;
  int rc = 0;
  if ((rc = dalloc(&dX, (size_t)n * k, "X")) || (rc = dalloc(&dA, (size_t)k * k, "Ainv")) ||
      (rc = upload(dX, X, (size_t)n * k, "X")) || (rc = upload(dA, Ainv.data(), (size_t)k * k, "Ainv"))) {
    for (void* p : {(void*)dX, (void*)dA}) if (p) (void)hipFree(p);
    return rc;
  }
  ImgVifArgs a{};
  a.Y = d_Y;
  a.ldy = ldy;
  a.n = n;
  a.k = k;
  a.V = V;
  a.X = dX;
  a.Ainv = dA;
  a.vif = d_vif;
  hipError_t e = launch_image_vif(a, y_f64 != 0, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  for (void* p : {(void*)dX, (void*)dA}) (void)hipFree(p);
  if (e != hipSuccess) return set_error(-(int)e, "image VIF kernel");
  return 0;
}

int tfce_se_of_slope(const double* d_sigma2, const double* invXX, int32_t k, int64_t V, float* d_se,
                     hipStream_t stream) {
  if (!d_sigma2 || !invXX || !d_se) return set_error(TFCE_EINVAL, "NULL argument");
  if (k < 1 || k > kOlsMaxK) return set_error(TFCE_EINVAL, "k must be in [1, 8]");
  double* dd = nullptr;
  std::vector<double> diag(k);
  for (int j = 0; j < k; ++j) diag[j] = invXX[j * k + j];
  int rc = dalloc(&dd, (size_t)k, "diag");
  if (!rc) rc = upload(dd, diag.data(), (size_t)k, "diag");
  if (rc) { if (dd) (void)hipFree(dd); return rc; }
  hipError_t e = launch_se_of_slope(d_sigma2, dd, k, V, d_se, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(dd);
  if (e != hipSuccess) return set_error(-(int)e, "se_of_slope kernel");
  return 0;
}

int tfce_total_ss(const float* d_Y, int32_t n, int64_t V, int64_t ldy, double* d_tss, hipStream_t stream) {
  if (!d_Y || !d_tss) return set_error(TFCE_EINVAL, "NULL argument");
  if (n < 1 || V < 1 || ldy < V) return set_error(TFCE_EINVAL, "need n >= 1, V >= 1 and ldy >= V");
  double* dm = nullptr;
  int rc = dalloc(&dm, 1, "grand mean");
  if (rc) return rc;
  hipError_t e = launch_total_ss(d_Y, n, V, ldy, d_tss, dm, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(dm);
  if (e != hipSuccess) return set_error(-(int)e, "total_ss kernels");
  return 0;
}

int tfce_sobel_z(const float* d_ta, const float* d_tb, int64_t n, int32_t alg, float ta_const, float* d_z,
                 hipStream_t stream) {
  if (!d_tb || !d_z) return set_error(TFCE_EINVAL, "NULL argument");
  if (alg < 0 || alg > 2) return set_error(TFCE_EINVAL, "alg must be 0 (aroian), 1 (sobel) or 2 (goodman)");
  hipError_t e = launch_sobel(d_ta, d_tb, n, alg, ta_const, d_z, stream);
  if (e != hipSuccess) return set_error(-(int)e, "sobel launch");
  return 0;
}

int tfce_mlp_param_count(void) { return kMlpParams; }

int tfce_mlp_forward(const float* d_params, const float* d_x, int32_t n, float* d_y, hipStream_t stream) {
  if (!d_params || !d_x || !d_y) return set_error(TFCE_EINVAL, "NULL argument");
  hipError_t e = launch_mlp_forward(d_params, d_x, n, d_y, stream);
  if (e != hipSuccess) return set_error(-(int)e, "mlp forward launch");
  return 0;
}

int tfce_mlp_train(float* d_params, float* d_adam_m, float* d_adam_v, const float* d_x, const float* d_y,
                   int32_t n, const int32_t* d_batch_idx, int32_t steps, int32_t batch, int32_t t0, float lr,
                   float weight_decay, float* d_loss, float* d_grad0, hipStream_t stream) {
  if (!d_params || !d_adam_m || !d_adam_v || !d_x || !d_y || !d_batch_idx)
    return set_error(TFCE_EINVAL, "NULL argument");
  if (batch < 1 || batch > kMlpBatch) return set_error(TFCE_EINVAL, "batch must be in [1, 128]");
  if (steps < 1) return set_error(TFCE_EINVAL, "steps must be >= 1");
  if (n < 1) return set_error(TFCE_EINVAL, "n must be >= 1");
  MlpTrainArgs a{};
  a.params = d_params;
  a.adam_m = d_adam_m;
  a.adam_v = d_adam_v;
  a.X = d_x;
  a.Y = d_y;
  a.batch_idx = d_batch_idx;
  a.n = n;
  a.steps = steps;
  a.batch = batch;
  a.t0 = t0;
  a.lr = lr;
  a.beta1 = 0.9f;
  a.beta2 = 0.999f;
  a.eps = 1e-8f;
  a.weight_decay = weight_decay;
  a.loss = d_loss;
  a.grad0 = d_grad0;
  hipError_t e = launch_mlp_train(a, stream);
  if (e != hipSuccess) return set_error(-(int)e, "mlp train launch");
  return 0;
}

int tfce_ols_info(const tfce_ols_model_t* md, int32_t* k, int32_t* n_fixed, int32_t* shared_gram) {
  if (!md) return set_error(TFCE_EINVAL, "NULL model");
  if (k) *k = md->k;
  if (n_fixed) *n_fixed = md->f;
  if (shared_gram) *shared_gram = md->shared ? 1 : 0;
  return 0;
}

}  // extern "C"
