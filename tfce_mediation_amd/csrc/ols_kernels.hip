// Permuted mass-univariate OLS t-values on CDNA4 (gfx950).
//
// What is computed (the reference's tval_int, cynumstats.pyx:59-64, applied to
// X[perm] as in vertex_tfce_multiple_regression_randomise.py:105-107):
//   a = (X'X)^-1 X'y,  sigma^2 = |y - X a|^2 / (n - k),
//   se_c = fl32(sqrt(sigma^2 [(X'X)^-1]_cc))  (se is stored as float32, :49),
//   t_c = a_c / se_c.
// How (my design): the design columns split into fixed ones F (the intercept,
// never moved by a permutation) and permuted ones Z.  For a batch of P
// permutations the only large product is Z_p' Y for every p at once - one dense
// contraction [P*rpp x n] . [n x V] on MFMA, rpp = one row per permuted column.
// With L the Cholesky factor of the permuted Gram matrix (X_p'X_p = L L'),
// h = L^-1 X_p'y gives |y - X a|^2 = y'y - |h|^2 and a = L^-T h, so the
// epilogue needs only h_Z = Linv_ZF g_F + Linv_ZZ g_Z per (permutation, element):
//   SSE = r0 - |h_Z|^2 with r0 = y'y - |Linv_FF g_F|^2 precomputed in FP64.
// Y is centred per element first (exact for every coefficient once the model
// holds an intercept; the intercept's own coefficient gets the mean back), so
// the bf16 split loses precision relative to the spread of y, not its offset.
#include <type_traits>

#include "ols_internal.h"

namespace tfcehip {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Split x into `pieces` bf16 terms, each the rounding of what is left.
__device__ __forceinline__ void split_bf16(float x, int pieces, uint16_t out[3]) {
  float r = x;
  for (int p = 0; p < 3; ++p) {
    if (p < pieces) {
      out[p] = bf16_rne(r);
      r = r - bf16_f32(out[p]);  // exact
    } else {
      out[p] = 0;
    }
  }
}

// Split a (scaled) value into fp16 hi and lo: hi = fl16(x), lo = fl16(x - hi);
// x - hi is exact in float, so hi + lo carries about 22 significant bits.
__device__ __forceinline__ void split_f16(float x, uint16_t out[3]) {
  const _Float16 hi = (_Float16)x;
  const _Float16 lo = (_Float16)(x - (float)hi);
  out[0] = __builtin_bit_cast(uint16_t, hi);
  out[1] = __builtin_bit_cast(uint16_t, lo);
  out[2] = 0;
}

// Power of two s with max_abs * s in [2^13, 2^14): fp16 operands well inside
// the normal range; exact to undo.
__device__ __forceinline__ float f16_scale(float max_abs) {
  if (!(max_abs > 0.0f) || !(max_abs < INFINITY)) return 1.0f;
  return ldexpf(1.0f, 13 - ilogbf(max_abs));
}

__global__ __launch_bounds__(256) void ols_prep_kernel(OlsPrepArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.Npad) return;
  const bool valid = v < a.V;
  double mu = 0.0;
  if (valid && a.center) {
    for (int i = 0; i < a.n; ++i) mu += (double)a.Y[(int64_t)i * a.ldy + v];
    mu /= (double)a.n;
  }
  float ys = 1.0f;
  if (a.f16) {
    float mx = 0.0f;
    if (valid)
      for (int i = 0; i < a.n; ++i) mx = fmaxf(mx, fabsf((float)((double)a.Y[(int64_t)i * a.ldy + v] - mu)));
    ys = f16_scale(mx);
    if (valid) a.yinv[v] = 1.0f / ys;
  }
  double yy = 0.0;
  const int nkg = a.Kp / 8;
  for (int kg = 0; kg < nkg; ++kg) {
    float vals[8];
    uint16_t pc[3][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = kg * 8 + j;
      float y = 0.0f;
      if (valid && i < a.n) y = (float)((double)a.Y[(int64_t)i * a.ldy + v] - mu);
      vals[j] = y;
      yy += (double)y * (double)y;
      uint16_t s[3];
      if (a.f16)
        split_f16(y * ys, s);
      else
        split_bf16(y, a.pieces, s);
      pc[0][j] = s[0];
      pc[1][j] = s[1];
      pc[2][j] = s[2];
    }
    const int64_t base = ((int64_t)kg * a.Npad + v) * 8;
    float4* dst = reinterpret_cast<float4*>(a.plane32 + base);
    dst[0] = make_float4(vals[0], vals[1], vals[2], vals[3]);
    dst[1] = make_float4(vals[4], vals[5], vals[6], vals[7]);
    for (int p = 0; p < a.pieces; ++p) {
      uint4 w;
      w.x = (uint32_t)pc[p][0] | ((uint32_t)pc[p][1] << 16);
      w.y = (uint32_t)pc[p][2] | ((uint32_t)pc[p][3] << 16);
      w.z = (uint32_t)pc[p][4] | ((uint32_t)pc[p][5] << 16);
      w.w = (uint32_t)pc[p][6] | ((uint32_t)pc[p][7] << 16);
      *reinterpret_cast<uint4*>(a.bf[p] + base) = w;
    }
  }
  if (valid) {
    a.mu[v] = a.center ? mu : 0.0;
    a.yy[v] = yy;
  }
}

__global__ __launch_bounds__(256) void ols_fixed_kernel(OlsFixedArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.V) return;
  double g[kOlsMaxK];
#pragma unroll
  for (int q = 0; q < kOlsMaxK; ++q) g[q] = 0.0;
  for (int i = 0; i < a.n; ++i) {
    const double y = (double)a.plane32[((int64_t)(i >> 3) * a.Npad + v) * 8 + (i & 7)];
#pragma unroll
    for (int q = 0; q < kOlsMaxK; ++q)
      if (q < a.f) g[q] += a.F[i * a.f + q] * y;
  }
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < kOlsMaxK; ++q) {
    if (q < a.f) {
      double hq = 0.0;
      for (int j = 0; j <= q; ++j) hq += a.LFF[q * a.f + j] * g[j];
      ss += hq * hq;
      a.gF[(int64_t)q * a.V + v] = (float)g[q];
      a.hF[(int64_t)q * a.V + v] = (float)hq;
    }
  }
  a.r0[v] = a.yy[v] - ss;
}

// One workgroup per permutation: Gram matrix of X_p = [F, Z[perm]], its
// Cholesky factor and the inverse factor (FP64), stored as float for the
// epilogue, plus [(X'X)^-1]_cc per contrast in FP64.
__global__ __launch_bounds__(64) void ols_factor_kernel(OlsFactorArgs a) {
  __shared__ double C[kOlsMaxK * kOlsMaxK];
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  const int32_t* perm = a.perm ? a.perm + (int64_t)p * a.n : nullptr;
  if (tid < a.f * a.m) {
    const int q = tid / a.m, b = tid % a.m;
    double s = 0.0;
    for (int i = 0; i < a.n; ++i) {
      int src = perm ? perm[i] : i;
      if (src < 0 || src >= a.n) { s = NAN; break; }
      s += a.F[i * a.f + q] * a.Z[src * a.m + b];
    }
    C[q * kOlsMaxK + b] = s;
  }
  __syncthreads();
  if (tid != 0) return;
  const int k = a.f + a.m;
  double G[kOlsMaxK][kOlsMaxK], L[kOlsMaxK][kOlsMaxK], Li[kOlsMaxK][kOlsMaxK];
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      double g;
      if (i < a.f && j < a.f) g = a.FF[i * a.f + j];
      else if (i < a.f) g = C[i * kOlsMaxK + (j - a.f)];
      else if (j < a.f) g = C[j * kOlsMaxK + (i - a.f)];
      else g = a.ZZ[(i - a.f) * a.m + (j - a.f)];
      G[i][j] = g;
      L[i][j] = 0.0;
      Li[i][j] = 0.0;
    }
  for (int j = 0; j < k; ++j) {
    double d = G[j][j];
    for (int q = 0; q < j; ++q) d -= L[j][q] * L[j][q];
    d = d > 0.0 ? sqrt(d) : NAN;
    L[j][j] = d;
    for (int i = j + 1; i < k; ++i) {
      double s = G[i][j];
      for (int q = 0; q < j; ++q) s -= L[i][q] * L[j][q];
      L[i][j] = s / d;
    }
  }
  for (int j = 0; j < k; ++j) {
    Li[j][j] = 1.0 / L[j][j];
    for (int i = j + 1; i < k; ++i) {
      double s = 0.0;
      for (int q = j; q < i; ++q) s += L[i][q] * Li[q][j];
      Li[i][j] = -s / L[i][i];
    }
  }
  float* out = a.Linv + (int64_t)p * 64;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) out[i * 8 + j] = (i < k && j < k) ? (float)Li[i][j] : 0.0f;
  for (int c = 0; c < kOlsMaxContrasts; ++c) {
    double s = 0.0;
    if (c < a.n_con)
      for (int j = a.con[c]; j < k; ++j) s += Li[j][a.con[c]] * Li[j][a.con[c]];
    a.dc[(int64_t)p * kOlsMaxContrasts + c] = s;
  }
  // Block b with columns B: beta = Li' h, and the extra sum of squares of B
  // given the other columns is beta_B' [(X'X)^-1_BB]^-1 beta_B = |C^-1 beta_B|^2
  // with C C' = (X'X)^-1_BB = (Li' Li)_BB; row r of Rb is sum_s C^-1[r][s]
  // Li[.][B_s], so |Rb h|^2 is that sum of squares.
  for (int b = 0; b < a.n_blk && a.Rb; ++b) {
    int B[kOlsMaxK], w = 0;
    for (int j = 0; j < k; ++j)
      if ((a.blk_mask[b] >> j) & 1u) B[w++] = j;
    double S[kOlsMaxK][kOlsMaxK], C[kOlsMaxK][kOlsMaxK], Ci[kOlsMaxK][kOlsMaxK];
    for (int r = 0; r < w; ++r)
      for (int q = 0; q < w; ++q) {
        double s = 0.0;
        for (int j = (B[r] > B[q] ? B[r] : B[q]); j < k; ++j) s += Li[j][B[r]] * Li[j][B[q]];
        S[r][q] = s;
        C[r][q] = 0.0;
        Ci[r][q] = 0.0;
      }
    for (int j = 0; j < w; ++j) {
      double d = S[j][j];
      for (int q = 0; q < j; ++q) d -= C[j][q] * C[j][q];
      d = d > 0.0 ? sqrt(d) : NAN;
      C[j][j] = d;
      for (int i = j + 1; i < w; ++i) {
        double s = S[i][j];
        for (int q = 0; q < j; ++q) s -= C[i][q] * C[j][q];
        C[i][j] = s / d;
      }
    }
    for (int j = 0; j < w; ++j) {
      Ci[j][j] = 1.0 / C[j][j];
      for (int i = j + 1; i < w; ++i) {
        double s = 0.0;
        for (int q = j; q < i; ++q) s += C[i][q] * Ci[q][j];
        Ci[i][j] = -s / C[i][i];
      }
    }
    float* R = a.Rb + ((int64_t)p * a.n_blk + b) * 64;
    for (int r = 0; r < 8; ++r)
      for (int j = 0; j < 8; ++j) {
        double s = 0.0;
        if (r < w && j < k)
          for (int q = 0; q <= r; ++q) s += Ci[r][q] * Li[j][B[q]];
        R[r * 8 + j] = (float)s;
      }
  }
}

__global__ __launch_bounds__(256) void ols_build_a_kernel(OlsBuildAArgs a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nkg = a.Kp / 8;
  if (idx >= (int64_t)nkg * a.Mpad) return;
  const int kg = (int)(idx / a.Mpad);
  const int64_t row = idx - (int64_t)kg * a.Mpad;
  const int per_tile = a.tile_rows / a.rpp;
  const int64_t rin = row % a.tile_rows;
  const int p = (int)((row / a.tile_rows) * per_tile + rin / a.rpp), j = (int)(rin % a.rpp);
  const bool used = rin < (int64_t)per_tile * a.rpp;
  float vals[8];
  uint16_t pc[3][8];
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const int i = kg * 8 + jj;
    float z = 0.0f;
    if (used && p < a.P && j < a.m && i < a.n) {
      const int src = a.perm ? a.perm[(int64_t)p * a.n + i] : i;
      z = (src >= 0 && src < a.n) ? a.Z32[src * a.m + j] : NAN;
    }
    vals[jj] = z;
    uint16_t s[3];
    if (a.f16)
      split_f16(j < 8 ? z * a.zscale[j] : 0.0f, s);
    else
      split_bf16(z, a.pieces, s);
    pc[0][jj] = s[0];
    pc[1][jj] = s[1];
    pc[2][jj] = s[2];
  }
  const int64_t base = ((int64_t)kg * a.Mpad + row) * 8;
  if (a.pieces == 0) {
    float4* dst = reinterpret_cast<float4*>(a.a32 + base);
    dst[0] = make_float4(vals[0], vals[1], vals[2], vals[3]);
    dst[1] = make_float4(vals[4], vals[5], vals[6], vals[7]);
  } else {
    for (int q = 0; q < a.pieces; ++q) {
      uint4 w;
      w.x = (uint32_t)pc[q][0] | ((uint32_t)pc[q][1] << 16);
      w.y = (uint32_t)pc[q][2] | ((uint32_t)pc[q][3] << 16);
      w.z = (uint32_t)pc[q][4] | ((uint32_t)pc[q][5] << 16);
      w.w = (uint32_t)pc[q][6] | ((uint32_t)pc[q][7] << 16);
      *reinterpret_cast<uint4*>(a.abf[q] + base) = w;
    }
  }
}

__device__ __forceinline__ bf16x8 ld_bf16x8(const uint16_t* p) {
  uint4 u = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, u);
}

// Epilogue for one (permutation, element): gz holds X_p'y for the permuted
// columns (FP32 accumulators), everything else comes from per-element and
// per-permutation tables.
// UNIFORM: the caller guarantees p is the same in every lane of the wave, so
// the permutation's factor is read with scalar loads.  SSE, se and t follow
// tval_int: se = fl32(sqrt(SSE / (n - k) * [(X'X)^-1]_cc)) in FP64
// (cynumstats.pyx:47-64), t = a / se in FP64, returned as float.
// FF / MM > 0: the fixed / permuted column counts as compile-time constants
// (every loop unrolls with constant bounds); 0: read them from the arguments.
// BLK: write the partial F of the column blocks instead (a kernel instance of
// its own, so the t epilogue's register budget does not carry it).
template <bool UNIFORM, int FF, int MM, bool BLK>
__device__ __forceinline__ void ols_epilogue(const OlsGemmArgs& a, int p, int64_t col, const float* gz,
                                             const float* gF, const float* hF, double r0, double mu) {
  const int f = FF > 0 ? FF : a.f, m = MM > 0 ? MM : a.m, k = f + m;
  const int fi = a.shared_factor ? 0 : (UNIFORM ? __builtin_amdgcn_readfirstlane(p) : p);
  const float* L = a.Linv + (int64_t)fi * 64;
  float hz[kOlsMaxK];
  double ssz = 0.0;
#pragma unroll
  for (int i = 0; i < kOlsMaxK; ++i) {
    hz[i] = 0.0f;
    if (i < m) {
      const int row = f + i;
      float s = 0.0f;
#pragma unroll
      for (int j = 0; j < kOlsMaxK; ++j)
        if (j < f) s = fmaf(L[row * 8 + j], gF[j], s);
#pragma unroll
      for (int j = 0; j < kOlsMaxK; ++j)
        if (j <= i) s = fmaf(L[row * 8 + f + j], gz[j], s);
      hz[i] = s;
      ssz += (double)s * (double)s;
    }
  }
  double sse = r0 - ssz;
  if (sse < 0.0) sse = 0.0;
  if constexpr (BLK) {
    // partial F of each column block: |Rb h|^2 / w over SSE / (n - k)
    // (pyfunc.py:2340-2352 computes the same extra sum of squares as the
    // difference of the reduced and the full model's residual sums)
    const float* R = a.Rb + (int64_t)(a.blk_shared ? 0 : fi) * a.n_blk * 64;
    const double s2 = sse / (double)(a.n - k);
#pragma unroll 1
    for (int b = 0; b < a.n_blk; ++b) {
      double ex = 0.0;
#pragma unroll 1
      for (int r = 0; r < a.blk_w[b]; ++r) {
        const float* Rr = R + b * 64 + r * 8;
        float u = 0.0f;
#pragma unroll
        for (int j = 0; j < kOlsMaxK; ++j)
          if (j < k) u = fmaf(Rr[j], (j < f) ? hF[j] : hz[(j - f) & (kOlsMaxK - 1)], u);
        ex += (double)u * (double)u;
      }
      a.t[((int64_t)b * a.P + p) * a.ldt + col] = (float)(ex / (double)a.blk_w[b] / s2);
    }
    return;
  }
  if (a.tss) {
    // F of the whole model against the caller's total sum of squares
    // (cynumstats.pyx:31-36), as sqrt(max(F, 0)) (voxel randomise :85-87)
    const double F = ((a.tss[col] - sse) / (double)(k - 1)) / (sse / (double)(a.n - k));
    a.t[(int64_t)p * a.ldt + col] = (float)(F < 0.0 ? 0.0 : sqrt(F));
    return;
  }
  const double s2 = sse / (double)(a.n - k);
  for (int c = 0; c < a.n_con; ++c) {
    const int ci = a.con[c];
    float ac = 0.0f;
#pragma unroll
    for (int j = 0; j < kOlsMaxK; ++j) {
      if (j >= ci && j < k) {
        const float hj = (j < f) ? hF[j < kOlsMaxK ? j : 0] : hz[(j - f) & (kOlsMaxK - 1)];
        ac = fmaf(L[j * 8 + ci], hj, ac);
      }
    }
    double av = (double)ac;
    if (a.centered && a.con_is_intercept[c]) av += mu;
    const float se = (float)sqrt(s2 * a.dc[(int64_t)fi * kOlsMaxContrasts + c]);
    a.t[((int64_t)c * a.P + p) * a.ldt + col] = (float)(av / (double)se);
  }
}

// Sobel-type z of the indirect effect from the two path t-values
// (pyfunc.py:150-161, calc_indirect :2678-2709): 0 aroian, 1 sobel, 2 goodman.
__device__ __forceinline__ float sobel_z(double a, double b, int alg) {
  const double ia = 1.0 / (a * a), ib = 1.0 / (b * b), iab = 1.0 / (a * a * b * b);
  double d;
  if (alg == 0) d = ib + ia + iab;
  else if (alg == 1) d = ib + ia;
  else d = ib + ia - iab;
  return (float)(1.0 / sqrt(d));
}

// t of contrast column ci of a model with f fixed columns (FF > 0: a
// compile-time f) and one permuted column whose X_p'y is gz0: the arithmetic
// of ols_epilogue for m = 1, operation for operation.
template <int FF>
__device__ __forceinline__ float single_t(int f_rt, int k, int n, const float* L, double dcv, int ci, int is_icpt,
                                          int centered, const float* gF, const float* hF, double r0, double mu,
                                          float gz0) {
  const int f = FF > 0 ? FF : f_rt;
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j)
    if (j < f) s = fmaf(L[f * 8 + j], gF[j], s);
  s = fmaf(L[f * 8 + f], gz0, s);
  const double ssz = 0.0 + (double)s * (double)s;
  double sse = r0 - ssz;
  if (sse < 0.0) sse = 0.0;
  const double s2 = sse / (double)(n - k);
  float ac = 0.0f;
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j) {
    if (j >= ci && j < k) {
      const float hj = (j < f) ? hF[j] : s;
      ac = fmaf(L[j * 8 + ci], hj, ac);
    }
  }
  double av = (double)ac;
  if (centered && is_icpt) av += mu;
  const float se = (float)sqrt(s2 * dcv);
  return (float)(av / (double)se);
}

// Epilogue through LDS.  The accumulators of a wave cover, for tile (rt, ct),
// rows row0 + 32 rt + 8 (i / 4) + 4 h + (i % 4) (i = accumulator index 0..15)
// of column 32 ct + r.  The workgroup writes one 32-column slice of every
// wave's tiles into an LDS block ep[256 rows][ldc], then each thread owns one
// column of the block and walks permutations: it reads the permutation's
// rpp rows (X_p'y for the permuted columns) and runs the per-element
// epilogue.  Unrolled register code stays small (only the LDS write is
// unrolled), so the accumulators never leave registers for scratch.
template <int RT, int CT>
__device__ __forceinline__ void epi_put(float* ep, int ldc, const f32x16 (&acc)[RT][CT], int sel, int row0,
                                        int c0, int h, int r) {
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    if (ct != sel) continue;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int i = 0; i < 16; ++i)
        ep[(row0 + 32 * rt + 8 * (i >> 2) + 4 * h + (i & 3)) * ldc + c0 + r] = acc[rt][ct][i];
  }
}

// Thread (cl, slot): column cl of the block (global element `col`),
// permutations slot, slot + step, ... of the workgroup's row tile.
template <int LDC, int FF, int MM, bool BLK = false>
__device__ __forceinline__ void epi_loop(const OlsGemmArgs& a, const float* ep, int cl, int64_t col, int p0, int ppt,
                                         int slot, int step, const float* gF, const float* hF, double r0, double mu,
                                         float yinv) {
  const int m = MM > 0 ? MM : a.m, rpp = MM > 0 ? MM : a.rpp;
  for (int pp = slot; pp < ppt; pp += step) {
    const int p = p0 + pp;
    if (p >= a.P) break;
    float gz[kOlsMaxK];
#pragma unroll
    for (int q = 0; q < kOlsMaxK; ++q) {
      gz[q] = q < m ? ep[(pp * rpp + q) * LDC + cl] : 0.0f;
      if (a.yinv) gz[q] = gz[q] * a.zinv[q] * yinv;  // powers of two: exact
      if (!a.centered && q < m) gz[q] += (float)(a.colsumZ[q] * mu);
    }
    // with 64 columns per block a wave's lanes share one permutation
    ols_epilogue<LDC == 64, FF, MM, BLK>(a, p, col, gz, gF, hF, r0, mu);
  }
}

// The paired mediation epilogue: thread (cl, slot) walks its permutations,
// computes t of both models from the one permuted column's X_p'y, and writes z.
template <int LDC, int F1, int F2>
__device__ __forceinline__ void epi_loop_pair(const OlsGemmArgs& a, const float* ep, int cl, int64_t col, int p0,
                                              int ppt, int slot, int step, double mu, float yinv) {
  float gF1[kOlsMaxK], hF1[kOlsMaxK], gF2[kOlsMaxK], hF2[kOlsMaxK];
#pragma unroll
  for (int q = 0; q < kOlsMaxK; ++q) {
    gF1[q] = q < a.f ? a.gF[(int64_t)q * a.V + col] : 0.0f;
    hF1[q] = q < a.f ? a.hF[(int64_t)q * a.V + col] : 0.0f;
    gF2[q] = q < a.f2 ? a.gF2[(int64_t)q * a.V + col] : 0.0f;
    hF2[q] = q < a.f2 ? a.hF2[(int64_t)q * a.V + col] : 0.0f;
  }
  const double r01 = a.r0[col], r02 = a.r02[col];
  for (int pp = slot; pp < ppt; pp += step) {
    const int p = p0 + pp;
    if (p >= a.P) break;
    float gz0 = ep[pp * LDC + cl];
    if (a.yinv) gz0 = gz0 * a.zinv[0] * yinv;  // powers of two: exact
    if (!a.centered) gz0 += (float)(a.colsumZ[0] * mu);
    const int fi1 = a.shared_factor ? 0 : p, fi2 = a.shared2 ? 0 : p;
    const float ta = single_t<F1>(a.f, a.k, a.n, a.Linv + (int64_t)fi1 * 64, a.dc[(int64_t)fi1 * kOlsMaxContrasts],
                                  a.con[0], a.con_is_intercept[0], a.centered, gF1, hF1, r01, mu, gz0);
    const float tb = single_t<F2>(a.f2, a.k2, a.n, a.Linv2 + (int64_t)fi2 * 64,
                                  a.dc2[(int64_t)fi2 * kOlsMaxContrasts], a.con2, a.con2_is_intercept, a.centered,
                                  gF2, hF2, r02, mu, gz0);
    a.t[(int64_t)p * a.ldt + col] = sobel_z((double)ta, (double)tb, a.alg);
  }
}

template <int LDC, bool BLK>
__device__ __forceinline__ void epi_run(const OlsGemmArgs& a, const float* ep, int cl, int64_t col, int64_t mrow0,
                                        int rows, int slot, int step) {
  const bool cvalid = col < a.V;
  if (!cvalid) return;
  float gF[kOlsMaxK], hF[kOlsMaxK];
#pragma unroll
  for (int q = 0; q < kOlsMaxK; ++q) {
    gF[q] = q < a.f ? a.gF[(int64_t)q * a.V + col] : 0.0f;
    hF[q] = q < a.f ? a.hF[(int64_t)q * a.V + col] : 0.0f;
  }
  const double r0 = a.r0[col], mu = a.mu[col];
  const float yinv = a.yinv ? a.yinv[col] : 1.0f;
  const int ppt = rows / a.rpp;
  const int p0 = (int)(mrow0 / rows) * ppt;
  if constexpr (BLK) {
    epi_loop<LDC, 0, 0, true>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv);
    return;
  }
  if (a.pair) {
    // mediation types M and I: path A [1, x], path B [1, y, x]
    if (a.f == 1 && a.f2 == 2) epi_loop_pair<LDC, 1, 2>(a, ep, cl, col, p0, ppt, slot, step, mu, yinv);
    else epi_loop_pair<LDC, 0, 0>(a, ep, cl, col, p0, ppt, slot, step, mu, yinv);
    return;
  }
  // the design shapes of the drivers (one fixed intercept column and up to
  // seven permuted ones, rpp = m; the mediation models' intercept and
  // outcome with one or two permuted columns) get unrolled copies; others
  // the general loop
  if (a.f == 2 && a.rpp == a.m) {
    switch (a.m) {
      case 1: epi_loop<LDC, 2, 1>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 2: epi_loop<LDC, 2, 2>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      default: break;
    }
  }
  if (a.f == 1 && a.rpp == a.m) {
    switch (a.m) {
      case 1: epi_loop<LDC, 1, 1>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 2: epi_loop<LDC, 1, 2>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 3: epi_loop<LDC, 1, 3>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 4: epi_loop<LDC, 1, 4>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 5: epi_loop<LDC, 1, 5>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 6: epi_loop<LDC, 1, 6>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      case 7: epi_loop<LDC, 1, 7>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv); return;
      default: break;
    }
  }
  epi_loop<LDC, 0, 0>(a, ep, cl, col, p0, ppt, slot, step, gF, hF, r0, mu, yinv);
}

// Block -> tile mapping shared by both contraction kernels: the n_mtiles row
// tiles of one column tile are consecutive blocks of one XCD (blockIdx mod 8),
// so that XCD's L2 holds the column tile of Y while all of them read it.
__device__ __forceinline__ bool gemm_tile(const OlsGemmArgs& a, int& mt, int& nt) {
  const int b = blockIdx.x, xcd = b & 7, jb = b >> 3;
  mt = jb % a.n_mtiles;
  nt = (jb / a.n_mtiles) * 8 + xcd;
  return nt < a.n_ntiles;
}

// Plain FP32 contraction (v_mfma_f32_32x32x2f32): each wave holds 64 rows x
// 128 columns and reads its fragments straight from global memory.
template <bool BLK>
__global__ __launch_bounds__(256) void ols_gemm_fp32_kernel(OlsGemmArgs a) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5, r = l & 31;
  int mt, nt;
  if (!gemm_tile(a, mt, nt)) return;
  const int64_t m0 = (int64_t)mt * kGemmRows + w * 64;
  const int64_t n0 = (int64_t)nt * kGemmCols;
  f32x16 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;
  const int nkg = a.Kp / 8;
  for (int kg0 = 0; kg0 < nkg; kg0 += 2) {
    const int kg = kg0 + h;
    float A[2][8], B[4][8];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      const float4* s = reinterpret_cast<const float4*>(a.a32 + ((int64_t)kg * a.Mpad + m0 + 32 * rt + r) * 8);
      float4 x0 = s[0], x1 = s[1];
      A[rt][0] = x0.x; A[rt][1] = x0.y; A[rt][2] = x0.z; A[rt][3] = x0.w;
      A[rt][4] = x1.x; A[rt][5] = x1.y; A[rt][6] = x1.z; A[rt][7] = x1.w;
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const float4* s = reinterpret_cast<const float4*>(a.b32 + ((int64_t)kg * a.Npad + n0 + 32 * ct + r) * 8);
      float4 x0 = s[0], x1 = s[1];
      B[ct][0] = x0.x; B[ct][1] = x0.y; B[ct][2] = x0.z; B[ct][3] = x0.w;
      B[ct][4] = x1.x; B[ct][5] = x1.y; B[ct][6] = x1.z; B[ct][7] = x1.w;
    }
#pragma unroll
    for (int jj = 0; jj < 8; ++jj)
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
          acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[rt][jj], B[ct][jj], acc[rt][ct], 0, 0, 0);
  }
  // epilogue: four passes of 32 columns through a [256][32] LDS block
  __shared__ float ep[kGemmRows * 32];
#pragma unroll
  for (int pass = 0; pass < 4; ++pass) {
    __syncthreads();
    epi_put<2, 4>(ep, 32, acc, pass, w * 64, 0, h, r);
    __syncthreads();
    epi_run<32, BLK>(a, ep, tid & 31, n0 + 32 * pass + (tid & 31), (int64_t)mt * kGemmRows, kGemmRows, tid >> 5, 8);
  }
}

// Split-bf16 / split-fp16 contraction staged through LDS.  The waves of a
// workgroup form a WR x WC grid of 128 x 64 wave tiles (4 x 2 MFMA 32x32x16
// tiles each), so the workgroup tile is (128 WR) rows x (64 WC) columns; the
// shape is a launch-time choice (the autotuner's).  One k step is a pair of
// k-groups (16 k); its A fragments ([2][rows][8] per plane) and B fragments
// ([2][columns][8]) are copied global -> LDS by the LDS DMA
// (global_load_lds_dwordx4: each wave instruction moves 64 consecutive
// 16-byte fragments, landing at consecutive LDS addresses), so every A and B
// element crosses from L2 once per workgroup instead of once per wave.  Two
// LDS stages: step s issues the DMA of step s+1 into the other stage, then
// reads its own fragments into registers and runs its MFMAs (each group of
// products waits only for the fragments it uses) while that DMA is in
// flight; one barrier per step separates the DMA of a stage from the reads
// of the step before.
template <int NP, int WR, int WC>
struct GemmStage {
  static constexpr int kRows = 128 * WR, kCols = 64 * WC, kThreads = 64 * WR * WC;
  static constexpr int kA = 2 * kRows * 16;   // bytes of A per plane
  static constexpr int kB = 2 * kCols * 16;   // bytes of B per plane
  static constexpr int kPlane = kA + kB;
  static constexpr int kBytes = NP * kPlane;
  static constexpr int kEpi = kRows * 32 * WC * 4;  // epilogue block: rows x (32 WC) floats
  // LDS stages: the loop is written for 2 or 3 (two copies in flight behind
  // the one being read); 3 measured slower on MI355X at C3 (less occupancy,
  // no gain in overlap), so 2.
  static constexpr int kStages = 2;
  static constexpr int kSmem = kStages * kBytes > kEpi ? kStages * kBytes : kEpi;
  // fewest copies any wave issues per stage (the round-robin leaves the last
  // waves one short when a chunk count is not a multiple of the wave count)
  static constexpr int kMinCopies = NP * ((2 * kRows / 64) / (WR * WC) + (2 * kCols / 64) / (WR * WC));
};

// s_waitcnt immediate for vmcnt(n) with lgkmcnt and expcnt left free (gfx9
// encoding: vmcnt[3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
template <int N>
struct VmCnt {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  static constexpr int kImm = (N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8);
};

// The 1 KiB copies of one stage (64 rows or columns of one k-group of one
// plane each) go round-robin over the waves.
template <int NP, int WR, int WC>
__device__ __forceinline__ void gemm_stage_copy(const OlsGemmArgs& a, uint8_t* stage, int kg0, int64_t mrow0,
                                                int64_t ncol0, int w) {
  using S = GemmStage<NP, WR, WC>;
  typedef __attribute__((address_space(3))) void* lds_ptr;
  constexpr int W = WR * WC, nA = 2 * S::kRows / 64, nB = 2 * S::kCols / 64;
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < NP; ++q) {
#pragma unroll
    for (int c0 = 0; c0 < nA; c0 += W) {
      const int c = c0 + w;
      if (c < nA) {
        const int j = c / (S::kRows / 64), rc = (c % (S::kRows / 64)) * 64;
        const uint16_t* src = a.abf[q] + ((int64_t)(kg0 + j) * a.Mpad + mrow0 + rc + l) * 8;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (lds_ptr)(stage + q * S::kPlane + (j * S::kRows + rc) * 16), 16, 0, 0);
      }
    }
#pragma unroll
    for (int c0 = 0; c0 < nB; c0 += W) {
      const int c = c0 + w;
      if (c < nB) {
        const int j = c / (S::kCols / 64), nc = (c % (S::kCols / 64)) * 64;
        const uint16_t* src = a.bbf[q] + ((int64_t)(kg0 + j) * a.Npad + ncol0 + nc + l) * 8;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (lds_ptr)(stage + q * S::kPlane + S::kA + (j * S::kCols + nc) * 16), 16,
                                         0, 0);
      }
    }
  }
}

template <int NP, bool F16, int WR, int WC, bool BLK = false>
__global__ __launch_bounds__(64 * WR * WC) void ols_gemm_split_kernel(OlsGemmArgs a) {
  using S = GemmStage<NP, WR, WC>;
  __shared__ __attribute__((aligned(16))) uint8_t smem[S::kSmem];
  const int tid = threadIdx.x, l = tid & 63, h = l >> 5, r = l & 31;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = w / WC, wc = w % WC;
  int mt, nt;
  if (!gemm_tile(a, mt, nt)) return;
  const int64_t mrow0 = (int64_t)mt * S::kRows, ncol0 = (int64_t)nt * S::kCols;
  f32x16 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  const int nsteps = a.Kp / 16;  // k steps of two k-groups
  constexpr int D = S::kStages;
#pragma unroll
  for (int s = 0; s < D - 1; ++s)
    if (s < nsteps) gemm_stage_copy<NP, WR, WC>(a, smem + s * S::kBytes, 2 * s, mrow0, ncol0, w);
  for (int step = 0; step < nsteps; ++step) {
    // stage `step` has landed once at most the copies of the stages issued
    // after it are outstanding (counted with the fewest copies per wave, so
    // a wave with more copies waits a little longer than it must)
    if (D == 3 && step + 1 < nsteps)
      __builtin_amdgcn_s_waitcnt(VmCnt<S::kMinCopies>::kImm);
    else
      __builtin_amdgcn_s_waitcnt(VmCnt<0>::kImm);
    // Every wave's copies of this stage have landed (each wave waited for its
    // own above), and every wave has finished reading the stage the next copy
    // reuses.  A bare s_barrier: __syncthreads' release fence would also wait
    // for the copies still in flight.  The empty asm keeps the compiler from
    // moving LDS reads across it.
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    // the copy of a later stage goes first, into the buffer every wave
    // finished reading before the barrier, so that the fragment reads below
    // are the last LDS operations before the products and can be waited for
    // in part
    if (step + D - 1 < nsteps)
      gemm_stage_copy<NP, WR, WC>(a, smem + ((step + D - 1) % D) * S::kBytes, 2 * (step + D - 1), mrow0, ncol0, w);
    const uint8_t* cur = smem + (step % D) * S::kBytes;
    typedef typename std::conditional<F16, f16x8, bf16x8>::type op8;
    op8 A[NP][4], B[NP][2];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
        A[q][rt] = *reinterpret_cast<const op8*>(cur + q * S::kPlane +
                                                  (h * S::kRows + wr * 128 + 32 * rt + r) * 16);
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
        B[q][ct] = *reinterpret_cast<const op8*>(cur + q * S::kPlane + S::kA +
                                                  (h * S::kCols + wc * 64 + 32 * ct + r) * 16);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        f32x16 c = acc[rt][ct];
        if constexpr (F16) {
          c = __builtin_amdgcn_mfma_f32_32x32x16_f16(A[1][rt], B[0][ct], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_32x32x16_f16(A[0][rt], B[1][ct], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_32x32x16_f16(A[0][rt], B[0][ct], c, 0, 0, 0);
        } else {
          if constexpr (NP == 3) {
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2][rt], B[0][ct], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1][rt], B[1][ct], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][rt], B[2][ct], c, 0, 0, 0);
          }
          c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1][rt], B[0][ct], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][rt], B[1][ct], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][rt], B[0][ct], c, 0, 0, 0);
        }
        acc[rt][ct] = c;
      }
  }
  // epilogue: two passes (column tile ct of every wave) through an LDS block
  // [rows][32 WC] that reuses the staging buffer
  float* ep = reinterpret_cast<float*>(smem);
  constexpr int ldc = 32 * WC;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    __syncthreads();
    epi_put<4, 2>(ep, ldc, acc, pass, wr * 128, wc * 32, h, r);
    __syncthreads();
    const int cl = tid % ldc;
    epi_run<ldc, BLK>(a, ep, cl, ncol0 + (cl >> 5) * 64 + 32 * pass + (cl & 31), mrow0, S::kRows, tid / ldc,
                 S::kThreads / ldc);
  }
}

}  // namespace

hipError_t launch_ols_prep(const OlsPrepArgs& a, hipStream_t s) {
  const int64_t blocks = (a.Npad + 255) / 256;
  hipLaunchKernelGGL(ols_prep_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ols_fixed(const OlsFixedArgs& a, hipStream_t s) {
  const int64_t blocks = (a.V + 255) / 256;
  hipLaunchKernelGGL(ols_fixed_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ols_factors(const OlsFactorArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(ols_factor_kernel, dim3(a.P), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ols_build_a(const OlsBuildAArgs& a, hipStream_t s) {
  const int64_t n = (int64_t)(a.Kp / 8) * a.Mpad;
  hipLaunchKernelGGL(ols_build_a_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int WR, int WC, bool BLK = false>
static void launch_split(const OlsGemmArgs& a, int precision, unsigned blocks, hipStream_t s) {
  const dim3 grid(blocks), block(64 * WR * WC);
  if (precision == kPrecBF16x3)
    hipLaunchKernelGGL((ols_gemm_split_kernel<2, false, WR, WC, BLK>), grid, block, 0, s, a);
  else if (precision == kPrecFP16x3)
    hipLaunchKernelGGL((ols_gemm_split_kernel<2, true, WR, WC, BLK>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((ols_gemm_split_kernel<3, false, WR, WC, BLK>), grid, block, 0, s, a);
}

hipError_t launch_ols_gemm(const OlsGemmArgs& a, int precision, int tile, hipStream_t s) {
  const int64_t ntp = (a.n_ntiles + 7) / 8 * 8;
  const unsigned blocks = (unsigned)(ntp * a.n_mtiles);
  if (a.n_blk > 0) {
    // block F: the 256 x 128 tile only (the caller sizes rows for it)
    if (precision == kPrecFP32)
      hipLaunchKernelGGL(ols_gemm_fp32_kernel<true>, dim3(blocks), dim3(256), 0, s, a);
    else
      launch_split<2, 2, true>(a, precision, blocks, s);
    return hipGetLastError();
  }
  if (precision == kPrecFP32)
    hipLaunchKernelGGL(ols_gemm_fp32_kernel<false>, dim3(blocks), dim3(256), 0, s, a);
  else if (tile == 1)
    launch_split<4, 1>(a, precision, blocks, s);
  else if (tile == 2)
    launch_split<4, 2>(a, precision, blocks, s);
  else if (tile == 3)
    launch_split<2, 4>(a, precision, blocks, s);
  else
    launch_split<2, 2>(a, precision, blocks, s);
  return hipGetLastError();
}

}  // namespace tfcehip

// ---------------------------------------------------------------------------
// FP64 single-design OLS, one thread per element (the reference's numpy path
// for one design: cynumstats.pyx:28-112).  a = invXX_a X'y, residuals, SSE,
// se_j = fl32(sqrt(SSE/(n-k) invXX_se[j][j])), t = a / se.
namespace tfcehip {
namespace {

__global__ __launch_bounds__(256) void ols_direct_kernel(OlsDirectArgs a) {
  __shared__ double sX[kOlsDirectMaxN * kOlsMaxK];
  for (int i = threadIdx.x; i < a.n * a.k; i += blockDim.x) sX[i] = a.X[i];
  __syncthreads();
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.V) return;
  const int n = a.n, k = a.k;
  double xty[kOlsMaxK], coef[kOlsMaxK];
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j) xty[j] = 0.0;
  double sy = 0.0, syy = 0.0;
  for (int i = 0; i < n; ++i) {
    const double y = (double)a.Y[(int64_t)i * a.ldy + v];
    sy += y;
    syy += y * y;
#pragma unroll
    for (int j = 0; j < kOlsMaxK; ++j)
      if (j < k) xty[j] = fma(sX[i * k + j], y, xty[j]);
  }
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j) {
    double s = 0.0;
    if (j < k)
      for (int q = 0; q < k; ++q) s = fma(a.Ainv[j * k + q], xty[q], s);
    coef[j] = s;
  }
  double sse = 0.0;
  for (int i = 0; i < n; ++i) {
    const double y = (double)a.Y[(int64_t)i * a.ldy + v];
    double fit = 0.0;
#pragma unroll
    for (int j = 0; j < kOlsMaxK; ++j)
      if (j < k) fit = fma(sX[i * k + j], coef[j], fit);
    const double r = y - fit;
    sse = fma(r, r, sse);
    if (a.resid) a.resid[(int64_t)i * a.V + v] = r;
  }
  const double s2 = sse / (double)(n - k);
#pragma unroll
  for (int j = 0; j < kOlsMaxK; ++j) {
    if (j < k) {
      const float se = (float)sqrt(s2 * a.Sinv[j * k + j]);
      if (a.coef) a.coef[(int64_t)j * a.V + v] = coef[j];
      if (a.se) a.se[(int64_t)j * a.V + v] = se;
      if (a.t) a.t[(int64_t)j * a.V + v] = coef[j] / (double)se;
    }
  }
  if (a.sse) a.sse[v] = sse;
  if (a.sumy) a.sumy[v] = sy;
  if (a.sumyy) a.sumyy[v] = syy;
}

// Sobel-type z of the indirect effect from the two path t-values
// (pyfunc.py:150-161, calc_indirect :2678-2709): 0 aroian, 1 sobel, 2 goodman.
__global__ void sobel_kernel(const float* ta, const float* tb, int64_t n, int alg, float ta_const, float* z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = ta ? (double)ta[i] : (double)ta_const;
  z[i] = sobel_z(a, (double)tb[i], alg);
}

}  // namespace

hipError_t launch_ols_direct(const OlsDirectArgs& a, hipStream_t s) {
  const int64_t blocks = (a.V + 255) / 256;
  hipLaunchKernelGGL(ols_direct_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sobel(const float* ta, const float* tb, int64_t n, int alg, float ta_const, float* z,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(sobel_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ta, tb, n, alg, ta_const, z);
  return hipGetLastError();
}

}  // namespace tfcehip

namespace tfcehip {
namespace {
// se[j][v] = fl32(sqrt(sigma2[v] * invXX[j][j]))  (cynumstats.pyx:47-52)
__global__ void se_of_slope_kernel(const double* sigma2, const double* diag, int k, int64_t V, float* se) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  for (int j = 0; j < k; ++j) se[(int64_t)j * V + v] = (float)sqrt(sigma2[v] * diag[j]);
}
}  // namespace
// Total sum of squares of calcF (cynumstats.pyx:34): per element, about the
// grand mean of all of Y, in FP64.  Three passes: column sums, the grand mean
// (one workgroup, fixed reduction order), then sum_i (y_iv - mean)^2.
__global__ __launch_bounds__(256) void colsum_kernel(const float* Y, int32_t n, int64_t V, int64_t ldy, double* out) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= V) return;
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) s += (double)Y[i * ldy + v];
  out[v] = s;
}

__global__ __launch_bounds__(1024) void grand_mean_kernel(const double* sums, int64_t V, int32_t n, double* mean) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int64_t v = threadIdx.x; v < V; v += 1024) s += sums[v];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) mean[0] = red[0] / ((double)n * (double)V);
}

__global__ __launch_bounds__(256) void tss_kernel(const float* Y, int32_t n, int64_t V, int64_t ldy, const double* mean,
                                                 double* tss) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= V) return;
  const double m = mean[0];
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    const double d = (double)Y[i * ldy + v] - m;
    s += d * d;
  }
  tss[v] = s;
}

hipError_t launch_total_ss(const float* Y, int32_t n, int64_t V, int64_t ldy, double* tss, double* mean,
                           hipStream_t s) {
  const unsigned blocks = (unsigned)((V + 255) / 256);
  hipLaunchKernelGGL(colsum_kernel, dim3(blocks), dim3(256), 0, s, Y, n, V, ldy, tss);
  hipLaunchKernelGGL(grand_mean_kernel, dim3(1), dim3(1024), 0, s, tss, V, n, mean);
  hipLaunchKernelGGL(tss_kernel, dim3(blocks), dim3(256), 0, s, Y, n, V, ldy, mean, tss);
  return hipGetLastError();
}

hipError_t launch_se_of_slope(const double* sigma2, const double* diag, int k, int64_t V, float* se,
                              hipStream_t s) {
  if (V <= 0) return hipSuccess;
  hipLaunchKernelGGL(se_of_slope_kernel, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, sigma2, diag, k, V, se);
  return hipGetLastError();
}
}  // namespace tfcehip
