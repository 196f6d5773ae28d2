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
// contraction [P*rpp x n] . [n x V] on MFMA, rpp = 4 or 8 rows per permutation.
// With L the Cholesky factor of the permuted Gram matrix (X_p'X_p = L L'),
// h = L^-1 X_p'y gives |y - X a|^2 = y'y - |h|^2 and a = L^-T h, so the
// epilogue needs only h_Z = Linv_ZF g_F + Linv_ZZ g_Z per (permutation, element):
//   SSE = r0 - |h_Z|^2 with r0 = y'y - |Linv_FF g_F|^2 precomputed in FP64.
// Y is centred per element first (exact for every coefficient once the model
// holds an intercept; the intercept's own coefficient gets the mean back), so
// the bf16 split loses precision relative to the spread of y, not its offset.
#include "ols_internal.h"

namespace tfcehip {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

__global__ __launch_bounds__(256) void ols_prep_kernel(OlsPrepArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.Npad) return;
  const bool valid = v < a.V;
  double mu = 0.0;
  if (valid && a.center) {
    for (int i = 0; i < a.n; ++i) mu += (double)a.Y[(int64_t)i * a.ldy + v];
    mu /= (double)a.n;
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
}

__global__ __launch_bounds__(256) void ols_build_a_kernel(OlsBuildAArgs a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nkg = a.Kp / 8;
  if (idx >= (int64_t)nkg * a.Mpad) return;
  const int kg = (int)(idx / a.Mpad);
  const int64_t row = idx - (int64_t)kg * a.Mpad;
  const int p = (int)(row / a.rpp), j = (int)(row % a.rpp);
  float vals[8];
  uint16_t pc[3][8];
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const int i = kg * 8 + jj;
    float z = 0.0f;
    if (p < a.P && j < a.m && i < a.n) {
      const int src = a.perm ? a.perm[(int64_t)p * a.n + i] : i;
      z = (src >= 0 && src < a.n) ? a.Z32[src * a.m + j] : NAN;
    }
    vals[jj] = z;
    uint16_t s[3];
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
__device__ __forceinline__ void ols_epilogue(const OlsGemmArgs& a, int p, int64_t col, const float* gz,
                                             const float* gF, const float* hF, double r0, double mu) {
  const int fi = a.shared_factor ? 0 : p;
  const float* L = a.Linv + (int64_t)fi * 64;
  float hz[kOlsMaxK];
  double ssz = 0.0;
#pragma unroll
  for (int i = 0; i < kOlsMaxK; ++i) {
    hz[i] = 0.0f;
    if (i < a.m) {
      const int row = a.f + i;
      float s = 0.0f;
#pragma unroll
      for (int j = 0; j < kOlsMaxK; ++j)
        if (j < a.f) s = fmaf(L[row * 8 + j], gF[j], s);
#pragma unroll
      for (int j = 0; j < kOlsMaxK; ++j)
        if (j <= i) s = fmaf(L[row * 8 + a.f + j], gz[j], s);
      hz[i] = s;
      ssz += (double)s * (double)s;
    }
  }
  double sse = r0 - ssz;
  if (sse < 0.0) sse = 0.0;
  const double s2 = sse / (double)(a.n - a.k);
  for (int c = 0; c < a.n_con; ++c) {
    const int ci = a.con[c];
    float ac = 0.0f;
#pragma unroll
    for (int j = 0; j < kOlsMaxK; ++j) {
      if (j >= ci && j < a.k) {
        const float hj = (j < a.f) ? hF[j < kOlsMaxK ? j : 0] : hz[(j - a.f) & (kOlsMaxK - 1)];
        ac = fmaf(L[j * 8 + ci], hj, ac);
      }
    }
    double av = (double)ac;
    if (a.centered && a.con_is_intercept[c]) av += mu;
    const float se = (float)sqrt(s2 * a.dc[(int64_t)fi * kOlsMaxContrasts + c]);
    a.t[((int64_t)c * a.P + p) * a.ldt + col] = (float)(av / (double)se);
  }
}

template <int PREC>
__global__ __launch_bounds__(256) void ols_gemm_kernel(OlsGemmArgs a) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, h = l >> 5, r = l & 31;
  const int b = blockIdx.x, xcd = b & 7, jb = b >> 3;
  const int mt = jb % a.n_mtiles;
  const int nt = (jb / a.n_mtiles) * 8 + xcd;
  if (nt >= a.n_ntiles) return;
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
    if constexpr (PREC == kPrecFP32) {
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
    } else {
      constexpr int NP = (PREC == kPrecBF16x6) ? 3 : 2;
      bf16x8 A[NP][2], B[NP][4];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
          A[q][rt] = ld_bf16x8(a.abf[q] + ((int64_t)kg * a.Mpad + m0 + 32 * rt + r) * 8);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
          B[q][ct] = ld_bf16x8(a.bbf[q] + ((int64_t)kg * a.Npad + n0 + 32 * ct + r) * 8);
      }
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          f32x16 c = acc[rt][ct];
          if constexpr (PREC == kPrecBF16x6) {
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2][rt], B[0][ct], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1][rt], B[1][ct], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][rt], B[2][ct], c, 0, 0, 0);
          }
          c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1][rt], B[0][ct], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][rt], B[1][ct], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0][rt], B[0][ct], c, 0, 0, 0);
          acc[rt][ct] = c;
        }
    }
  }

  // ---- epilogue: lane holds rows 32rt + 8(i/4) + 4h + (i%4) of column r ----
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) {
    const int64_t col = n0 + 32 * ct + r;
    const bool cvalid = col < a.V;
    float gF[kOlsMaxK], hF[kOlsMaxK];
    double r0 = 0.0, mu = 0.0;
#pragma unroll
    for (int q = 0; q < kOlsMaxK; ++q) {
      gF[q] = (cvalid && q < a.f) ? a.gF[(int64_t)q * a.V + col] : 0.0f;
      hF[q] = (cvalid && q < a.f) ? a.hF[(int64_t)q * a.V + col] : 0.0f;
    }
    if (cvalid) {
      r0 = a.r0[col];
      mu = a.mu[col];
    }
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
#pragma unroll
      for (int a4 = 0; a4 < 4; ++a4) {
        float G[4], O[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) G[q] = acc[rt][ct][4 * a4 + q];
        if (a.rpp == 4) {
          const int p = (int)((m0 + 32 * rt + 8 * a4 + 4 * h) >> 2);
          float gz[kOlsMaxK];
#pragma unroll
          for (int q = 0; q < kOlsMaxK; ++q) gz[q] = q < 4 ? G[q & 3] : 0.0f;
          if (!a.centered) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (q < a.m) gz[q] += (float)(a.colsumZ[q] * mu);
          }
          if (cvalid && p < a.P) ols_epilogue(a, p, col, gz, gF, hF, r0, mu);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) O[q] = __shfl_xor(G[q], 32, 64);
          const int p = (int)((m0 + 32 * rt + 8 * a4) >> 3);
          float gz[kOlsMaxK];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            gz[q] = h ? O[q] : G[q];
            gz[4 + q] = h ? G[q] : O[q];
          }
          if (!a.centered) {
#pragma unroll
            for (int q = 0; q < kOlsMaxK; ++q)
              if (q < a.m) gz[q] += (float)(a.colsumZ[q] * mu);
          }
          if (cvalid && p < a.P && (a4 & 1) == h) ols_epilogue(a, p, col, gz, gF, hF, r0, mu);
        }
      }
    }
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

hipError_t launch_ols_gemm(const OlsGemmArgs& a, int precision, hipStream_t s) {
  const int64_t ntp = (a.n_ntiles + 7) / 8 * 8;
  const unsigned blocks = (unsigned)(ntp * a.n_mtiles);
  if (precision == kPrecFP32)
    hipLaunchKernelGGL(ols_gemm_kernel<kPrecFP32>, dim3(blocks), dim3(256), 0, s, a);
  else if (precision == kPrecBF16x3)
    hipLaunchKernelGGL(ols_gemm_kernel<kPrecBF16x3>, dim3(blocks), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(ols_gemm_kernel<kPrecBF16x6>, dim3(blocks), dim3(256), 0, s, a);
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
  const double b = (double)tb[i];
  const double ia = 1.0 / (a * a), ib = 1.0 / (b * b), iab = 1.0 / (a * a * b * b);
  double d;
  if (alg == 0) d = ib + ia + iab;
  else if (alg == 1) d = ib + ia;
  else d = ib + ia - iab;
  z[i] = (float)(1.0 / sqrt(d));
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
hipError_t launch_se_of_slope(const double* sigma2, const double* diag, int k, int64_t V, float* se,
                              hipStream_t s) {
  if (V <= 0) return hipSuccess;
  hipLaunchKernelGGL(se_of_slope_kernel, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, sigma2, diag, k, V, se);
  return hipGetLastError();
}
}  // namespace tfcehip
