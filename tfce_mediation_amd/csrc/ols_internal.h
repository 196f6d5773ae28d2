// Internal declarations of the permuted-OLS kernels (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfcehip {

// Contraction precision of the permuted design with Y.
constexpr int kPrecFP32 = 0;    // FP32 MFMA (v_mfma_f32_32x32x2_f32)
constexpr int kPrecBF16x3 = 1;  // hi/lo bf16 split, hi*hi + hi*lo + lo*hi, FP32 accumulate
constexpr int kPrecBF16x6 = 2;  // three-piece bf16 split, every product of order <= 2
constexpr int kPrecFP16x3 = 3;  // hi/lo fp16 split of power-of-two scaled operands, three products

constexpr int kOlsMaxK = 8;         // design columns
constexpr int kOlsMaxContrasts = 8;
constexpr int kGemmRows = 256;      // rows (permutation x coefficient) per workgroup
constexpr int kGemmCols = 128;      // elements per workgroup
constexpr int kColAlign = 1024;     // element padding: 8 XCDs x 128 columns

inline int ols_pieces(int precision) {
  return precision == kPrecBF16x6 ? 3 : ((precision == kPrecBF16x3 || precision == kPrecFP16x3) ? 2 : 1);
}

struct OlsPrepArgs {
  const float* Y;   // [n][ldy]
  int32_t n;
  int32_t Kp;       // n rounded up to 16
  int64_t V, ldy, Npad;
  int32_t center;
  int32_t pieces;   // bf16 pieces to write (0 = none)
  int32_t f16;      // pieces are fp16 of y * yscale[v] (power of two per element)
  float* yinv;      // [V] 1 / yscale (fp16 mode)
  double* mu;       // [V]
  double* yy;       // [V]
  float* plane32;   // [Kp/8][Npad][8] centred float32 (always written)
  uint16_t* bf[3];  // [Kp/8][Npad][8] bf16 pieces
};

struct OlsFixedArgs {
  const float* plane32;  // data
  int32_t Kp, n;
  int64_t V, Npad;
  const double* yy;
  const double* F;       // [n][f] fixed columns (internal order)
  int32_t f;
  const double* LFF;     // [f][f] inverse Cholesky factor of F'F (lower)
  float* gF;             // [f][V]  F'Y
  float* hF;             // [f][V]  LFF * gF
  double* r0;            // [V]     yy - |hF|^2
};

struct OlsFactorArgs {
  const double* F;       // [n][f]
  const double* Z;       // [n][m]
  const double* FF;      // [f][f]
  const double* ZZ;      // [m][m]
  const int32_t* perm;   // [P][n] or null (identity)
  int32_t n, f, m, P;
  int32_t n_con;
  int32_t con[kOlsMaxContrasts];  // internal column index of each contrast
  float* Linv;           // [P][8][8]
  double* dc;            // [P][kOlsMaxContrasts]
  // Column blocks whose F against the fit without them is wanted: block b's
  // internal columns are the set bits of blk_mask[b]; Rb[p][b] (8 x 8, rows
  // < the block's width) maps h = L^-1 X_p'y to a vector whose squared norm
  // is the block's extra sum of squares.
  int32_t n_blk;
  uint32_t blk_mask[kOlsMaxK];
  float* Rb;             // [P][n_blk][8][8] or null
};

struct OlsBuildAArgs {
  const float* Z32;      // [n][m] permuted columns, float
  const int32_t* perm;   // [P][n] or null
  int32_t n, m, P, Kp, rpp;
  int32_t tile_rows;     // a row tile holds tile_rows / rpp whole permutations
  int64_t Mpad;
  int32_t pieces;
  int32_t f16;           // pieces are fp16 of z * zscale[j]
  float zscale[8];       // power of two per permuted column (fp16 mode)
  float* a32;            // [Kp/8][Mpad][8] (FP32 mode)
  uint16_t* abf[3];      // [Kp/8][Mpad][8] bf16 pieces
};

struct OlsGemmArgs {
  const float* a32;
  const uint16_t* abf[3];
  const float* b32;
  const uint16_t* bbf[3];
  int32_t Kp;
  int64_t Mpad, Npad, V;
  int32_t rpp;           // rows per permutation (= permuted columns, >= 1)
  int32_t P;             // valid permutations
  int32_t n, k, f, m;
  double inv_dof;        // 1 / (n - k)
  int32_t n_con;
  int32_t con[kOlsMaxContrasts];
  int32_t con_is_intercept[kOlsMaxContrasts];
  int32_t shared_factor; // 1: factors at index 0 serve every permutation
  const float* Linv;     // [P or 1][8][8]
  const double* dc;      // [P or 1][kOlsMaxContrasts]
  const float* gF;       // [f][V]
  const float* hF;       // [f][V]
  const double* r0;      // [V]
  const double* mu;      // [V]
  int32_t centered;
  const double* colsumZ; // [m]  (uncentred mode correction)
  float* t;              // [n_con][P][ldt]
  int64_t ldt;
  int32_t n_mtiles, n_ntiles;
  const float* yinv;     // [V] per-element inverse data scale, or null
  float zinv[8];         // per permuted column inverse design scale (fp16 mode)
  const double* tss;     // [V] total sum of squares: write sqrt(max(F, 0)) to t[p][col]; null: t-values
  // Paired mediation epilogue (pair != 0): a second model on the same data with
  // the same single permuted column (its fixed columns differ); the t of
  // contrast 0 of each model combine into the Sobel-type z of alg (0 aroian,
  // 1 sobel, 2 goodman), written to t[p][col].
  int32_t pair, alg;
  int32_t f2, k2, con2, con2_is_intercept, shared2;
  const float* Linv2;    // [P or 1][8][8]
  const double* dc2;     // [P or 1][kOlsMaxContrasts]
  const float* gF2;      // [f2][V]
  const float* hF2;      // [f2][V]
  const double* r02;     // [V]
  // Partial F of column blocks (n_blk > 0): F_b = |Rb[p][b] h|^2 / w_b over
  // SSE / (n - k), written to t[(b * P + p) * ldt + col].
  int32_t n_blk, blk_shared;
  int32_t blk_w[kOlsMaxK];
  const float* Rb;       // [P or 1][n_blk][8][8]
};

constexpr int kOlsDirectMaxN = 2048;  // subjects held in LDS by the FP64 kernel

struct OlsDirectArgs {
  const float* Y;        // [n][ldy]
  int32_t n, k;
  int64_t V, ldy;
  const double* X;       // [n][k] (device)
  const double* Ainv;    // [k][k] inverse used for the coefficients
  const double* Sinv;    // [k][k] inverse whose diagonal scales se
  double* coef;          // [k][V] or null
  float* se;             // [k][V] or null
  double* t;             // [k][V] or null
  double* resid;         // [n][V] or null
  double* sse;           // [V] or null
  double* sumy;          // [V] or null
  double* sumyy;         // [V] or null
};

hipError_t launch_ols_direct(const OlsDirectArgs& a, hipStream_t s);

// Element-wise regressor (image_regression): t of the r shared columns and of
// the element's own regressor, written to T[j][ldt] (j = r: the regressor).
struct ImgRegArgs {
  const float* Y;     // [n][ldy]
  const float* Z;     // [n][ldz] the element-wise regressor
  int64_t ldy, ldz;
  int32_t n, r;       // subjects, shared columns (intercept first), r + 1 <= kOlsMaxK
  int64_t V;
  const double* R;    // [n][r] (device)
  const double* Ainv; // [r][r] inverse of R'R (device)
  double std_min;     // regressors with a smaller standard deviation give t = 0
  float* T;           // [r + 1][ldt]
  int64_t ldt;
};

// Variance inflation image (image_reg_VIF): 1 / (1 - R2) of y on X.
struct ImgVifArgs {
  const void* Y;      // [n][ldy] float or double
  int64_t ldy;
  int32_t n, k;
  int64_t V;
  const double* X;    // [n][k] (device), intercept first
  const double* Ainv; // [k][k] inverse of X'X (device)
  double* vif;        // [V]
};

hipError_t launch_image_regression(const ImgRegArgs& a, hipStream_t s);
hipError_t launch_image_vif(const ImgVifArgs& a, bool f64, hipStream_t s);
hipError_t launch_se_of_slope(const double* sigma2, const double* diag, int k, int64_t V, float* se,
                              hipStream_t s);
hipError_t launch_total_ss(const float* Y, int32_t n, int64_t V, int64_t ldy, double* tss, double* mean,
                           hipStream_t s);
hipError_t launch_sobel(const float* ta, const float* tb, int64_t n, int alg, float ta_const, float* z,
                        hipStream_t s);
hipError_t launch_ols_prep(const OlsPrepArgs& a, hipStream_t s);
hipError_t launch_ols_fixed(const OlsFixedArgs& a, hipStream_t s);
hipError_t launch_ols_factors(const OlsFactorArgs& a, hipStream_t s);
hipError_t launch_ols_build_a(const OlsBuildAArgs& a, hipStream_t s);
// Workgroup tile of the split contraction: 0 = 256 x 128 (2 x 2 waves),
// 1 = 512 x 64 (4 x 1 waves), 2 = 512 x 128 (4 x 2 waves), 3 = 256 x 256
// (2 x 4 waves).  FP32 uses 256 x 128.
constexpr int kGemmTiles = 4;
inline int gemm_tile_rows(int tile) { return (tile == 0 || tile == 3) ? 256 : 512; }
inline int gemm_tile_cols(int tile) { return tile == 1 ? 64 : (tile == 3 ? 256 : 128); }
hipError_t launch_ols_gemm(const OlsGemmArgs& a, int precision, int tile, hipStream_t s);

}  // namespace tfcehip
