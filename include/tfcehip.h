/*
 * libtfcehip — C ABI of the MI355X permutation hot path of TFCE_mediation.
 *
 * Conventions
 *   - every function returns int: 0 = OK, < 0 = error; -1..-999 are negated
 *     hipError_t values, <= -1000 are validation codes below;
 *     tfce_last_error() gives a thread-local message for the last failure;
 *   - d_* arguments are device pointers owned by the caller; host arrays are
 *     read during the call only;
 *   - each call enqueues on the given stream and returns without a host
 *     synchronisation, unless its comment says otherwise;
 *   - handles are immutable after creation, so one handle may serve calls on
 *     several streams at once (scratch memory is kept per (device, stream));
 *   - graph capture: a call made while its stream captures a graph allocates
 *     nothing and fails with TFCE_EINVAL unless the same call ran eagerly on
 *     that stream before (the stream's scratch is large enough); the graph
 *     then holds that scratch, which the library never frees while a later
 *     call on the stream grows it.  Replays share the capture stream's
 *     scratch with each other and with later calls on that stream: run them
 *     in order, not concurrently with such calls.
 */
#ifndef TFCEHIP_H
#define TFCEHIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFCE_OK 0
#define TFCE_EINVAL (-1000)      /* bad argument */
#define TFCE_EADJ (-1001)        /* adjacency index out of range / malformed CSR */
#define TFCE_EDEGENERATE (-1002) /* max(image) >= minT with a non-positive step */
#define TFCE_ELEVELS (-1003)     /* more than 1024 threshold levels */
#define TFCE_ENOTSUP (-1004)

/* ---------------------------------------------------------------- errors */
const char *tfce_last_error(void);
const char *tfce_version(void);

/* ------------------------------------------------------------- adjacency */
typedef struct tfce_adj tfce_adj_t;

/* Adjacency flags (tfce_adj_create_ex) */
#define TFCE_ADJ_IMPLICIT_ZERO 2   /* the full image also holds zeros outside these elements
                                      (masked vertices): they enter max() for the levels */
#define TFCE_ADJ_ZERO_IN_SCALE 4   /* ... and for the max/100 rescale */

/* Replaces CreateAdjSet(H, E, adjacency) (tfce_mediation/tfce.pyx:24-42): CSR
 * copy of the ragged adjacency plus the float H and E the operator stores
 * (tfce.pyx:27-28).  row_ptr has V+1 entries, col holds indices in [0, V).
 * Symmetry is detected here; an asymmetric list is honoured the way
 * fast_tfce.hpp:44-66 uses it (an edge counts when it is listed by the later-
 * activated end).  Blocking: uploads the CSR before returning, with a
 * fixed-width copy of the lists (longest list <= 32) and, for a symmetric
 * adjacency, which entries of each list are adjacent to each other (within
 * groups of 8 entries; host work O(nnz * 8 log degree)).  When every list
 * holds at most 6 neighbours and V < 2^18 - 1 (closed triangle meshes such
 * as fsaverage), the lists are put in ring order instead and packed with the
 * adjacency of ring-consecutive neighbours into one 16-byte row per vertex.
 * The sweep uses these to read a vertex's neighbours in one row and to join
 * each group of already-connected neighbours once. */
int tfce_adj_create(const int32_t *row_ptr, const int32_t *col, int32_t V, int64_t nnz,
                    float H, float E, tfce_adj_t **out);
int tfce_adj_create_ex(const int32_t *row_ptr, const int32_t *col, int32_t V, int64_t nnz,
                       float H, float E, uint32_t flags, tfce_adj_t **out);
int tfce_adj_destroy(tfce_adj_t *adj);

/* Locality order of an adjacency, host only: new_of_old[v] = new index of
 * element v (reverse Cuthill-McKee per component from a pseudo-peripheral
 * vertex).  No counterpart in the reference, whose TFCE walks the caller's
 * element order; relabelling a symmetric adjacency and its maps with this
 * order changes no partition and no TFCE value, and puts neighbours close in
 * memory, which cuts the random HBM traffic of the threshold sweep. */
int tfce_locality_order(const int32_t *row_ptr, const int32_t *col, int32_t V, int32_t *new_of_old);
int tfce_adj_info(const tfce_adj_t *adj, int32_t *V, int64_t *nnz, int32_t *symmetric);

/* Replaces create_adjac_voxel(data_index, data_mask, num_voxel, dirtype)
 * (pyfunc.py:48-76) on the GPU, as CSR: d_mask [X][Y][Z] uint8 in C order
 * (nonzero = in the mask); the elements are the in-mask voxels in C order
 * (np.where order); row v lists the labels of its in-mask neighbours in the
 * 3x3x3 window (dirtype 26) or its six face neighbours (dirtype 6), ascending,
 * with the reference's rules: label 0 is never a neighbour, row 0 is empty, and
 * with dirtype 6 a voxel whose window is clipped by the volume edge has none.
 * With d_row_ptr and d_col NULL the call only writes *V and *nnz; then, with
 * d_row_ptr [V+1] and d_col [nnz], it fills them.  Blocking. */
int tfce_adj_voxel_csr(const uint8_t *d_mask, int32_t X, int32_t Y, int32_t Z, int32_t dirtype,
                       int32_t *d_row_ptr, int32_t *d_col, int32_t *V, int64_t *nnz, hipStream_t stream);

/* Replaces create_adjac_vertex(vertices, faces) (pyfunc.py:37-46) on the GPU,
 * as CSR: d_faces [F][3] int32; row v lists, ascending and once each, every
 * vertex that shares a face with v (v itself when a face repeats it, as the
 * reference's set does).  Same two-call protocol as tfce_adj_voxel_csr;
 * TFCE_EADJ if a face names a vertex outside [0, n_vertices).  Blocking. */
int tfce_adj_mesh_csr(const int32_t *d_faces, int64_t F, int32_t n_vertices, int32_t *d_row_ptr,
                      int32_t *d_col, int64_t *nnz, hipStream_t stream);

/* Replaces adjacency.compute(v, f, thresholds) and compute_distance_parallel
 * (adjacency.pyx:44-79, :82-106: the exact geodesic distance of lib/geodesic
 * from every vertex i to every vertex j > i): every pair (i < j) of a
 * triangle mesh (float32 vertices [n_vertices][3], int32 faces
 * [n_faces][3]) whose distance along the surface is below max_distance, with
 * that distance, in (i, j) order, for the sources i in
 * [source_begin, source_end) (source_end <= 0: every vertex).  Host code,
 * n_threads host threads (0: all cores); the result is read with
 * tfce_geodesic_count / _copy and freed with tfce_geodesic_destroy.
 * TFCE_EADJ if a face names a vertex out of range. */
typedef struct tfce_geodesic tfce_geodesic_t;
int tfce_geodesic_pairs(const float *vertices, int32_t n_vertices, const int32_t *faces, int32_t n_faces,
                        double max_distance, int32_t source_begin, int32_t source_end, int32_t n_threads,
                        tfce_geodesic_t **out);
int tfce_geodesic_count(const tfce_geodesic_t *pairs, int64_t *n_pairs);
int tfce_geodesic_copy(const tfce_geodesic_t *pairs, int32_t *i, int32_t *j, double *distance);
int tfce_geodesic_destroy(tfce_geodesic_t *pairs);

/* Replaces CreateAdjSet.run(image, enhn) (tfce.pyx:44-45 -> fast_tfce.hpp:11-95):
 * enhn += TFCE(image) over V float32 elements (accumulate != 0), or
 * enhn = TFCE(image) (accumulate == 0).  dh == 0 gives the reference's step
 * fl(max/100); dh > 0 an explicit step.  minT is 0 as in tfce.pyx:45.  A
 * symmetric surface of 32,768 elements or more is swept over the whole GPU
 * (a cooperative launch, levels in lock step), a smaller one by one
 * workgroup; tfce_set_single_image_mode changes the choice, not the values. */
int tfce_run(tfce_adj_t *adj, const float *d_image, float *d_enhn, float dh, int accumulate,
             hipStream_t stream);

/* tfce_run with every knob: min_t, optional per-call status (levels visited, or
 * a negative TFCE_E* code) and component labels after level labels_level
 * (d_labels[v] = a representative vertex of v's component, -1 inactive; the
 * caller fills it with -1 first). */
int tfce_run_ex(tfce_adj_t *adj, const float *d_image, float *d_enhn, float min_t, float dh,
                int accumulate, int32_t labels_level, int32_t *d_labels, int32_t *d_status,
                hipStream_t stream);

/* Batched permutation maxima for one adjacency, replacing the per-permutation
 * pair of write_perm_maxTFCE_voxel / _vertex calls (pyfunc.py:107-126):
 *   d_max[2p + 0] = max_v fl(fl(TFCE(+map_p)_v * fl(max(+map_p)/100)) * vd_v)
 *   d_max[2p + 1] = same for -map_p
 * d_maps is P x V row-major.  sign_mask bit0 = +, bit1 = - (an unrequested
 * sign writes 0).  d_vdensity: per-element weights or NULL (= 1). */
int tfce_batch_max(tfce_adj_t *adj, const float *d_maps, int P, int sign_mask,
                   const float *d_vdensity, float *d_max, hipStream_t stream);

/* tfce_batch_max that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The scratch
 * is one device array of TFCE_NOALLOC_SCRATCH_BYTES per GPU, held from module
 * load until the process ends; its address is looked up on the first call on
 * each device.  A call fails with TFCE_EINVAL, before any launch, when the
 * per-item records of 2*P items plus one union-find slot of the adjacency do
 * not fit (about 8.2 KB per item, and 4*(8 + w)*V bytes per slot with w = 0, 4,
 * 8 or 16 words of neighbour rows); within the limit fewer workgroups run
 * when fewer slots fit.  Every call on one GPU shares the array: calls that
 * overlap on one GPU race.  Same outputs as tfce_batch_max, bit for bit. */
#define TFCE_NOALLOC_SCRATCH_BYTES (256u << 20)
int tfce_batch_max_noalloc(tfce_adj_t *adj, const float *d_maps, int P, int sign_mask,
                           const float *d_vdensity, float *d_max, hipStream_t stream);

/* ---------------------------------------------------------- surface sets */
/* Several adjacencies laid out along one map row (lh|rh hemispheres, or the
 * merged image of tm_func.py:54-123 / create_position_array :556-562). */
typedef struct tfce_surfset tfce_surfset_t;
int tfce_surfset_create(const tfce_adj_t *const *adjs, int32_t n, const int64_t *offsets,
                        tfce_surfset_t **out);
int tfce_surfset_destroy(tfce_surfset_t *set);
/* Launch order of a set (largest surface first; surfaces the LDS-resident
 * sweep takes after all others): user_index[k] = caller's index of the k-th
 * surface; outputs of tfce_batch follow this order. */
int tfce_surfset_order(const tfce_surfset_t *set, int32_t *user_index);

/* One launch over every (surface, map, sign) item, largest surface first.
 * Outputs are indexed item = (surface * P + p) * n_signs + sign_index:
 *   d_scaled: rescaled maximum as tfce_batch_max; d_raw: max of the TFCE sum;
 *   d_own: max of sign*map over the surface; d_status: levels or negative code.
 * d_T0 (optional, [n_items]) replaces each item's first threshold (the shared-
 * threshold mode of tm_func.py:54-123, where deltaT comes from the merged
 * image).  d_vdensity (optional) holds per-element weights in map-column
 * indexing.  Any output pointer except d_scaled may be NULL. */
int tfce_batch(const tfce_surfset_t *set, const float *d_maps, int64_t map_stride, int P,
               int sign_mask, const float *d_T0, const float *d_vdensity, float *d_scaled,
               float *d_raw, float *d_own, int32_t *d_status, hipStream_t stream);

/* ----------------------------------------------------------- permuted OLS */
/* Precision of the permuted-design contraction (every mode accumulates in
 * FP32; Y'Y, the residual sum of squares and the k x k factors are FP64). */
#define TFCE_PREC_FP32 0    /* FP32 matrix cores */
#define TFCE_PREC_BF16X3 1  /* hi/lo bf16 split, three products */
#define TFCE_PREC_BF16X6 2  /* three-piece bf16 split, six products */
#define TFCE_PREC_FP16X3 3  /* hi/lo fp16 split of operands scaled by powers of two
                               (per element of Y, per permuted design column), three products */

typedef struct tfce_ols_data tfce_ols_data_t;
typedef struct tfce_ols_model tfce_ols_model_t;

/* Device-resident copy of Y (n subjects x V elements, row stride ldy, the
 * merge_y layout of STEP_1_vertex_tfce_multiple_regression.py:263), centred
 * per element when center != 0 (exact for models with an intercept), with
 * FP64 y'y and the split planes the chosen precision needs.  Enqueued on
 * stream; d_Y may be freed once the stream has passed this call. */
int tfce_ols_data_create(const float *d_Y, int32_t n, int64_t V, int64_t ldy, int32_t precision,
                         int32_t center, tfce_ols_data_t **out, hipStream_t stream);
int tfce_ols_data_destroy(tfce_ols_data_t *data);

/* Design X (host, n x k row-major float64, k <= 8) bound to data.  Bit j of
 * perm_mask marks column j as permuted (the reference permutes every row of
 * X, vertex_tfce_multiple_regression_randomise.py:105, i.e. all columns; -v
 * permutes a subset, :93-97).  contrasts lists the columns whose t is wanted
 * (the reference's tnum = 1..k-1, :114-117).  Blocking. */
int tfce_ols_model_create(const tfce_ols_data_t *data, const double *X, int32_t k, uint32_t perm_mask,
                          const int32_t *contrasts, int32_t n_contrasts, tfce_ols_model_t **out,
                          hipStream_t stream);
int tfce_ols_model_destroy(tfce_ols_model_t *model);
int tfce_ols_info(const tfce_ols_model_t *model, int32_t *k, int32_t *n_fixed, int32_t *shared_gram);

/* Replaces tval_int (cynumstats.pyx:59-64) on X[perm_p] for P permutations at
 * once: d_t[(c * P + p) * ldt + v] = t of contrast c, permutation p, element v
 * (float32, se rounded to float32 as cynumstats.pyx:49 does).  d_perm is P x n
 * int32 (row i of the permuted design is X[perm[i]]); NULL = identity, P = 1. */
int tfce_ols_perm_t(const tfce_ols_model_t *model, const int32_t *d_perm, int32_t P, float *d_t,
                    int64_t ldt, hipStream_t stream);

/* Replaces calc_sobelz (pyfunc.py:130-162) of mediation types M and I for a
 * batch of permutations of x (vertex_tfce_mediation_randomise.py:82-91):
 * path_a (y_img ~ [1, x]) and path_b (y_img ~ [1, y, x] for type M, [1, x, y]
 * for type I) are models on one data handle that permute the same single
 * column (x), each with one contrast (x for path A; y for type M, x for type
 * I).  X_p'Y is contracted once for both; the epilogue computes both t-values
 * as tfce_ols_perm_t would and writes d_z[p * ldz + v] = the Sobel-type z of
 * alg (0 aroian, 1 sobel, 2 goodman) as tfce_sobel_z would, bit for bit. */
int tfce_ols_perm_sobel(const tfce_ols_model_t *path_a, const tfce_ols_model_t *path_b, const int32_t *d_perm,
                        int32_t P, int32_t alg, float *d_z, int64_t ldz, hipStream_t stream);

/* Replaces the ancova branch of the voxel randomise loop
 * (voxel_tfce_multiple_regression_randomise.py:79-89: calcF of X[perm],
 * negative F set to 0, square root; calcF is cynumstats.pyx:31-36):
 * d_f[p*ldf + v] = fl32(sqrt(max(F, 0))) with F = ((tss[v] - SSE)/(k-1)) /
 * (SSE/(n-k)) for the whole model of X[perm_p] (k >= 2).  d_tss[V] (FP64) is
 * the caller's total sum of squares, sum_i (y_iv - mean(y))^2 about the grand
 * mean of every element, as calcF takes it.  The contrasts of the model are
 * not used. */
int tfce_ols_perm_f(const tfce_ols_model_t *model, const int32_t *d_perm, int32_t P, const double *d_tss,
                    float *d_f, int64_t ldf, hipStream_t stream);

/* Replaces the per-variable F of glm_typeI (pyfunc.py:2282-2360) inside the
 * tm-models GLM randomise loop (tm_models_randomise.py:186-196: rows of the
 * whole design permuted by rand_array): for each block b of caller columns
 * [block_start[b], block_start[b] + block_width[b]) (host arrays),
 * F_b = ((SSE of X[perm_p] without the block) - SSE) / block_width[b] /
 * (SSE / (n - k)), written to d_f[(b * P + p) * ldf + v].  Blocks must not
 * overlap, must leave at least one column, and may not hold the intercept of
 * centred data; 1 <= n_blocks <= 8.  A one-column block's F is the square of
 * that column's t.  The contrasts of the model are not used. */
int tfce_ols_perm_partial_f(const tfce_ols_model_t *model, const int32_t *d_perm, int32_t P,
                            const int32_t *block_start, const int32_t *block_width, int32_t n_blocks,
                            float *d_f, int64_t ldf, hipStream_t stream);

/* The total sum of squares calcF takes (cynumstats.pyx:34, TSS = sum((y -
 * np.mean(y))**2) summed over rows): d_tss[v] = sum_i (Y[i][v] - m)^2 in FP64,
 * m the grand mean of all n x V values of Y (row stride ldy).  The input to
 * tfce_ols_perm_f.  Blocking. */
int tfce_total_ss(const float *d_Y, int32_t n, int64_t V, int64_t ldy, double *d_tss, hipStream_t stream);

/* Workgroup tile of the split-precision contraction (process-wide; the
 * autotuner's knob): 0 = 256 rows x 128 elements (default), 1 = 512 x 64,
 * 2 = 512 x 128, 3 = 256 x 256.  No counterpart in the reference. */
int tfce_ols_set_tile(int32_t tile);
int tfce_ols_tile(void);

/* One design, FP64 throughout (the numpy path of cynumstats.pyx:28-112 for a
 * single X): coefficients with inv(X'X) computed here, residuals, SSE and
 * se_j = fl32(sqrt(SSE/(n-k) * S[j][j])) with S = invXX_se (the invXX the
 * caller passes to tval_int, cynumstats.pyx:59-64) or inv(X'X) when NULL;
 * t = a / se.  Also the per-element sum and sum of squares of y (for calcF's
 * grand-mean TSS, :31-36).  X is host n x k row-major (n <= 2048, k <= 8);
 * outputs are device arrays [k][V] (coef, se, t), [n][V] (resid), [V], any
 * may be NULL.  Blocking. */
int tfce_ols_direct(const float *d_Y, int32_t n, int64_t V, int64_t ldy, const double *X, int32_t k,
                    const double *invXX_se, double *d_coef, float *d_se, double *d_t, double *d_resid,
                    double *d_sse, double *d_sumy, double *d_sumyy, hipStream_t stream);

/* Element-wise regressor (pyfunc.py:1024-1086 image_regression, called by
 * STEP_1_vertex_tfce_multiple_regression.py:276-289 and
 * STEP_1_voxel_tfce_multiple_regression.py:162-183): for every element v the
 * design is the shared columns R (host n x r row-major, intercept first,
 * 1 <= r <= 7) plus the element's own regressor Z[.][v]; the t-values of all
 * r + 1 coefficients (FP64 least squares, sigma2 = SSE / (n - r - 1), se =
 * sqrt(sigma2 * diag(inv(X'X))), t = a / se) go to d_t[j * ldt + v] as float,
 * j = r for the element's regressor.  Elements whose regressor holds a NaN or
 * has a population standard deviation below std_min (the reference: 0.01)
 * get 0 in every column.  Y and Z are device [n][ldy] / [n][ldz] float;
 * r + 1 < n <= 2048.  Blocking. */
int tfce_image_regression(const float *d_Y, int64_t ldy, const float *d_Z, int64_t ldz, const double *R,
                          int32_t n, int32_t r, int64_t V, double std_min, float *d_t, int64_t ldt,
                          hipStream_t stream);

/* Variance inflation image (pyfunc.py:1088-1114 image_reg_VIF): per element,
 * R2 = 1 - RSS / TSS of y on X (host n x k row-major, intercept first,
 * k <= 8, k < n <= 2048; RSS from the FP64 least-squares fit), VIF =
 * 1 / (1 - R2), NaN set to 0, written to the device array d_vif[V] (double).
 * d_Y is a device [n][ldy] array of float (y_f64 = 0) or double (y_f64 = 1);
 * TSS = sum((y - mean)^2) is formed in that type, subject by subject, as
 * numpy reduces the first axis.  Blocking. */
int tfce_image_vif(const void *d_Y, int32_t y_f64, int64_t ldy, const double *X, int32_t n, int32_t k, int64_t V,
                   double *d_vif, hipStream_t stream);

/* se[j][v] = fl32(sqrt(sigma2[v] * invXX[j][j])) for j < k (cynumstats.pyx:47-52);
 * invXX is host k x k, sigma2 and se are device arrays.  Blocking. */
int tfce_se_of_slope(const double *d_sigma2, const double *invXX, int32_t k, int64_t V, float *d_se,
                     hipStream_t stream);

/* Indirect-effect z from the path t-values, elementwise (pyfunc.py:150-161):
 * alg 0 Aroian 1/sqrt(1/tb^2 + 1/ta^2 + 1/(ta^2 tb^2)), 1 Sobel, 2 Goodman.
 * d_ta == NULL uses the scalar ta_const (mediation type Y, pyfunc.py:142). */
int tfce_sobel_z(const float *d_ta, const float *d_tb, int64_t n, int32_t alg, float ta_const,
                 float *d_z, hipStream_t stream);

/* One-call forms (blocking; they build and release the handles above).
 * ols_perm_t: t of column `contrast` of X[perm_p] (X host, n x k row-major
 * float64; every non-constant column moves, as the reference's full-row
 * permutation X[perm] does, vertex_tfce_multiple_regression_randomise.py:105-107)
 * for P permutations d_perm[P][n]; d_t[P][V].  The reference's tval_int takes
 * invXX besides X; row permutation leaves X'X unchanged, so X alone fixes it.
 * sobel_perm_z: per permutation of x, the Sobel-family z of calc_sobelz
 * (pyfunc.py:130-162, vertex_tfce_mediation_randomise.py:82-91) for mediator
 * images d_M[n][V]: medtype 0 = M, 1 = I, 2 = Y (type Y moves x and y
 * together); alg 0 aroian, 1 sobel, 2 goodman; d_z[P][V]. */
int ols_perm_t(const double *X, const float *d_Y, const int32_t *d_perm, int32_t n, int32_t k, int64_t V,
               int32_t contrast, int32_t precision, int32_t P, float *d_t, hipStream_t stream);
int sobel_perm_z(const float *d_M, int32_t n, int64_t V, const double *x, const double *y, int32_t medtype,
                 int32_t alg, const int32_t *d_perm, int32_t P, int32_t precision, float *d_z, hipStream_t stream);

/* ------------------------------------------------------------------ FWER */
/* 1 - p(FWE) of each element against the sorted (ascending) null, FP64.
 * mode 0: calculate_fweP_vertex.py:37-69 / calculate_fweP_voxel.py:23-57
 *   (elements with value > 0; idx = #(null < value); (idx-1)/N floored at 0;
 *   other elements 0).
 * mode 1: the nearest-neighbour find_nearest of tm_func.py:566-573 (used by
 *   apply_mfwer, tm_func.py:470-487) for every element. */
int tfce_fwe_p(const double *d_sorted_null, int32_t n_null, const float *d_values, int64_t V,
               int32_t mode, double *d_out, hipStream_t stream);
/* The same lookups for float64 element values (apply_mfwer's standardised
 * log-TFCE, tm_func.py:431-446). */
int tfce_fwe_p64(const double *d_sorted_null, int32_t n_null, const double *d_values, int64_t V,
                 int32_t mode, double *d_out, hipStream_t stream);

/* tfce_batch with options:
 *   TFCE_BATCH_SHARED_THRESHOLDS - every surface of a (map, sign) uses the
 *     threshold sequence of the maximum over the whole set, as when the
 *     surfaces are TFCE'd as one merged image (tm_func.py:54-123 with its
 *     merged adjacency, tm_multimodality_multisurface_regression.py:443-444);
 *     the rescale still uses each surface's own maximum (tm_func.py:83-91);
 *   TFCE_BATCH_ZERO_IN_IMAGE - that merged image also holds zeros (vertex
 *     masks in create_full_mask, tm_func.py:502-510). */
#define TFCE_BATCH_SHARED_THRESHOLDS 1
#define TFCE_BATCH_ZERO_IN_IMAGE 2
int tfce_batch_ex(const tfce_surfset_t *set, const float *d_maps, int64_t map_stride, int P,
                  int sign_mask, uint32_t flags, const float *d_vdensity, float *d_scaled, float *d_raw,
                  float *d_own, int32_t *d_status, hipStream_t stream);

/* How tfce_run / tfce_run_ex sweep one image of a symmetric surface: mode -1
 * (default) over the whole GPU for surfaces of at least 32,768 elements and in
 * one workgroup below that, 0 always in one workgroup, 1 always over the whole
 * GPU; max_workgroups > 0 caps the whole-GPU grid (0: one workgroup per
 * compute unit).  Both give the same values bit for bit; the labels debug
 * output of tfce_run_ex always takes the one-workgroup sweep.  Process-wide.
 * No counterpart in the reference. */
int tfce_set_single_image_mode(int32_t mode, int32_t max_workgroups);

/* Number of resident workgroups per CU used by the TFCE sweep (autotuner knob). */
int tfce_set_occupancy(int32_t wgs_per_cu);

/* Batched maxima (tfce_batch*, tfce_batch_max*) of symmetric surfaces of at
 * most 13,248 elements with a neighbour table run in an LDS-resident sweep
 * (the whole union-find of an image in one compute unit's LDS) unless this
 * process-wide knob is 0; the maxima are the same bit for bit either way.
 * No counterpart in the reference. */
int tfce_set_lds_sweep(int32_t enable);

/* Batched maxima of both signs (sign_mask 3) over ring-packed surfaces (closed
 * triangle meshes of at most six neighbours per vertex) run both signs of a map
 * in one workgroup, sharing one pass over the map and the vertices, unless
 * this process-wide knob is 0; the maxima are the same bit for bit either way.
 * No counterpart in the reference. */
int tfce_set_pair_sweep(int32_t enable);

/* ---------------------------------------------------- learned early stop */
/* No counterpart in the reference (nearest: the parametric null fits of
 * bin/tm_maths:532-552).  Tail model 32 -> 64 -> 64 -> 2, tanh, FP32, mapping
 * 32 standardised quantiles of the first null values to the standardised
 * 95% / 99% critical values (tools/plot_permutations.py:81-85 definition). */
int tfce_mlp_param_count(void);
int tfce_mlp_forward(const float *d_params, const float *d_x, int32_t n, float *d_y, hipStream_t stream);
/* `steps` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, step count from t0+1)
 * on MSE over minibatches d_batch_idx[steps][batch] (batch <= 128) of the
 * examples d_x[n][32] -> d_y[n][2]; params, adam_m, adam_v are updated in
 * place.  Optional d_loss[steps] (minibatch loss before each update) and
 * d_grad0 (gradient of the first step).  One workgroup; deterministic.  A
 * minibatch holding an index outside [0, n) is skipped (no update) and its
 * d_loss entry is NaN. */
int tfce_mlp_train(float *d_params, float *d_adam_m, float *d_adam_v, const float *d_x, const float *d_y,
                   int32_t n, const int32_t *d_batch_idx, int32_t steps, int32_t batch, int32_t t0, float lr,
                   float weight_decay, float *d_loss, float *d_grad0, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TFCEHIP_H */
