/*
 * ebsd_amd.h - C ABI of the MI355X (gfx950) dictionary-indexing library
 * (libebsd_amd.so).
 *
 * Conventions
 *   - Every buffer is a device pointer owned by the caller; the library never
 *     allocates or frees caller memory. Scratch space is passed in by the
 *     caller, sized by the matching *_workspace query.
 *   - Every compute call takes a HIP stream (as void*, NULL = default stream)
 *     and is stream ordered. No call synchronises the device.
 *   - Return value: 0 = success, negative = argument error (EBSD_ERR_*),
 *     positive = a hipError_t raised by a launch.
 *   - Dictionary indices are int32 on the device (dictionaries < 2^31 rows).
 *   - Pattern matrices are row-major with a leading dimension `ld` that is a
 *     multiple of ebsd_pad_multiple(dtype); columns [n_pix, ld) hold zeros.
 *
 * Each entry point names the part of ZacharyVarley/ebsdtorch (v0.0.64) that it
 * replaces; paths are relative to the ebsdtorch repository root.
 */
#ifndef EBSD_AMD_H
#define EBSD_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EBSD_ABI_VERSION 1

/* status codes */
#define EBSD_OK 0
#define EBSD_ERR_NULL (-1)        /* a required pointer is NULL            */
#define EBSD_ERR_SHAPE (-2)       /* a size or leading dimension is invalid */
#define EBSD_ERR_DTYPE (-3)       /* unsupported dtype for this call        */
#define EBSD_ERR_TOPK (-4)        /* top-k outside [1, EBSD_MAX_TOPK]       */
#define EBSD_ERR_WORKSPACE (-5)   /* workspace missing or too small         */
#define EBSD_ERR_ALIGN (-6)       /* pointer not 16-byte aligned            */

#define EBSD_MAX_TOPK 16

/* element types of pattern matrices */
#define EBSD_DTYPE_F32 0
#define EBSD_DTYPE_BF16 1
#define EBSD_DTYPE_F16 2
#define EBSD_DTYPE_FP8_E4M3 3 /* OCP e4m3fn, per-row scale in a float array */

/* flags of ebsd_project / ebsd_prep_patterns */
#define EBSD_FLAG_ZERO_MEAN 1   /* subtract the per-pattern mean            */
#define EBSD_FLAG_UNIT_NORM 2   /* divide by the per-pattern L2 norm         */
#define EBSD_FLAG_RENORM_DIRS 4 /* ebsd_project: re-normalise rotated rays   */

/* ABI version of the loaded library (EBSD_ABI_VERSION it was built with). */
int ebsd_abi_version(void);

/* Static description of a status code. */
const char* ebsd_status_string(int status);

/* Leading-dimension granularity (in elements) required for `dtype`. */
int ebsd_pad_multiple(int dtype);

/*
 * Fused dictionary projection: for every orientation n and detector pixel d,
 * rotate ray d by quaternion n, map the direction to the side-by-side
 * Rosca-Lambert square, sample the master pattern bilinearly (border padding,
 * align_corners=False), then (flags) zero-mean and/or unit-norm the pattern
 * and store row n of `out` in `out_dtype`.
 *   mp       (mp_rows, mp_cols) f32 side-by-side Lambert image, northern
 *            hemisphere in the left half, southern in the right half
 *            (mp_cols = 2 * mp_rows, or the average-pooled shape of one)
 *   rays     (n_pix, 3) f32 sample-frame directions
 *   quats    (n_ori, 4) f32 unit quaternions (w, x, y, z)
 *   out      (n_ori, ld_out) of out_dtype; out_scale (n_ori,) f32 is written
 *            for FP8 output (value = stored * scale) and may be NULL otherwise
 * Replaces: ebsd/indexing.py:117-143 (qu_apply + MasterPattern.interpolate +
 * mean subtraction), s2_and_so3/quaternions.py:186-213 (qu_apply),
 * s2_and_so3/sphere.py:145-185 (rosca_lambert_side_by_side),
 * ebsd/master_pattern.py:141-209 (interpolate -> grid_sample).
 */
int ebsd_project(const float* mp, int mp_rows, int mp_cols, const float* rays,
                 int n_pix, const float* quats, int n_ori, void* out,
                 int out_dtype, int ld_out, float* out_scale, int flags,
                 void* stream);

/*
 * Master-pattern lookup at arbitrary directions: out[i] = bilinear sample
 * (border, align_corners=False) of the side-by-side Lambert master pattern at
 * direction dirs[i] (re-normalised first when renorm != 0).
 *   dirs (n, 3) f32; out (n,) f32
 * Replaces: ebsd/master_pattern.py:141-209 (MasterPattern.interpolate with
 * mode="bilinear", padding_mode="border", align_corners=False).
 */
int ebsd_sample_dirs(const float* mp, int mp_rows, int mp_cols, const float* dirs,
                     long long n, float* out, int renorm, void* stream);

/*
 * Experimental-pattern preparation: gather pixels (pix_idx may be NULL for
 * the identity), optionally zero-mean / unit-norm each row, convert to
 * out_dtype and zero-pad to ld_out.
 *   pats     (n, ld_in) f32, row stride ld_in elements
 *   pix_idx  (n_pix,) int32 column indices into a row of pats, or NULL
 * Replaces: ebsd/indexing.py:175-187 (get_patterns + signal mask + mean
 * subtraction of each query chunk) and the .to(match_dtype) of
 * utils/knn.py:234.
 */
int ebsd_prep_patterns(const float* pats, int n, int ld_in, const int* pix_idx,
                       int n_pix, void* out, int out_dtype, int ld_out,
                       float* out_scale, int flags, void* stream);

/* Bytes of scratch that ebsd_sim_topk needs for these sizes. */
int ebsd_sim_topk_workspace(int m, int n, int topk, int dtype, size_t* bytes);

/*
 * Similarity + streaming top-k: scores S = Q * Dict^T (angular metric, FP32
 * accumulation; FP8 scores are multiplied by q_scale[i] * d_scale[j]) are
 * reduced on the fly to the best `topk` dictionary rows per query; the score
 * matrix is never written. Indices are index_base + row. With merge != 0 the
 * existing (top_val, top_idx) state, sorted descending, is merged with the
 * new chunk; with merge == 0 it is overwritten. Output rows are sorted by
 * descending score, ties broken by the lower index.
 *   q        (m, ld) of dtype;  dict (n, ld) of dtype
 *   top_val  (m, topk) f32;      top_idx (m, topk) int32
 * Replaces: utils/knn.py:80-103 (knn_batch, metric "angular"),
 * utils/knn.py:163-266 (ChunkedKNN.set_data_chunk / query_all / query_chunk)
 * and utils/knn.py:268-287 (retrieve_topk sort).
 */
int ebsd_sim_topk(const void* q, const float* q_scale, int m, const void* dict,
                  const float* d_scale, int n, int ld, int dtype, int topk,
                  float* top_val, int* top_idx, int index_base, int merge,
                  void* workspace, size_t workspace_bytes, void* stream);

/*
 * FP32 re-scoring of low-precision candidates: for each query row i, the
 * n_cand dictionary rows cand[i, :] (int32; entries outside [0, n) are
 * empty slots) are scored exactly in FP32 against q[i, :] and the best
 * `topk` are written, sorted descending (ties: lower index first).
 *   q (m, ld) f32; dict (n, ld) f32; cand (m, n_cand) int32
 *   out_val (m, topk) f32; out_idx (m, topk) int32
 * No counterpart in the reference: it is the second stage of the precision
 * form of utils/knn.py:80-103, which restores the FP32 ranking of the
 * candidates a BF16/FP16/FP8 similarity pass proposes.
 */
int ebsd_rescore(const float* q, int m, const float* dict, int n, int ld, const int* cand,
                 int n_cand, int topk, float* out_val, int* out_idx, void* stream);

/*
 * Dense product C = alpha * A * B^T + beta * C (FP32 accumulation).
 *   a (m, ld), b (n, ld) of dtype F32 / BF16 / F16, row-major, ld a multiple
 *   of ebsd_pad_multiple(dtype) with zero padding; c (m, ldc) f32.
 * Replaces: ebsd/indexing.py:401, :421, :452 (torch.matmul of patterns with
 * the PCA component matrix; b holds the components as rows).
 */
int ebsd_gemm_nt(const void* a, int m, const void* b, int n, int ld, int dtype,
                 float* c, int ldc, float alpha, float beta, void* stream);

/* Bytes of scratch that ebsd_cov_update needs for a batch of n rows of d. */
int ebsd_cov_update_workspace(int n, int d, size_t* bytes);

/*
 * Streaming covariance update (Welford / Chan batched form): merges the
 * batch x (n, ld_x) f32 (first d columns used) into the running mean (d,)
 * and aggregate cov (d, ld_cov) f32, given n_seen rows merged before.
 *   cov += (x - mean_old)^T (x - mean_new); mean <- mean_new
 * Only the block tiles (128 x 128) on or above the diagonal of cov are
 * updated; the lower triangle is the transpose of the upper one.
 * workspace: ebsd_cov_update_workspace bytes, 256-byte aligned.
 * Replaces: utils/pca_covmat.py:53-91 (update_covmat) and :120-131
 * (OnlineCovMatrix.forward).
 */
int ebsd_cov_update(const float* x, int n, int d, int ld_x, float* mean,
                    float* cov, int ld_cov, long long n_seen, void* workspace,
                    size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EBSD_AMD_H */
