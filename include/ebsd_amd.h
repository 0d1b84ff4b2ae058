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

#define EBSD_ABI_VERSION 3

/* status codes */
#define EBSD_OK 0
#define EBSD_ERR_NULL (-1)        /* a required pointer is NULL            */
#define EBSD_ERR_SHAPE (-2)       /* a size or leading dimension is invalid */
#define EBSD_ERR_DTYPE (-3)       /* unsupported dtype for this call        */
#define EBSD_ERR_TOPK (-4)        /* top-k outside [1, EBSD_MAX_TOPK]       */
#define EBSD_ERR_WORKSPACE (-5)   /* workspace missing or too small         */
#define EBSD_ERR_ALIGN (-6)       /* pointer not 16-byte aligned            */
#define EBSD_ERR_SCRATCH (-7)     /* the sizes need more than the static scratch of an allocation-free call */
#define EBSD_ERR_INTERNAL (-8)    /* the library's own work plan is inconsistent (a library bug) */

#define EBSD_MAX_TOPK 16

/* element types of pattern matrices */
#define EBSD_DTYPE_F32 0
#define EBSD_DTYPE_BF16 1
#define EBSD_DTYPE_F16 2
#define EBSD_DTYPE_FP8_E4M3 3 /* OCP e4m3fn, per-row power-of-two scale in a float array */

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
 *            for FP8 output (value = stored * scale, scale = the smallest
 *            power of two with max|value| / scale <= 448) and may be NULL
 *            otherwise
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
 * ebsd_project reading the master pattern through its F32 quad table
 * (ebsd_mp_quad_table, dtype F32): one 16-byte load per bilinear sample
 * instead of four, same taps and blend, so the output equals ebsd_project's
 * bit for bit. quads (mp_rows, mp_cols, 4) f32, 16-byte aligned.
 * order: NULL, or a permutation of [0, n_ori) (int32, device) giving the
 * order in which rows are processed; it changes nothing in the output, only
 * which patterns run together (rows sorted by ebsd_view_keys read the same
 * master-pattern region at the same time, which keeps the gathers in L2).
 * Replaces: as ebsd_project.
 */
int ebsd_project_quads(const float* quads, int mp_rows, int mp_cols,
                       const float* rays, int n_pix, const float* quats,
                       int n_ori, void* out, int out_dtype, int ld_out,
                       float* out_scale, int flags, const int* order, void* stream);

/*
 * Locality key of each orientation for ebsd_project_quads' processing order:
 * keys[n] = Z-order index of the 128 x 64 cell of the side-by-side Lambert
 * image where direction (dir_x, dir_y, dir_z) (the detector's central ray)
 * lands after rotation by quats[n].
 *   quats (n_ori, 4) f32, 16-byte aligned; keys (n_ori,) int32
 * No counterpart in the reference (scheduling only).
 */
int ebsd_view_keys(const float* quats, int n_ori, float dir_x, float dir_y, float dir_z,
                   int* keys, void* stream);

/*
 * ebsd_project with a per-orientation pattern-centre shift: ray d of row n
 * is normalise(coords[d] + offsets[n]) before the rotation. coords are the
 * un-normalised sample-frame pixel positions, offsets (n_ori, 3) f32 the
 * translation a moved pattern centre applies to every pixel (the derivative
 * of geometry.py:75-89 with respect to the pattern centre times the shift).
 * Used to simulate training patterns with a jittered pattern centre.
 * No counterpart in the reference (the learned-proposal form's simulator).
 */
int ebsd_project_shifted(const float* mp, int mp_rows, int mp_cols,
                         const float* coords, int n_pix, const float* quats,
                         const float* offsets, int n_ori, void* out,
                         int out_dtype, int ld_out, float* out_scale, int flags,
                         void* stream);

/* grid_sample modes and paddings of ebsd_sample_dirs_mode */
#define EBSD_SAMPLE_BILINEAR 0
#define EBSD_SAMPLE_NEAREST 1
#define EBSD_SAMPLE_BICUBIC 2
#define EBSD_PAD_ZEROS 0
#define EBSD_PAD_BORDER 1
#define EBSD_PAD_REFLECTION 2

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
 * Master-pattern lookup at arbitrary directions in any grid_sample mode:
 * mode EBSD_SAMPLE_BILINEAR / NEAREST / BICUBIC, padding EBSD_PAD_ZEROS /
 * BORDER / REFLECTION, align_corners 0 or 1, each as torch's grid_sample
 * defines it (nearest rounds half to even; bicubic is Keys' cubic
 * convolution with A = -0.75, padding applied to each of its 16 taps).
 *   dirs (n, 3) f32; out (n,) f32
 * Replaces: ebsd/master_pattern.py:141-209 (MasterPattern.interpolate with
 * the other modes; refining.py:203-209 uses bilinear and :338-344 bicubic,
 * both with align_corners=True).
 */
int ebsd_sample_dirs_mode(const float* mp, int mp_rows, int mp_cols, const float* dirs, long long n,
                          float* out, int renorm, int mode, int padding, int align_corners, void* stream);

/*
 * ebsd_sample_dirs_mode without re-normalisation, with the gradient of each
 * value by its direction: grad[i] = d out[i] / d dirs[i] (3 floats), through
 * the Lambert map and the sampling as torch's autograd differentiates
 * grid_sample (nearest: 0; border and reflection: 0 on and beyond the
 * borders).
 *   dirs (n, 3) f32; out (n,) f32; grad (n, 3) f32
 * Replaces: the backward of ebsd/master_pattern.py:141-209 that
 * refining.py:318-365 (orientation_grad_refinement) takes through
 * MasterPattern.interpolate.
 */
int ebsd_sample_dirs_grad(const float* mp, int mp_rows, int mp_cols, const float* dirs, long long n,
                          float* out, float* grad, int mode, int padding, int align_corners, void* stream);

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

/*
 * ebsd_prep_patterns with two outputs from one read of the patterns: the
 * finished rows in f32 (out_f32, ld_f32) and the same values converted to
 * out2_dtype (EBSD_DTYPE_BF16 or EBSD_DTYPE_F16; out2, ld2), bit for bit what
 * two ebsd_prep_patterns calls would write.
 * Replaces: ebsd/indexing.py:175-187, as ebsd_prep_patterns; the second
 * output has no counterpart in the reference (operands of the projection of
 * queries onto the principal directions of a compressed search).
 */
int ebsd_prep_patterns_pair(const float* pats, int n, int ld_in, const int* pix_idx, int n_pix,
                            float* out_f32, int ld_f32, void* out2, int out2_dtype, int ld2, int flags,
                            void* stream);

/* Bytes of scratch that ebsd_sim_topk needs for these sizes. */
int ebsd_sim_topk_workspace(int m, int n, int topk, int dtype, size_t* bytes);

/*
 * Work schedule of ebsd_sim_topk for these sizes, walked on the host with the
 * kernels' own block -> work mapping (no device work). info receives
 * EBSD_SCHEDULE_INFO ints: kernel (0 = 128 x 128, 1 = 192 x 256, 2 = 192 x 128 with
 * two workgroups per CU), grid
 * (workgroups), partial lists per query, list length, query rows per tile,
 * dictionary rows per tile, query tiles, dictionary tiles, and the grouping
 * A, B, R. If given, tile_visits (query tiles x dictionary tiles, row-major)
 * receives how many workgroups score each tile pair, and list_writes (query
 * tiles x partial lists) how many write each partial list; a correct plan
 * has every entry equal to 1. EBSD_ERR_WORKSPACE when an array is too short.
 * No counterpart in the reference: it lets a host check the schedule.
 */
#define EBSD_SCHEDULE_INFO 11
int ebsd_sim_schedule(int m, int n, int topk, int dtype, int ld, int* info, int* tile_visits,
                      long long visits_len, int* list_writes, long long writes_len);

/*
 * Similarity + streaming top-k: scores S = Q * Dict^T (angular metric, FP32
 * accumulation; FP8 scores are multiplied by q_scale[i] * d_scale[j], each
 * scale taken as its power of two: only the exponent is used, as the E8M0
 * block scale of the FP8 matrix instructions) are
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
 * Allocation-free form of ebsd_sim_topk: same arguments and results, with the
 * partial top-k lists in a static device buffer of EBSD_NOALLOC_SCRATCH_BYTES
 * per GPU (held from the library's load until the process ends) instead of a
 * caller workspace, so the call makes no allocation: for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). Sizes whose partial lists exceed the buffer (FP32 with
 * top-k > 1: m > 65,536 queries; BF16 / FP16 / FP8 on the 192 x 256 kernel:
 * m > 262,144; FP8 rows of 256 bytes on the 192 x 128 two-workgroup kernel:
 * m > 524,288) return EBSD_ERR_SCRATCH before any device work. Calls that
 * overlap on one GPU share the buffer and race: serialise them.
 * No counterpart in the reference.
 */
#define EBSD_NOALLOC_SCRATCH_BYTES (64u << 20)
int ebsd_sim_topk_noalloc(const void* q, const float* q_scale, int m, const void* dict,
                          const float* d_scale, int n, int ld, int dtype, int topk,
                          float* top_val, int* top_idx, int index_base, int merge, void* stream);

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
 * Detector noise: out[n, d] = Poisson(mean_counts[n] * clean[n, d] /
 * mean_d(clean[n, :])) / mean_counts[n] + sigma * N(0, 1), for the first n_pix
 * columns of each row (mean_counts[n] <= 0: no Poisson stage, out = clean +
 * sigma * N(0, 1)); Philox4x32-10 counter-based random numbers keyed by
 * seed, counted by (offset, row, column), so the result does not depend on
 * the launch. No counterpart in the reference (synthetic workload and
 * learned-proposal training data).
 */
int ebsd_poisson_gaussian_noise(const float* clean, int n, int n_pix, int ld_in,
                                float* out, int ld_out, const float* mean_counts,
                                float sigma, unsigned long long seed,
                                unsigned long long offset, void* stream);

/*
 * Dense product C = alpha * A * B^T + beta * C (FP32 accumulation).
 *   a (m, ld), b (n, ld) of dtype F32 / BF16 / F16, row-major, ld a multiple
 *   of ebsd_pad_multiple(dtype) with zero padding; c (m, ldc) f32.
 * Replaces: ebsd/indexing.py:401, :421, :452 (torch.matmul of patterns with
 * the PCA component matrix; b holds the components as rows).
 */
int ebsd_gemm_nt(const void* a, int m, const void* b, int n, int ld, int dtype,
                 float* c, int ldc, float alpha, float beta, void* stream);

/*
 * ebsd_gemm_nt with the K loop split `splits` ways into partial products
 * (workspace: splits * m * n floats, sized by ebsd_gemm_nt_splitk_workspace)
 * that are then summed in a fixed order: for long reductions with small
 * outputs (weight gradients). No counterpart in the reference.
 */
int ebsd_gemm_nt_splitk_workspace(int m, int n, int splits, size_t* bytes);
int ebsd_gemm_nt_splitk(const void* a, int m, const void* b, int n, int ld,
                        int dtype, float* c, int ldc, float alpha, float beta,
                        int splits, void* workspace, size_t workspace_bytes,
                        void* stream);

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

/*
 * Learned orientation proposal: pieces of the convolutional network's forward
 * and backward passes (3x3 stride-2 padding-1 convolutions on NHWC float32
 * activations run as im2col + ebsd_gemm_nt). None of these has a counterpart
 * in the reference, which has no network; they serve the learned-proposal
 * form, whose supports in the reference are orientations.py:940-963 (6D ->
 * rotation) and laue_generators.py:54-249 (symmetry operators).
 *   ebsd_nn_im2col     x (b, h, w, c) f32 -> col (b*ho*wo, ld) F32/BF16,
 *                      column (ky*3 + kx)*c + ch, values times scale, zero
 *                      padding to ld; transposed != 0 stores col^T
 *                      (9c, ld), ld >= b*ho*wo, zero columns beyond
 *   ebsd_nn_conv_c1    first convolution (one input channel) done directly:
 *                      y (b*ho*wo, c_out) f32 = ReLU(bias + w . taps), taps
 *                      scale * x rounded to dtype (F32/BF16), weight
 *                      (c_out, ld_w) of dtype, columns ky*3 + kx; the values
 *                      of im2col + GEMM + bias_act up to the fp32 sum order;
 *                      c_out a multiple of 4, <= 256, y 16-byte aligned
 *   ebsd_nn_conv_gemm  convolution as an implicit GEMM with bias + ReLU: the
 *                      im2col operand of x (b, h, w, c) f32 is built in LDS
 *                      per K stage, never stored; y (b*ho*wo, c_out) f32
 *                      equals ebsd_nn_im2col + ebsd_gemm_nt_bias_act bit for
 *                      bit. weight (c_out, ld_w) of dtype F32/BF16, ld_w >= 9c
 *                      a multiple of ebsd_pad_multiple(dtype), zero beyond 9c;
 *                      c a multiple of 4 (F32) / 8 (BF16); x, weight 16-byte
 *                      aligned
 *   ebsd_gemm_nt_bias_act  c = act(a b^T + bias[col]) (ebsd_gemm_nt's operand
 *                      rules; act = ReLU if relu): GEMM + bias_act in one pass
 *   ebsd_nn_cast       dst (rows, ld_dst) F32/BF16 = src (rows, cols) f32,
 *                      zero columns beyond cols
 *   ebsd_nn_col2im     dcol (b*ho*wo, ld) f32 -> dx (b, h, w, c) f32 (gather)
 *   ebsd_nn_bias_act   y[r, :cols] = act(y + bias), act = ReLU if relu
 *   ebsd_nn_relu_bwd   dy *= (y > 0), n elements
 *   ebsd_nn_colsum     out[c] = sum_r x[r, c] (deterministic two-level sum)
 *   ebsd_nn_transpose  dst (cols, ld_dst) = src^T (F32/BF16), zero rows beyond
 *   ebsd_nn_geodesic_loss  per sample: 6D -> R (Gram-Schmidt), loss = min over
 *                      the n_ops symmetry rotations G of the geodesic angle
 *                      between R and G R(q_true); grad = d(mean loss)/d(6D)
 *   ebsd_nn_adam       Adam step with bias correction on FP32 weights, also
 *                      writing a BF16 copy when w_bf16 != NULL
 *   ebsd_nn_rot6d_to_quat  6D -> the loss's rotation R -> unit quaternion
 *                      (w >= 0) with quat_matrix(q) = R, (b, 4) f32
 */
int ebsd_nn_conv_c1(const float* x, int b, int h, int w, float scale,
                    const void* weight, int ld_w, int dtype, const float* bias,
                    int c_out, float* y, void* stream);
int ebsd_nn_conv_gemm(const float* x, int b, int h, int w, int c,
                      const void* weight, int ld_w, int c_out, int dtype,
                      const float* bias, float* y, void* stream);
int ebsd_gemm_nt_bias_act(const void* a, int m, const void* b, int n, int ld,
                          int dtype, float* c, int ldc, const float* bias,
                          int relu, void* stream);
int ebsd_nn_im2col(const float* x, int b, int h, int w, int c, float scale,
                   void* col, int ld, int dtype, int transposed, void* stream);
int ebsd_nn_cast(const float* src, int rows, int cols, int ld_src, void* dst,
                 int ld_dst, int dtype, void* stream);
int ebsd_nn_col2im(const float* dcol, int ld, int b, int h, int w, int c,
                   float* dx, void* stream);
int ebsd_nn_bias_act(float* y, int rows, int cols, int ld, const float* bias,
                     int relu, void* stream);
int ebsd_nn_relu_bwd(float* dy, const float* y, long long n, void* stream);
int ebsd_nn_colsum_workspace(int rows, int cols, size_t* bytes);
int ebsd_nn_colsum(const float* x, int rows, int cols, int ld, float* out,
                   void* workspace, size_t workspace_bytes, void* stream);
int ebsd_nn_transpose(const float* src, int rows, int cols, int ld_src,
                      void* dst, int ld_dst, int dtype, void* stream);
int ebsd_nn_geodesic_loss(const float* z6, int ldz, int b, const float* q_true,
                          const float* ops, int n_ops, float* loss, float* grad,
                          int ldg, void* stream);
int ebsd_nn_adam(float* w, const float* g, float* m, float* v, long long n,
                 float lr, float beta1, float beta2, float eps, int step,
                 float weight_decay, void* w_bf16, void* stream);
int ebsd_nn_rot6d_to_quat(const float* z6, int ldz, int b, float* q, void* stream);

/*
 * Quad table of a master pattern for ebsd_local_refine: quads (mp_rows,
 * mp_cols, 4) of dtype F32 or F16, 16-byte aligned, texel (y, x) = the
 * bilinear taps mp[y][x], mp[y][x1], mp[y1][x], mp[y1][x1] with
 * x1 = min(x + 1, cols - 1), y1 = min(y + 1, rows - 1). No counterpart in the
 * reference (a layout of the master pattern that ebsd/master_pattern.py:197-206
 * samples).
 */
int ebsd_mp_quad_table(const float* mp, int mp_rows, int mp_cols, void* quads,
                       int dtype, void* stream);

/*
 * Local-dictionary refinement (learned-proposal form): for each query row i
 * (unit-norm f32, n_pix values, row stride ld) and each of the n_grid small
 * rotations g_j, the candidate orientation qu_prod(g_j, centers[i]) is
 * projected on the fly (rotation, side-by-side Lambert, bilinear border
 * lookup of the master pattern through its quad table) and scored by the
 * cosine between the zero-mean candidate pattern and the query; the best
 * `topk` per query are written sorted descending (ties: lower j), out_idx
 * holding j (-1 for empty slots). No candidate pattern is stored.
 *   quads (mp_rows, mp_cols, 4) of quad_dtype (F32 or F16) from
 *   ebsd_mp_quad_table, fewer than 2^24 texels
 *   rays (n_pix, 3) f32; centers (m, 4) f32 proposals
 *   grid (n_grid, 4) f32 local rotations, n_grid <= 4096
 *   out_val (m, topk) f32; out_idx (m, topk) int32
 * Replaces (as a fused kernel): the candidate generation and scoring of
 * ebsd/refining.py:104-135 (local grid composed with the current orientation,
 * projected and dotted with the experimental pattern).
 */
int ebsd_local_refine(const void* quads, int quad_dtype, int mp_rows, int mp_cols,
                      const float* rays, int n_pix, const float* q, int m, int ld,
                      const float* centers, const float* grid, int n_grid,
                      int topk, float* out_val, int* out_idx, void* stream);

/*
 * Rescoring of listed local candidates: for each query row i and each of its
 * k candidates cand[i][c] (grid indices, -1 = none), the cosine of
 * qu_prod(grid[cand[i][c]], centers[i]) projected as in ebsd_local_refine,
 * against the query at this call's resolution (rays, n_pix). Typically the
 * top-k of ebsd_local_refine on binned patterns, rescored on the full
 * patterns. Writes best_val[i] / best_idx[i] (the best candidate, higher
 * score then lower list position; -inf / -1 when no candidate is valid) and,
 * if out_val is not NULL, every score out_val (m, k) (-inf for invalid
 * entries).
 *   cand (m, k) int32, 1 <= k <= 64; other operands as in ebsd_local_refine
 *   (no limit on n_grid here)
 * No counterpart in the reference: its refinement (ebsd/refining.py:104-135)
 * scores candidates at the virtually binned resolution only; this is the
 * same candidate projection and cosine, at the detector's own size.
 */
int ebsd_local_rescore(const void* quads, int quad_dtype, int mp_rows, int mp_cols,
                       const float* rays, int n_pix, const float* q, int m, int ld,
                       const float* centers, const float* grid, int n_grid,
                       const int* cand, int k, float* out_val, float* best_val,
                       int* best_idx, void* stream);

/*
 * One iteration of orientation grid refinement for m patterns: for pattern
 * i and each local rotation g_j of its level's grid
 * (grids[level[i]] (n_grid, 4) f32, the cubochoric local grid shrunk
 * level[i] times, as unit quaternions), the candidate qu_std(g_j * c_i) is
 * moved to the fundamental zone (right product with the first of the n_ops
 * symmetry operators `ops` (n_ops, 4) f32 that maximises |w|, then w >= 0),
 * projected through the bilinear, border, align_corners=True lookup of the
 * master pattern (its quad table, mp_rows x mp_cols, F32 or F16; the
 * virtually binned master pattern when the patterns are binned) along
 * `rays` (n_pix, 3) f32, and scored by the dot product of the zero-mean
 * candidate with the pattern row q (zero-mean, not normalised). Writes the
 * best score best_val (m,), its grid index best_idx (m,) (ties: lower j) and
 * its reduced quaternion best_q (m, 4).
 *   q (m, ld) f32; centers (m, 4) f32; level (m,) int32 in [0, n_levels)
 *   (clamped to that range)
 * Replaces: the body of the iteration loop of refining.py:172-237
 * (orientation_grid_refinement: cu2qu of the grid is done once per level by
 * the caller; qu_prod, ori_to_fz_laue, qu_apply, MasterPattern.interpolate
 * with align_corners=True, the einsum dot products and torch.max fused).
 */
int ebsd_grid_refine_step(const void* quads, int quad_dtype, int mp_rows, int mp_cols,
                          const float* rays, int n_pix, const float* q, int m, int ld,
                          const float* centers, const float* grids, int n_grid,
                          int n_levels, const int* level, const float* ops, int n_ops,
                          float* best_val, int* best_idx, float* best_q, void* stream);

/*
 * NLPAR (non-local pattern averaging) of a 2D scan x (scan_h, scan_w, n_pix)
 * f32, window radius k_rad in [1, 5], offsets (n_off, 2) int32 (dy, dx) with
 * |dy|, |dx| <= k_rad, n_off = (2 k_rad + 1)^2, in the caller's order
 * (clamped to the window). Points outside the scan are zero patterns.
 *   ebsd_nlpar_distances  dists (scan_h * scan_w, n_off) f32:
 *                         sum over pixels of (x[p + offset] - x[p])^2
 *   ebsd_nlpar_average    out (scan_h, scan_w, n_pix) f32:
 *                         sum over offsets, in order, of x[p + offset] *
 *                         weights[p][offset] (weights (points, n_off) f32)
 * Replaces: the two loops over the window of preprocessing/nlpar.py:105-117
 * (squared differences) and :187-191 (weighted average); the weights between
 * them (nlpar.py:119-184) are computed by the caller.
 */
int ebsd_nlpar_distances(const float* x, int scan_h, int scan_w, int n_pix, int k_rad,
                         const int* offsets, float* dists, void* stream);
int ebsd_nlpar_average(const float* x, int scan_h, int scan_w, int n_pix, int k_rad,
                       const int* offsets, const float* weights, float* out,
                       void* stream);

/*
 * CLAHE of n patterns x (n, h, w) f32 (values in [0, 1]; others are counted
 * in the end bins) on a grid_h x grid_w tile grid (grid_h * grid_w <= 64)
 * with n_bins <= 64 and clip_limit (<= 0: no clipping): reflect padding to a
 * multiple of the grid, per-tile histograms of trunc(v (n_bins - 1)),
 * clipping at max(clip_limit * tile_pixels // n_bins, 1) with the excess
 * redistributed, clamped CDFs, and each output pixel the trilinear
 * (row, column, value) lookup of the CDF volume (align_corners=False,
 * border). out (n, h, w) f32; the padding must be smaller than h and w.
 * Replaces: preprocessing/clahe.py:18-130 (clahe_grayscale).
 */
int ebsd_clahe(const float* x, int n, int h, int w, int grid_h, int grid_w, int n_bins,
               float clip_limit, float* out, void* stream);

/*
 * Per-pattern normalisation in place of x (n, n_pix) f32: mode 0 minmax
 * (x - min) / (1e-4 + max - min), 1 zero mean, 2 standard (zero mean, unit
 * unbiased standard deviation).
 * Replaces: ebsd/experiment_pats.py:256-281 (normalize_per_pattern).
 */
int ebsd_normalize_patterns(float* x, int n, int n_pix, int mode, void* stream);

/*
 * Mutual-information maps of the pattern-centre fitter's "mi" metric: for
 * each of n_maps (dictionary pattern, experimental pattern) pairs, corr holds
 * the 5 x 5 circular cross-correlations (n_maps, 5, 5, n_pix) f32 of their
 * level-indicator images (channel 0 the foreground, 1-4 the value bins of
 * trunc(4 v)); out (n_maps, n_pix) f32 receives H_dst + H_src - H_joint per
 * shift, counts below 1 dropped, NaN / infinite / negative values set to 0.
 * Replaces: utils/mutual_information.py:304-673 (cmif, bins 4 + 4,
 * constant_marginal=False, norm_mode "none") after its FFT products, as
 * ebsd/fitting.py:486-496 calls it.
 */
int ebsd_cmif_maps(const float* corr, long long n_maps, int n_pix, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EBSD_AMD_H */
