// Kernels of the learned orientation-proposal network (gfx950).
//
// The network is 4 convolutions (3x3, stride 2, padding 1, ReLU) and a
// 2-layer MLP producing a 6D rotation. Activations are NHWC float32;
// convolutions run as im2col + MFMA GEMM (ebsd_gemm_nt), so this file holds
// the data-movement and element-wise pieces around the GEMMs: im2col /
// col2im, bias + ReLU forward and backward, column sums (bias gradients),
// transpose-and-cast (GEMM operand layouts), the symmetric geodesic loss and
// the Adam update.

#include "common.h"

#include <hip/hip_bf16.h>

namespace {

template <typename T>
__device__ __forceinline__ T cast_out(float v);
template <>
__device__ __forceinline__ float cast_out<float>(float v) { return v; }
template <>
__device__ __forceinline__ __hip_bfloat16 cast_out<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

// col[(b*Ho + oy)*Wo + ox][(ky*3 + kx)*C + c] = scale * X[b][2 oy + ky - 1][2 ox + kx - 1][c]
// (zero outside the image and in the padding columns [9C, ld)). With
// TRANS the matrix is stored transposed, colT[k][row] with rows padded to ld
// (the layout of the weight-gradient product, whose reduction runs over rows).
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void im2col_kernel(const float* __restrict__ X, int B, int H, int W, int C,
                                                     int Ho, int Wo, float scale, T* __restrict__ col, int ld) {
  const int64_t n_rows = (int64_t)B * Ho * Wo;
  const int64_t total = TRANS ? (int64_t)9 * C * ld : n_rows * ld;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
  const int k = TRANS ? (int)(idx / ld) : (int)(idx % ld);
  const int64_t row = TRANS ? idx % ld : idx / ld;
  float v = 0.0f;
  if (k < 9 * C && row < n_rows) {
    const int c = k % C;
    const int kk = k / C;
    const int ky = kk / 3, kx = kk % 3;
    const int ox = (int)(row % Wo);
    const int oy = (int)((row / Wo) % Ho);
    const int b = (int)(row / ((int64_t)Wo * Ho));
    const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
    if (y >= 0 && y < H && x >= 0 && x < W) v = scale * X[(((int64_t)b * H + y) * W + x) * C + c];
  }
  col[idx] = cast_out<T>(v);
  }
}

// dst[r][c] = src[r][c] for c < cols, 0 for cols <= c < ld_dst.
template <typename T>
__global__ __launch_bounds__(256) void cast_kernel(const float* __restrict__ src, int rows, int cols, int ld_src,
                                                   T* __restrict__ dst, int ld_dst) {
  const int64_t total = (int64_t)rows * ld_dst;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int c = (int)(idx % ld_dst);
    const int64_t r = idx / ld_dst;
    dst[idx] = cast_out<T>(c < cols ? src[r * ld_src + c] : 0.0f);
  }
}

// dX[b][y][x][c] = sum over the (ky, kx) whose window covers (y, x) of
// dcol[(b*Ho + oy)*Wo + ox][(ky*3 + kx)*C + c]; a gather, so no atomics.
__global__ __launch_bounds__(256) void col2im_kernel(const float* __restrict__ dcol, int ld, int B, int H, int W,
                                                     int C, int Ho, int Wo, float* __restrict__ dX) {
  const int64_t total = (int64_t)B * H * W * C;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
  const int c = (int)(idx % C);
  const int x = (int)((idx / C) % W);
  const int y = (int)((idx / ((int64_t)C * W)) % H);
  const int b = (int)(idx / ((int64_t)C * W * H));
  float s = 0.0f;
#pragma unroll
  for (int ky = 0; ky < 3; ++ky) {
    const int ty = y + 1 - ky;  // = 2 oy
    if (ty < 0 || (ty & 1)) continue;
    const int oy = ty >> 1;
    if (oy >= Ho) continue;
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int tx = x + 1 - kx;
      if (tx < 0 || (tx & 1)) continue;
      const int ox = tx >> 1;
      if (ox >= Wo) continue;
      s += dcol[(((int64_t)b * Ho + oy) * Wo + ox) * ld + (ky * 3 + kx) * C + c];
    }
  }
  dX[idx] = s;
  }
}

// Y[r][c] = relu(Y[r][c] + bias[c]) in place (relu optional).
__global__ __launch_bounds__(256) void bias_act_kernel(float* __restrict__ Y, int rows, int cols, int ld,
                                                       const float* __restrict__ bias, int relu) {
  const int64_t total = (int64_t)rows * cols;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int c = (int)(idx % cols);
    const int64_t r = idx / cols;
    float v = Y[r * ld + c] + bias[c];
    if (relu) v = fmaxf(v, 0.0f);
    Y[r * ld + c] = v;
  }
}

// dZ[r][c] = dY[r][c] * (Y[r][c] > 0) (Y post-activation), in place on dY.
__global__ __launch_bounds__(256) void relu_bwd_kernel(float* __restrict__ dY, const float* __restrict__ Y,
                                                       int64_t n) {
  const int64_t total = n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    if (!(Y[idx] > 0.0f)) dY[idx] = 0.0f;
  }
}

// out[c] (+)= sum_r X[r][c], two levels: per-chunk partials then a fixed-
// order sum, so the result is deterministic.
__global__ __launch_bounds__(256) void colsum_partial_kernel(const float* __restrict__ X, int rows, int cols,
                                                             int ld, int chunk, float* __restrict__ part) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  const int r0 = blockIdx.y * chunk;
  const int r1 = min(r0 + chunk, rows);
  float s = 0.0f;
  for (int r = r0; r < r1; ++r) s += X[(int64_t)r * ld + c];
  part[(int64_t)blockIdx.y * cols + c] = s;
}

__global__ __launch_bounds__(256) void colsum_final_kernel(const float* __restrict__ part, int n_parts, int cols,
                                                           float* __restrict__ out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int p = 0; p < n_parts; ++p) s += part[(int64_t)p * cols + c];
  out[c] = s;
}

// dst[j][i] = src[i][j] for i < rows, j < cols; zero for rows <= i < ld_dst
// (dst is (cols, ld_dst) of T).
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ src, int rows, int cols,
                                                        int ld_src, T* __restrict__ dst, int ld_dst) {
  __shared__ float tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int ii = ty; ii < 32; ii += 8) {
    const int i = i0 + ii, j = j0 + tx;
    tile[ii][tx] = (i < rows && j < cols) ? src[(int64_t)i * ld_src + j] : 0.0f;
  }
  __syncthreads();
  for (int jj = ty; jj < 32; jj += 8) {
    const int j = j0 + jj, i = i0 + tx;
    if (j < cols && i < ld_dst) dst[(int64_t)j * ld_dst + i] = cast_out<T>(tile[tx][jj]);
  }
}

// ---- 6D rotation, symmetric geodesic loss -------------------------------

__device__ __forceinline__ float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Unit quaternion (w, x, y, z) -> rotation matrix, active convention.
__device__ __forceinline__ void quat_matrix(const float* q, float* R) {
  const float w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z); R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y); R[7] = 2 * (y * z + w * x); R[8] = 1 - 2 * (x * x + y * y);
}

// Per sample: rows b1, b2, b3 of R from the 6D vector (a1, a2) by
// Gram-Schmidt (orientations.py:940-963, zh2om); the loss is the smallest
// geodesic angle between R and G R_true over the Laue rotations G (a pattern
// is unchanged when its orientation is multiplied by a crystal symmetry on
// the left):
//   theta_g = acos(clamp((tr(R^T G R_true) - 1) / 2)), loss = min_g theta_g,
// averaged over the batch. The gradient with respect to the 6D input is
// written to grad (already divided by the batch size). The clamp keeps the
// acos derivative finite at theta = 0 and pi.
__global__ __launch_bounds__(256) void geodesic_loss_kernel(const float* __restrict__ z6, int ldz, int B,
                                                            const float* __restrict__ q_true,
                                                            const float* __restrict__ ops, int n_ops,
                                                            float* __restrict__ loss, float* __restrict__ grad,
                                                            int ldg) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= B) return;
  const float* a = z6 + (int64_t)n * ldz;
  float a1[3] = {a[0], a[1], a[2]}, a2[3] = {a[3], a[4], a[5]};
  const float n1 = sqrtf(dot3(a1, a1)) + 1e-12f;
  float b1[3] = {a1[0] / n1, a1[1] / n1, a1[2] / n1};
  const float p = dot3(b1, a2);
  float u2[3] = {a2[0] - p * b1[0], a2[1] - p * b1[1], a2[2] - p * b1[2]};
  const float n2 = sqrtf(dot3(u2, u2)) + 1e-12f;
  float b2[3] = {u2[0] / n2, u2[1] / n2, u2[2] / n2};
  float b3[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
  // R rows: b1, b2, b3
  float Rt[9];
  quat_matrix(q_true + 4 * (int64_t)n, Rt);
  float best = 1e30f, best_c = 0.0f;
  float M[9];  // G R_true of the best operator
  for (int g = 0; g < n_ops; ++g) {
    float Rg[9], T[9];
    quat_matrix(ops + 4 * g, Rg);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        T[3 * i + j] = Rg[3 * i] * Rt[j] + Rg[3 * i + 1] * Rt[3 + j] + Rg[3 * i + 2] * Rt[6 + j];
    // tr(R^T T) = sum_ij R_ij T_ij
    const float tr = b1[0] * T[0] + b1[1] * T[1] + b1[2] * T[2] + b2[0] * T[3] + b2[1] * T[4] + b2[2] * T[5] +
                     b3[0] * T[6] + b3[1] * T[7] + b3[2] * T[8];
    const float c = fminf(fmaxf(0.5f * (tr - 1.0f), -1.0f + 1e-6f), 1.0f - 1e-6f);
    const float th = acosf(c);
    if (th < best) {
      best = th;
      best_c = c;
#pragma unroll
      for (int i = 0; i < 9; ++i) M[i] = T[i];
    }
  }
  loss[n] = best;
  // dL/dR = dtheta/dc * dc/dR = (-1 / sqrt(1 - c^2)) * 0.5 * M, divided by B
  const float s = -0.5f / sqrtf(1.0f - best_c * best_c) / (float)B;
  float g1[3] = {s * M[0], s * M[1], s * M[2]};
  float g2[3] = {s * M[3], s * M[4], s * M[5]};
  float g3[3] = {s * M[6], s * M[7], s * M[8]};
  // b3 = b1 x b2: dL/db1 += b2 x g3, dL/db2 += g3 x b1
  g1[0] += b2[1] * g3[2] - b2[2] * g3[1];
  g1[1] += b2[2] * g3[0] - b2[0] * g3[2];
  g1[2] += b2[0] * g3[1] - b2[1] * g3[0];
  g2[0] += g3[1] * b1[2] - g3[2] * b1[1];
  g2[1] += g3[2] * b1[0] - g3[0] * b1[2];
  g2[2] += g3[0] * b1[1] - g3[1] * b1[0];
  // b2 = u2 / |u2|: dL/du2 = (g2 - (g2.b2) b2) / n2
  const float g2b2 = dot3(g2, b2);
  float gu[3] = {(g2[0] - g2b2 * b2[0]) / n2, (g2[1] - g2b2 * b2[1]) / n2, (g2[2] - g2b2 * b2[2]) / n2};
  // u2 = a2 - (b1.a2) b1: dL/da2 = gu - (gu.b1) b1; dL/db1 += -(b1.a2) gu - (gu.b1) a2
  const float gub1 = dot3(gu, b1);
  float* go = grad + (int64_t)n * ldg;
  go[3] = gu[0] - gub1 * b1[0];
  go[4] = gu[1] - gub1 * b1[1];
  go[5] = gu[2] - gub1 * b1[2];
  g1[0] += -p * gu[0] - gub1 * a2[0];
  g1[1] += -p * gu[1] - gub1 * a2[1];
  g1[2] += -p * gu[2] - gub1 * a2[2];
  // b1 = a1 / |a1|: dL/da1 = (g1 - (g1.b1) b1) / n1
  const float g1b1 = dot3(g1, b1);
  go[0] = (g1[0] - g1b1 * b1[0]) / n1;
  go[1] = (g1[1] - g1b1 * b1[1]) / n1;
  go[2] = (g1[2] - g1b1 * b1[2]) / n1;
}

// 6D -> unit quaternion (w >= 0) of the same rotation the loss uses: rows
// b1, b2, b3 of R by Gram-Schmidt, then the matrix -> quaternion inverse of
// quat_matrix by Shepperd's method (largest of 1 + tr, 1 + 2 R_ii - tr
// chosen as the pivot, so the division is well conditioned).
__global__ __launch_bounds__(256) void rot6d_quat_kernel(const float* __restrict__ z6, int ldz, int B,
                                                         float* __restrict__ q) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= B) return;
  const float* a = z6 + (int64_t)n * ldz;
  float a1[3] = {a[0], a[1], a[2]}, a2[3] = {a[3], a[4], a[5]};
  const float n1 = sqrtf(dot3(a1, a1)) + 1e-12f;
  float b1[3] = {a1[0] / n1, a1[1] / n1, a1[2] / n1};
  const float p = dot3(b1, a2);
  float u2[3] = {a2[0] - p * b1[0], a2[1] - p * b1[1], a2[2] - p * b1[2]};
  const float n2 = sqrtf(dot3(u2, u2)) + 1e-12f;
  float b2[3] = {u2[0] / n2, u2[1] / n2, u2[2] / n2};
  float b3[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
  const float R00 = b1[0], R01 = b1[1], R02 = b1[2];
  const float R10 = b2[0], R11 = b2[1], R12 = b2[2];
  const float R20 = b3[0], R21 = b3[1], R22 = b3[2];
  const float tr = R00 + R11 + R22;
  float w, x, y, z;
  if (tr >= R00 && tr >= R11 && tr >= R22) {
    const float s = 2.0f * sqrtf(fmaxf(1.0f + tr, 0.0f));
    w = 0.25f * s;
    x = (R21 - R12) / s;
    y = (R02 - R20) / s;
    z = (R10 - R01) / s;
  } else if (R00 >= R11 && R00 >= R22) {
    const float s = 2.0f * sqrtf(fmaxf(1.0f + 2.0f * R00 - tr, 0.0f));
    w = (R21 - R12) / s;
    x = 0.25f * s;
    y = (R01 + R10) / s;
    z = (R02 + R20) / s;
  } else if (R11 >= R22) {
    const float s = 2.0f * sqrtf(fmaxf(1.0f + 2.0f * R11 - tr, 0.0f));
    w = (R02 - R20) / s;
    x = (R01 + R10) / s;
    y = 0.25f * s;
    z = (R12 + R21) / s;
  } else {
    const float s = 2.0f * sqrtf(fmaxf(1.0f + 2.0f * R22 - tr, 0.0f));
    w = (R10 - R01) / s;
    x = (R02 + R20) / s;
    y = (R12 + R21) / s;
    z = 0.25f * s;
  }
  const float inv = (w < 0.0f ? -1.0f : 1.0f) * rsqrtf(w * w + x * x + y * y + z * z);
  float* o = q + 4 * (int64_t)n;
  o[0] = w * inv;
  o[1] = x * inv;
  o[2] = y * inv;
  o[3] = z * inv;
}

// Adam with bias correction on FP32 master weights; optionally refreshes a
// BF16 copy used as the GEMM operand.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int64_t n,
                                                   float lr, float b1, float b2, float eps, float c1, float c2,
                                                   float wd, __hip_bfloat16* __restrict__ w_bf16) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w_bf16) w_bf16[i] = __float2bfloat16(wi);
  }
}

// 1D launches are grid-stride loops over at most 2^20 workgroups, so any
// element count fits the launch limits.
inline dim3 grid1d(int64_t n) {
  const int64_t blocks = (n + 255) / 256;
  return dim3((unsigned)(blocks < (1 << 20) ? (blocks > 0 ? blocks : 1) : (1 << 20)));
}

}  // namespace

extern "C" int ebsd_nn_im2col(const float* x, int b, int h, int w, int c, float scale, void* col, int ld,
                              int dtype, int transposed, void* stream) {
  if (!x || !col) return EBSD_ERR_NULL;
  if (b < 1 || h < 1 || w < 1 || c < 1) return EBSD_ERR_SHAPE;
  const int ho = (h + 1) / 2, wo = (w + 1) / 2;
  const int64_t n_rows = (int64_t)b * ho * wo;
  if (transposed ? (ld < n_rows) : (ld < 9 * c)) return EBSD_ERR_SHAPE;
  const int64_t total = transposed ? (int64_t)9 * c * ld : n_rows * ld;
  hipStream_t st = static_cast<hipStream_t>(stream);
#define EBSD_IM2COL(T, TR)                                                                                   \
  hipLaunchKernelGGL((im2col_kernel<T, TR>), grid1d(total), dim3(256), 0, st, x, b, h, w, c, ho, wo, scale,  \
                     static_cast<T*>(col), ld)
  if (dtype == EBSD_DTYPE_F32) {
    if (transposed) EBSD_IM2COL(float, true); else EBSD_IM2COL(float, false);
  } else if (dtype == EBSD_DTYPE_BF16) {
    if (transposed) EBSD_IM2COL(__hip_bfloat16, true); else EBSD_IM2COL(__hip_bfloat16, false);
  } else {
    return EBSD_ERR_DTYPE;
  }
#undef EBSD_IM2COL
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_cast(const float* src, int rows, int cols, int ld_src, void* dst, int ld_dst, int dtype,
                            void* stream) {
  if (!src || !dst) return EBSD_ERR_NULL;
  if (rows < 0 || cols < 1 || ld_src < cols || ld_dst < cols) return EBSD_ERR_SHAPE;
  if (rows == 0) return EBSD_OK;
  const int64_t total = (int64_t)rows * ld_dst;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(cast_kernel<float>, grid1d(total), dim3(256), 0, st, src, rows, cols, ld_src,
                       static_cast<float*>(dst), ld_dst);
  else if (dtype == EBSD_DTYPE_BF16)
    hipLaunchKernelGGL(cast_kernel<__hip_bfloat16>, grid1d(total), dim3(256), 0, st, src, rows, cols, ld_src,
                       static_cast<__hip_bfloat16*>(dst), ld_dst);
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_col2im(const float* dcol, int ld, int b, int h, int w, int c, float* dx, void* stream) {
  if (!dcol || !dx) return EBSD_ERR_NULL;
  if (b < 1 || h < 1 || w < 1 || c < 1 || ld < 9 * c) return EBSD_ERR_SHAPE;
  const int ho = (h + 1) / 2, wo = (w + 1) / 2;
  const int64_t total = (int64_t)b * h * w * c;
  hipLaunchKernelGGL(col2im_kernel, grid1d(total), dim3(256), 0, static_cast<hipStream_t>(stream), dcol, ld, b,
                     h, w, c, ho, wo, dx);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_bias_act(float* y, int rows, int cols, int ld, const float* bias, int relu, void* stream) {
  if (!y || !bias) return EBSD_ERR_NULL;
  if (rows < 0 || cols < 1 || ld < cols) return EBSD_ERR_SHAPE;
  if (rows == 0) return EBSD_OK;
  hipLaunchKernelGGL(bias_act_kernel, grid1d((int64_t)rows * cols), dim3(256), 0,
                     static_cast<hipStream_t>(stream), y, rows, cols, ld, bias, relu);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_relu_bwd(float* dy, const float* y, long long n, void* stream) {
  if (!dy || !y) return EBSD_ERR_NULL;
  if (n < 0) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  hipLaunchKernelGGL(relu_bwd_kernel, grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), dy, y,
                     (int64_t)n);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

// Rows per partial sum: 256, or more when that would need more than 65,535
// partials (the launch's y dimension).
static int colsum_chunk(int rows) {
  int chunk = 256;
  while ((rows + chunk - 1) / chunk > 65535) chunk *= 2;
  return chunk;
}

extern "C" int ebsd_nn_colsum_workspace(int rows, int cols, size_t* bytes) {
  if (!bytes) return EBSD_ERR_NULL;
  if (rows < 1 || cols < 1) return EBSD_ERR_SHAPE;
  const int chunk = colsum_chunk(rows);
  *bytes = (size_t)((rows + chunk - 1) / chunk) * cols * sizeof(float);
  return EBSD_OK;
}

extern "C" int ebsd_nn_colsum(const float* x, int rows, int cols, int ld, float* out, void* ws, size_t ws_bytes,
                              void* stream) {
  if (!x || !out) return EBSD_ERR_NULL;
  if (rows < 1 || cols < 1 || ld < cols) return EBSD_ERR_SHAPE;
  const int chunk = colsum_chunk(rows);
  const int parts = (rows + chunk - 1) / chunk;
  if (!ws || ws_bytes < (size_t)parts * cols * sizeof(float)) return EBSD_ERR_WORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  float* part = static_cast<float*>(ws);
  hipLaunchKernelGGL(colsum_partial_kernel, dim3((cols + 255) / 256, parts), dim3(256), 0, st, x, rows, cols, ld,
                     chunk, part);
  EBSD_LAUNCH_CHECK();
  hipLaunchKernelGGL(colsum_final_kernel, dim3((cols + 255) / 256), dim3(256), 0, st, part, parts, cols, out);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_transpose(const float* src, int rows, int cols, int ld_src, void* dst, int ld_dst,
                                 int dtype, void* stream) {
  if (!src || !dst) return EBSD_ERR_NULL;
  if (rows < 1 || cols < 1 || ld_src < cols || ld_dst < rows) return EBSD_ERR_SHAPE;
  const dim3 grid((ld_dst + 31) / 32, (cols + 31) / 32);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(transpose_kernel<float>, grid, dim3(256), 0, st, src, rows, cols, ld_src,
                       static_cast<float*>(dst), ld_dst);
  else if (dtype == EBSD_DTYPE_BF16)
    hipLaunchKernelGGL(transpose_kernel<__hip_bfloat16>, grid, dim3(256), 0, st, src, rows, cols, ld_src,
                       static_cast<__hip_bfloat16*>(dst), ld_dst);
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_geodesic_loss(const float* z6, int ldz, int b, const float* q_true, const float* ops,
                                     int n_ops, float* loss, float* grad, int ldg, void* stream) {
  if (!z6 || !q_true || !ops || !loss || !grad) return EBSD_ERR_NULL;
  if (b < 0 || b > (1 << 28) || ldz < 6 || ldg < 6 || n_ops < 1) return EBSD_ERR_SHAPE;
  if (b == 0) return EBSD_OK;
  hipLaunchKernelGGL(geodesic_loss_kernel, grid1d(b), dim3(256), 0, static_cast<hipStream_t>(stream), z6, ldz, b,
                     q_true, ops, n_ops, loss, grad, ldg);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float beta1,
                            float beta2, float eps, int step, float weight_decay, void* w_bf16, void* stream) {
  if (!w || !g || !m || !v) return EBSD_ERR_NULL;
  if (n < 0 || step < 1) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, grid1d(n), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v,
                     (int64_t)n, lr, beta1, beta2, eps, c1, c2, weight_decay,
                     static_cast<__hip_bfloat16*>(w_bf16));
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_rot6d_to_quat(const float* z6, int ldz, int b, float* q, void* stream) {
  if (!z6 || !q) return EBSD_ERR_NULL;
  if (b < 0 || b > (1 << 28) || ldz < 6) return EBSD_ERR_SHAPE;
  if (b == 0) return EBSD_OK;
  hipLaunchKernelGGL(rot6d_quat_kernel, grid1d(b), dim3(256), 0, static_cast<hipStream_t>(stream), z6, ldz, b, q);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}
