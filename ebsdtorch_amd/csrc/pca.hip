// PCA kernels for gfx950: the streaming covariance update of
// utils/pca_covmat.py:53-91 and the projection GEMMs of indexing.py:401,
// :421, :452.
//
// Covariance: for a batch X (n x d) the reference adds
//   (X - mu_old)^T (X - mu_new)                       (pca_covmat.py:81-87)
// to the aggregate. With the batch mean mu_b, Y = X - mu_b and
// a = mu_b - mu_old this equals Chan's batch identity
//   Y^T Y + (n_old n / (n_old + n)) a a^T,
// which is what is computed: column means (two-pass, fixed order, so the
// result is deterministic), a centred transposed copy Y^T in the workspace,
// then one MFMA GEMM Y^T (Y^T)^T that adds the rank-1 term in its epilogue.
// Only block tiles on or above the diagonal are computed; the lower triangle
// is filled from the upper one when the matrix is read out.
//
// GEMM: C = alpha A B^T + beta C (+ w u v^T) with A (m x K), B (n x K)
// row-major, K contiguous and a multiple of 32 floats; the 128 x 128 block
// product is the MFMA tile of mfma_tile.h.

#include "mfma_tile.h"

namespace {

using namespace ebsd_tile;

constexpr int kColChunk = 256;  // rows per partial column sum

// partial[c][j] = sum of X[r][j] over rows r in chunk c.
__global__ __launch_bounds__(256) void col_partial_kernel(const float* __restrict__ X, int n, int d, int ld,
                                                          float* __restrict__ partial) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (j >= d) return;
  const int r0 = c * kColChunk;
  const int r1 = min(r0 + kColChunk, n);
  float s = 0.0f;
  for (int r = r0; r < r1; ++r) s += X[(size_t)r * ld + j];
  partial[(size_t)c * d + j] = s;
}

// mu_b = sum of partials / n; a = mu_b - mean; mean <- mean + (n / (n_old + n)) a
// (the running-mean update of pca_covmat.py:78).
__global__ __launch_bounds__(256) void mean_update_kernel(const float* __restrict__ partial, int n_chunks, int d,
                                                          int n, double n_old, float* __restrict__ mean,
                                                          float* __restrict__ a_out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= d) return;
  double s = 0.0;
  for (int c = 0; c < n_chunks; ++c) s += (double)partial[(size_t)c * d + j];
  const double mu_b = s / (double)n;
  const double mu_o = (double)mean[j];
  const double a = mu_b - mu_o;
  a_out[j] = (float)a;
  mean[j] = (float)(mu_o + a * ((double)n / (n_old + (double)n)));
  a_out[d + j] = (float)mu_b;
}

// Yt[j][r] = X[r][j] - mu_b[j] for r < n, 0 for n <= r < n_pad (64 x 64 tiles
// through LDS so both the read and the write are coalesced).
__global__ __launch_bounds__(256) void center_transpose_kernel(const float* __restrict__ X, int n, int d, int ld,
                                                               const float* __restrict__ mu_b, int n_pad,
                                                               float* __restrict__ Yt) {
  __shared__ float tile[64][65];
  const int r0 = blockIdx.x * 64;
  const int j0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  for (int rr = ty; rr < 64; rr += 4) {
    const int r = r0 + rr, j = j0 + tx;
    float v = 0.0f;
    if (r < n && j < d) v = X[(size_t)r * ld + j] - mu_b[j];
    tile[rr][tx] = v;
  }
  __syncthreads();
  for (int jj = ty; jj < 64; jj += 4) {
    const int j = j0 + jj, r = r0 + tx;
    if (j < d && r < n_pad) Yt[(size_t)j * n_pad + r] = tile[tx][jj];
  }
}

// C[i][j] = act(alpha * (A B^T)[i][j] + beta * C[i][j] + w * u[i] * v[j] + bias[j])
// for the 128 x 128 block (blockIdx.x, blockIdx.y), act = ReLU if relu (bias
// NULL: no bias term). upper_only skips blocks strictly below the block
// diagonal.
// Split K: blockIdx.z = s covers stages [s * split_stages, (s + 1) *
// split_stages) of the K loop and writes its own output slice
// C + s * split_stride (alpha, beta, u, v apply per slice; the split-K path
// writes partial sums with beta = 0 and reduces them afterwards).
template <int DT>
__global__ __launch_bounds__(256, 2) void gemm_nt_kernel(const unsigned char* __restrict__ A, int M,
                                                         const unsigned char* __restrict__ B, int N, int ld_bytes,
                                                         float* __restrict__ C, int ldc, float alpha, float beta,
                                                         const float* __restrict__ u, const float* __restrict__ v,
                                                         float w, int upper_only, int split_stages,
                                                         int64_t split_stride, const float* __restrict__ bias,
                                                         int relu) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_BYTES];
  const int bi_tile = blockIdx.x, bj_tile = blockIdx.y;
  if (upper_only && bi_tile > bj_tile) return;  // uniform over the workgroup
  const int total_stages = ld_bytes / STAGE_BYTES;
  const int s0 = blockIdx.z * split_stages;
  const int ns = min(split_stages, total_stages - s0);
  if (ns <= 0) return;
  C += blockIdx.z * split_stride;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int i0 = bi_tile * TILE, j0 = bj_tile * TILE;
  const int lr = tid >> 3;
  const int lc = (tid & 7) * 16;
  const unsigned char* arow[4];
  const unsigned char* brow[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    arow[i] = A + (size_t)min(i0 + lr + 32 * i, M - 1) * ld_bytes + lc + (size_t)s0 * STAGE_BYTES;
    brow[i] = B + (size_t)min(j0 + lr + 32 * i, N - 1) * ld_bytes + lc + (size_t)s0 * STAGE_BYTES;
  }
  const int a_off = (lane & 31) * LDS_ROW + lane_k_off<DT>(lane);
  const int b_off = (wave * 32 + (lane & 31)) * LDS_ROW + lane_k_off<DT>(lane);
  const int st_off = lr * LDS_ROW + lc;
  floatx16 acc[4];
  tile_product<DT>(arow, brow, ns, smem, a_off, b_off, st_off, acc);
  const int col = j0 + wave * 32 + (lane & 31);
  if (col >= N) return;
  const float vj = (v && w != 0.0f) ? w * v[col] : 0.0f;
  const float bj = bias ? bias[col] : 0.0f;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = i0 + bi * 32 + mfma32_row(r, lane);
      if (row < M) {
        float* cp = C + (size_t)row * ldc + col;
        float out = alpha * acc[bi][r];
        if (beta != 0.0f) out += beta * *cp;
        if (vj != 0.0f) out += u[row] * vj;
        if (bias) out += bj;
        if (relu) out = fmaxf(out, 0.0f);
        *cp = out;
      }
    }
  }
}

int elem_bytes(int dtype) {
  switch (dtype) {
    case EBSD_DTYPE_F32: return 4;
    case EBSD_DTYPE_BF16:
    case EBSD_DTYPE_F16: return 2;
  }
  return 0;
}

// out[i][j] = alpha * sum_s part[s][i][j] + beta * out[i][j]
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const float* __restrict__ part, int splits, int m,
                                                            int n, float* __restrict__ out, int ldc, float alpha,
                                                            float beta) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += part[(int64_t)k * m * n + idx];
  float* o = out + (int64_t)i * ldc + j;
  *o = beta != 0.0f ? alpha * s + beta * *o : alpha * s;
}

int launch_gemm(const void* a, int m, const void* b, int n, int ld, int dtype, float* c, int ldc, float alpha,
                float beta, const float* u, const float* v, float w, int upper_only, hipStream_t st,
                int splits = 1, int split_stages = 0, int64_t split_stride = 0, const float* bias = nullptr,
                int relu = 0) {
  const int eb = elem_bytes(dtype);
  if (splits <= 1) split_stages = ld * eb / STAGE_BYTES;
  const dim3 grid((m + TILE - 1) / TILE, (n + TILE - 1) / TILE, splits > 1 ? splits : 1), block(256);
  const unsigned char* A = static_cast<const unsigned char*>(a);
  const unsigned char* B = static_cast<const unsigned char*>(b);
  switch (dtype) {
    case EBSD_DTYPE_F32:
      hipLaunchKernelGGL(gemm_nt_kernel<EBSD_DTYPE_F32>, grid, block, 0, st, A, m, B, n, ld * eb, c, ldc, alpha,
                         beta, u, v, w, upper_only, split_stages, split_stride, bias, relu);
      break;
    case EBSD_DTYPE_BF16:
      hipLaunchKernelGGL(gemm_nt_kernel<EBSD_DTYPE_BF16>, grid, block, 0, st, A, m, B, n, ld * eb, c, ldc, alpha,
                         beta, u, v, w, upper_only, split_stages, split_stride, bias, relu);
      break;
    case EBSD_DTYPE_F16:
      hipLaunchKernelGGL(gemm_nt_kernel<EBSD_DTYPE_F16>, grid, block, 0, st, A, m, B, n, ld * eb, c, ldc, alpha,
                         beta, u, v, w, upper_only, split_stages, split_stride, bias, relu);
      break;
    default:
      return EBSD_ERR_DTYPE;
  }
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

struct CovWs {
  size_t partial, vecs, yt, total;
  int n_chunks, n_pad;
};

CovWs cov_ws(int n, int d) {
  CovWs w;
  w.n_chunks = (n + kColChunk - 1) / kColChunk;
  w.n_pad = (int)round_up((size_t)n, 32);
  w.partial = 0;
  w.vecs = round_up((size_t)w.n_chunks * d * sizeof(float), 256);
  w.yt = w.vecs + round_up((size_t)2 * d * sizeof(float), 256);
  w.total = w.yt + (size_t)d * w.n_pad * sizeof(float);
  return w;
}

}  // namespace

extern "C" int ebsd_gemm_nt(const void* a, int m, const void* b, int n, int ld, int dtype, float* c, int ldc,
                            float alpha, float beta, void* stream) {
  if (!a || !b || !c) return EBSD_ERR_NULL;
  if (m < 0 || n < 0 || ld < 1 || ldc < n) return EBSD_ERR_SHAPE;
  const int eb = elem_bytes(dtype);
  if (eb == 0) return EBSD_ERR_DTYPE;
  if (ld % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) != 0) return EBSD_ERR_ALIGN;
  if (m == 0 || n == 0) return EBSD_OK;
  return launch_gemm(a, m, b, n, ld, dtype, c, ldc, alpha, beta, nullptr, nullptr, 0.0f, 0,
                     static_cast<hipStream_t>(stream));
}

extern "C" int ebsd_gemm_nt_bias_act(const void* a, int m, const void* b, int n, int ld, int dtype, float* c,
                                     int ldc, const float* bias, int relu, void* stream) {
  if (!a || !b || !c || !bias) return EBSD_ERR_NULL;
  if (m < 0 || n < 0 || ld < 1 || ldc < n) return EBSD_ERR_SHAPE;
  const int eb = elem_bytes(dtype);
  if (eb == 0) return EBSD_ERR_DTYPE;
  if (ld % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) != 0) return EBSD_ERR_ALIGN;
  if (m == 0 || n == 0) return EBSD_OK;
  return launch_gemm(a, m, b, n, ld, dtype, c, ldc, 1.0f, 0.0f, nullptr, nullptr, 0.0f, 0,
                     static_cast<hipStream_t>(stream), 1, 0, 0, bias, relu ? 1 : 0);
}

extern "C" int ebsd_gemm_nt_splitk_workspace(int m, int n, int splits, size_t* bytes) {
  if (!bytes) return EBSD_ERR_NULL;
  if (m < 1 || n < 1 || splits < 1) return EBSD_ERR_SHAPE;
  *bytes = (size_t)splits * m * n * sizeof(float);
  return EBSD_OK;
}

extern "C" int ebsd_gemm_nt_splitk(const void* a, int m, const void* b, int n, int ld, int dtype, float* c,
                                   int ldc, float alpha, float beta, int splits, void* workspace,
                                   size_t ws_bytes, void* stream) {
  if (!a || !b || !c) return EBSD_ERR_NULL;
  if (m < 1 || n < 1 || ld < 1 || ldc < n || splits < 1) return EBSD_ERR_SHAPE;
  const int eb = elem_bytes(dtype);
  if (eb == 0) return EBSD_ERR_DTYPE;
  if (ld % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) != 0) return EBSD_ERR_ALIGN;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int stages = ld * eb / STAGE_BYTES;
  const int per = (stages + splits - 1) / splits;
  const int used = (stages + per - 1) / per;
  if (used <= 1)
    return launch_gemm(a, m, b, n, ld, dtype, c, ldc, alpha, beta, nullptr, nullptr, 0.0f, 0, st);
  if (!workspace || ws_bytes < (size_t)used * m * n * sizeof(float)) return EBSD_ERR_WORKSPACE;
  float* part = static_cast<float*>(workspace);
  const int r = launch_gemm(a, m, b, n, ld, dtype, part, n, 1.0f, 0.0f, nullptr, nullptr, 0.0f, 0, st, used, per,
                            (int64_t)m * n);
  if (r) return r;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)(((int64_t)m * n + 255) / 256)), dim3(256), 0, st, part,
                     used, m, n, c, ldc, alpha, beta);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_cov_update_workspace(int n, int d, size_t* bytes) {
  if (!bytes) return EBSD_ERR_NULL;
  if (n < 1 || d < 1) return EBSD_ERR_SHAPE;
  *bytes = cov_ws(n, d).total;
  return EBSD_OK;
}

extern "C" int ebsd_cov_update(const float* x, int n, int d, int ld_x, float* mean, float* cov, int ld_cov,
                               long long n_seen, void* workspace, size_t ws_bytes, void* stream) {
  if (!x || !mean || !cov) return EBSD_ERR_NULL;
  if (n < 1 || d < 1 || ld_x < d || ld_cov < d || n_seen < 0) return EBSD_ERR_SHAPE;
  const CovWs w = cov_ws(n, d);
  if (!workspace || ws_bytes < w.total) return EBSD_ERR_WORKSPACE;
  if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0) return EBSD_ERR_ALIGN;
  hipStream_t st = static_cast<hipStream_t>(stream);
  unsigned char* base = static_cast<unsigned char*>(workspace);
  float* partial = reinterpret_cast<float*>(base + w.partial);
  float* a_mu = reinterpret_cast<float*>(base + w.vecs);  // a (d) then mu_b (d)
  float* yt = reinterpret_cast<float*>(base + w.yt);
  hipLaunchKernelGGL(col_partial_kernel, dim3((d + 255) / 256, w.n_chunks), dim3(256), 0, st, x, n, d, ld_x,
                     partial);
  EBSD_LAUNCH_CHECK();
  hipLaunchKernelGGL(mean_update_kernel, dim3((d + 255) / 256), dim3(256), 0, st, partial, w.n_chunks, d, n,
                     (double)n_seen, mean, a_mu);
  EBSD_LAUNCH_CHECK();
  hipLaunchKernelGGL(center_transpose_kernel, dim3(w.n_pad / 64 + (w.n_pad % 64 ? 1 : 0), (d + 63) / 64),
                     dim3(256), 0, st, x, n, d, ld_x, a_mu + d, w.n_pad, yt);
  EBSD_LAUNCH_CHECK();
  const float wgt = (float)((double)n_seen * (double)n / ((double)n_seen + (double)n));
  return launch_gemm(yt, d, yt, d, w.n_pad, EBSD_DTYPE_F32, cov, ld_cov, 1.0f, 1.0f, a_mu, a_mu, wgt, 1, st);
}
