// Kernels of the learned orientation-proposal network (gfx950).
//
// The network is 4 convolutions (3x3, stride 2, padding 1, ReLU) and a
// 2-layer MLP producing a 6D rotation. Activations are NHWC float32;
// convolutions run as im2col + MFMA GEMM with a bias + ReLU epilogue
// (ebsd_gemm_nt_bias_act), except the first (one input channel, 9 taps per
// output), which is computed directly. This file holds the data-movement and
// element-wise pieces around the GEMMs: the direct first convolution, im2col
// / col2im, bias + ReLU forward and backward, column sums (bias gradients),
// transpose-and-cast (GEMM operand layouts), the symmetric geodesic loss and
// the Adam update.

#include <algorithm>

#include "common.h"
#include "mfma_tile.h"

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

// Row-major im2col (the forward GEMM's A operand) with one thread per 8
// consecutive columns of one row and 32-bit index arithmetic (the entry
// point checks rows * ld / 8 < 2^31): with C a multiple of 8 the 8 columns
// are 8 consecutive channels of one tap, two 16-byte loads; otherwise (the
// first convolution, C = 1) they are gathered one by one. Same values as
// im2col_kernel<T, false>.
template <typename T>
__device__ __forceinline__ void store8(T* dst, const float (&v)[8]);
template <>
__device__ __forceinline__ void store8<float>(float* dst, const float (&v)[8]) {
  reinterpret_cast<float4*>(dst)[0] = make_float4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<float4*>(dst)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
template <>
__device__ __forceinline__ void store8<__hip_bfloat16>(__hip_bfloat16* dst, const float (&v)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
    w[e] = (uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * e])) |
           ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * e + 1])) << 16);
  *reinterpret_cast<uint4*>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
}

template <typename T>
__global__ __launch_bounds__(256) void im2col8_kernel(const float* __restrict__ X, int H, int W, int C, int Ho,
                                                      int Wo, float scale, T* __restrict__ col, int ld,
                                                      int n_units) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n_units) return;
  const int ld8 = ld >> 3;
  const int row = t / ld8;
  const int k0 = (t - row * ld8) << 3;
  const int ox = row % Wo;
  const int r2 = row / Wo;
  const int oy = r2 % Ho;
  const int b = r2 / Ho;
  float v[8];
  if ((C & 7) == 0) {
    const int kk = k0 / C;
    const int c = k0 - kk * C;
    const int ky = kk / 3, kx = kk - 3 * (kk / 3);
    const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
    if (kk < 9 && y >= 0 && y < H && x >= 0 && x < W) {
      const float4* p = reinterpret_cast<const float4*>(X + (((int64_t)b * H + y) * W + x) * C + c);
      const float4 a = p[0], d = p[1];
      v[0] = scale * a.x; v[1] = scale * a.y; v[2] = scale * a.z; v[3] = scale * a.w;
      v[4] = scale * d.x; v[5] = scale * d.y; v[6] = scale * d.z; v[7] = scale * d.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = 0.0f;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + e;
      const int kk = k / C;
      const int c = k - kk * C;
      const int ky = kk / 3, kx = kk - 3 * (kk / 3);
      const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
      v[e] = (kk < 9 && y >= 0 && y < H && x >= 0 && x < W)
                 ? scale * X[(((int64_t)b * H + y) * W + x) * C + c]
                 : 0.0f;
    }
  }
  store8<T>(col + (int64_t)t * 8, v);
}

// Direct convolution of a one-channel image (the first layer) with fused
// bias + ReLU: Y[row][co] = max(0, bias[co] + sum_t w[co][t] in_t) over the
// 9 taps t = ky*3 + kx, in_t = round_T(scale * X[2 oy + ky - 1][2 ox + kx - 1])
// (0 outside the image): the values of im2col<T> + the GEMM + bias_act, up to
// the order of the fp32 sum. One thread per output pixel (row): its taps and
// index arithmetic once, then all c_out channels with the weights and biases
// as wave-uniform (scalar) loads; grid-stride launch of 8 workgroups per CU.
template <typename T>
__device__ __forceinline__ float round_to(float v);
template <>
__device__ __forceinline__ float round_to<float>(float v) { return v; }
template <>
__device__ __forceinline__ float round_to<__hip_bfloat16>(float v) { return __bfloat162float(__float2bfloat16(v)); }

template <typename T>
__device__ __forceinline__ float load_float(const T* p);
template <>
__device__ __forceinline__ float load_float<float>(const float* p) { return *p; }
template <>
__device__ __forceinline__ float load_float<__hip_bfloat16>(const __hip_bfloat16* p) { return __bfloat162float(*p); }

constexpr int kConvC1MaxOut = 256;

// The 9 weights of a channel row as floats. BF16 rows are read as 5 dwords
// (two weights each; the row is 4-byte aligned and its 10th entry is padding)
// so that wave-uniform rows become scalar loads.
template <typename T>
__device__ __forceinline__ void load_taps(const T* __restrict__ w, float (&out)[9]);
template <>
__device__ __forceinline__ void load_taps<float>(const float* __restrict__ w, float (&out)[9]) {
#pragma unroll
  for (int t = 0; t < 9; ++t) out[t] = w[t];
}
template <>
__device__ __forceinline__ void load_taps<__hip_bfloat16>(const __hip_bfloat16* __restrict__ w, float (&out)[9]) {
  const uint32_t* d = reinterpret_cast<const uint32_t*>(w);
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    const uint32_t v = d[t];
    out[2 * t] = __uint_as_float(v << 16);
    if (2 * t + 1 < 9) out[2 * t + 1] = __uint_as_float(v & 0xffff0000u);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void conv_c1_kernel(const float* __restrict__ X, int H, int W, int Ho, int Wo,
                                                      float scale, const T* __restrict__ wgt, int ld_w,
                                                      const float* __restrict__ bias, int c_out,
                                                      float* __restrict__ Y, int n_rows) {
  for (int row = blockIdx.x * 256 + threadIdx.x; row < n_rows; row += gridDim.x * 256) {
    const int ox = row % Wo;
    const int r2 = row / Wo;
    const int oy = r2 % Ho;
    const int b = r2 / Ho;
    const float* __restrict__ img = X + (int64_t)b * H * W;
    float in[9];
#pragma unroll
    for (int ky = 0; ky < 3; ++ky)
#pragma unroll
      for (int kx = 0; kx < 3; ++kx) {
        const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
        in[ky * 3 + kx] = (y >= 0 && y < H && x >= 0 && x < W) ? round_to<T>(scale * img[y * W + x]) : 0.0f;
      }
    float* __restrict__ out = Y + (int64_t)row * c_out;
    // co is wave-uniform: the weights and biases are scalar loads
    for (int co = 0; co < c_out; co += 4) {
      float r[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float wv[9];
        load_taps<T>(wgt + (size_t)(co + q) * ld_w, wv);
        float s = wv[0] * in[0];
#pragma unroll
        for (int t = 1; t < 9; ++t) s = fmaf(wv[t], in[t], s);
        r[q] = fmaxf(s + bias[co + q], 0.0f);
      }
      *reinterpret_cast<float4*>(out + co) = make_float4(r[0], r[1], r[2], r[3]);
    }
  }
}

// Implicit-GEMM convolution (3x3, stride 2, padding 1) with bias + ReLU:
// Y[row][co] = max(0, bias[co] + sum_k A[row][k] Wt[co][k]) where A is the
// im2col matrix of the NHWC float32 activations X, which is never stored:
// each K stage of the 128 x 128 MFMA tile (the GEMM's LDS layout and K order)
// is filled straight from X, converted to the operand type as im2col does,
// so Y equals im2col + the GEMM + bias/ReLU bit for bit. C must be a
// multiple of the 16-byte run of a stage (8 BF16 / 4 FP32 channels), so a
// thread's run is consecutive channels of one tap; K columns past 9 C read
// as zeros (the weights are zero there too).
template <int DT>
__device__ __forceinline__ ebsd_tile::floatx4 conv_a_run(const float* __restrict__ X, bool valid, int64_t off) {
  using ebsd_tile::floatx4;
  if constexpr (DT == EBSD_DTYPE_F32) {
    if (!valid) return floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    return *reinterpret_cast<const floatx4*>(X + off);
  } else {
    float v[8];
    if (valid) {
      const float4 a = reinterpret_cast<const float4*>(X + off)[0];
      const float4 b = reinterpret_cast<const float4*>(X + off)[1];
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = 0.0f;
    }
    uint32_t w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      w[e] = (uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * e])) |
             ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * e + 1])) << 16);
    return __builtin_bit_cast(floatx4, make_uint4(w[0], w[1], w[2], w[3]));
  }
}

template <int DT>
__global__ __launch_bounds__(256, 2) void conv_gemm_kernel(const float* __restrict__ X, int H, int W, int C, int Ho,
                                                           int Wo, int M, const unsigned char* __restrict__ Wt,
                                                           int N, int ld_bytes, const float* __restrict__ bias,
                                                           float* __restrict__ Y) {
  using namespace ebsd_tile;
  __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_BYTES];
  constexpr int EB = DT == EBSD_DTYPE_F32 ? 4 : 2;
  constexpr int STAGE_ELEMS = STAGE_BYTES / EB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int i0 = blockIdx.x * TILE, j0 = blockIdx.y * TILE;
  const int lr = tid >> 3;
  const int lc = (tid & 7) * 16;
  const int ke = lc / EB;  // first K element of this thread's run within a stage
  // the thread's 4 A rows: image base row b*H, top-left input corner
  int rbh[4], ry[4], rx[4];
  bool rv[4];
  const unsigned char* brow[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = i0 + lr + 32 * i;
    rv[i] = row < M;
    const int r = min(row, M - 1);
    const int ox = r % Wo;
    const int t = r / Wo;
    const int oy = t % Ho;
    rbh[i] = (t / Ho) * H;
    ry[i] = 2 * oy - 1;
    rx[i] = 2 * ox - 1;
    brow[i] = Wt + (size_t)min(j0 + lr + 32 * i, N - 1) * ld_bytes + lc;
  }
  const int nk = ld_bytes / STAGE_BYTES;
  floatx4 ra[4], rb[4];
  auto load = [&](int kt) {
    const int k = kt * STAGE_ELEMS + ke;
    const int tap = k / C;
    const int c = k - tap * C;
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int y = ry[i] + ky, x = rx[i] + kx;
      const bool ok = rv[i] && tap < 9 && y >= 0 && y < H && x >= 0 && x < W;
      ra[i] = conv_a_run<DT>(X, ok, ((int64_t)(rbh[i] + y) * W + x) * C + c);
      rb[i] = *reinterpret_cast<const floatx4*>(brow[i] + (size_t)kt * STAGE_BYTES);
    }
  };
  const int a_off = (lane & 31) * LDS_ROW + lane_k_off<DT>(lane);
  const int b_off = (wave * 32 + (lane & 31)) * LDS_ROW + lane_k_off<DT>(lane);
  const int st_off = lr * LDS_ROW + lc;
  floatx16 acc[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[a][r] = 0.0f;
  load(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    *reinterpret_cast<floatx4*>(smem + st_off + 32 * i * LDS_ROW) = ra[i];
    *reinterpret_cast<floatx4*>(smem + LDS_TILE + st_off + 32 * i * LDS_ROW) = rb[i];
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) load(kt + 1);
    const unsigned char* As = smem + cur * 2 * LDS_TILE;
    stage_mfma<DT>(As, As + LDS_TILE, a_off, b_off, acc);
    if (more) {
      unsigned char* Ad = smem + (cur ^ 1) * 2 * LDS_TILE;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        *reinterpret_cast<floatx4*>(Ad + st_off + 32 * i * LDS_ROW) = ra[i];
        *reinterpret_cast<floatx4*>(Ad + LDS_TILE + st_off + 32 * i * LDS_ROW) = rb[i];
      }
    }
    __syncthreads();
  }
  const int col = j0 + wave * 32 + (lane & 31);
  if (col >= N) return;
  const float bj = bias[col];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = i0 + bi * 32 + mfma32_row(r, lane);
      if (row < M) Y[(size_t)row * N + col] = fmaxf(acc[bi][r] + bj, 0.0f);
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
  const bool vec = !transposed && (ld & 7) == 0 && total / 8 < INT_MAX &&
                   (reinterpret_cast<uintptr_t>(col) & 15) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  if (vec && (dtype == EBSD_DTYPE_F32 || dtype == EBSD_DTYPE_BF16)) {
    const int units = (int)(total / 8);
    const dim3 g((unsigned)((units + 255) / 256));
    if (dtype == EBSD_DTYPE_F32)
      hipLaunchKernelGGL(im2col8_kernel<float>, g, dim3(256), 0, st, x, h, w, c, ho, wo, scale,
                         static_cast<float*>(col), ld, units);
    else
      hipLaunchKernelGGL(im2col8_kernel<__hip_bfloat16>, g, dim3(256), 0, st, x, h, w, c, ho, wo, scale,
                         static_cast<__hip_bfloat16*>(col), ld, units);
    EBSD_LAUNCH_CHECK();
    return EBSD_OK;
  }
#define EBSD_IM2COL(T, TR)                                                                                 \
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

extern "C" int ebsd_nn_conv_c1(const float* x, int b, int h, int w, float scale, const void* weight, int ld_w,
                               int dtype, const float* bias, int c_out, float* y, void* stream) {
  if (!x || !weight || !bias || !y) return EBSD_ERR_NULL;
  if (b < 1 || h < 1 || w < 1 || ld_w < 9 || c_out < 4 || c_out > kConvC1MaxOut || (c_out & 3) != 0)
    return EBSD_ERR_SHAPE;
  if ((reinterpret_cast<uintptr_t>(y) & 15) != 0) return EBSD_ERR_ALIGN;
  if (dtype == EBSD_DTYPE_BF16 && (ld_w < 10 || (ld_w & 1) != 0 || (reinterpret_cast<uintptr_t>(weight) & 3) != 0))
    return EBSD_ERR_ALIGN;  // BF16 rows are read as dwords, the 10th entry included
  const int ho = (h + 1) / 2, wo = (w + 1) / 2;
  const int64_t units = (int64_t)b * ho * wo;
  if (units >= INT_MAX - (1 << 20)) return EBSD_ERR_SHAPE;
  const dim3 g((unsigned)std::min<int64_t>((units + 255) / 256, 2048));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(conv_c1_kernel<float>, g, dim3(256), 0, st, x, h, w, ho, wo, scale,
                       static_cast<const float*>(weight), ld_w, bias, c_out, y, (int)units);
  else if (dtype == EBSD_DTYPE_BF16)
    hipLaunchKernelGGL(conv_c1_kernel<__hip_bfloat16>, g, dim3(256), 0, st, x, h, w, ho, wo, scale,
                       static_cast<const __hip_bfloat16*>(weight), ld_w, bias, c_out, y, (int)units);
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nn_conv_gemm(const float* x, int b, int h, int w, int c, const void* weight, int ld_w, int c_out,
                                 int dtype, const float* bias, float* y, void* stream) {
  if (!x || !weight || !bias || !y) return EBSD_ERR_NULL;
  if (dtype != EBSD_DTYPE_F32 && dtype != EBSD_DTYPE_BF16) return EBSD_ERR_DTYPE;
  const int run = dtype == EBSD_DTYPE_F32 ? 4 : 8;
  if (b < 1 || h < 1 || w < 1 || c < 1 || c % run != 0 || c_out < 1) return EBSD_ERR_SHAPE;
  if (ld_w < 9 * c || ld_w % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  if (((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(weight)) & 15) != 0) return EBSD_ERR_ALIGN;
  const int ho = (h + 1) / 2, wo = (w + 1) / 2;
  const int64_t m = (int64_t)b * ho * wo;
  if (m >= INT_MAX || (int64_t)b * h >= INT_MAX) return EBSD_ERR_SHAPE;
  const dim3 grid((unsigned)((m + ebsd_tile::TILE - 1) / ebsd_tile::TILE),
                  (unsigned)((c_out + ebsd_tile::TILE - 1) / ebsd_tile::TILE));
  const int eb = dtype == EBSD_DTYPE_F32 ? 4 : 2;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned char* wt = static_cast<const unsigned char*>(weight);
  if (dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(conv_gemm_kernel<EBSD_DTYPE_F32>, grid, dim3(256), 0, st, x, h, w, c, ho, wo, (int)m, wt,
                       c_out, ld_w * eb, bias, y);
  else
    hipLaunchKernelGGL(conv_gemm_kernel<EBSD_DTYPE_BF16>, grid, dim3(256), 0, st, x, h, w, c, ho, wo, (int)m, wt,
                       c_out, ld_w * eb, bias, y);
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
