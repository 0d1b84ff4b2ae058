// Per-pattern preprocessing of experimental patterns (gfx950): CLAHE
// (preprocessing/clahe.py:18-130) and per-pattern normalisation
// (experiment_pats.py:256-281). One workgroup per pattern; the pattern stays
// in cache between its passes.
//
// CLAHE, for a pattern of H x W values in [0, 1] and a gh x gw tile grid:
// the pattern is reflect-padded to a multiple of the grid (tiles of
// ceil(H/gh) x ceil(W/gw) pixels, row-major), each tile's n_bins histogram of
// trunc(v (n_bins - 1)) is built with LDS atomics, clipped at `limit` with the
// excess redistributed evenly and the remainder one count each to the lowest
// bins, and turned into a clamped CDF (count / tile pixels). Each output pixel
// is the trilinear, border-clamped, align_corners=False lookup of the
// (gh, gw, n_bins) CDF volume at (row, column, value) coordinates
// (linspace(-1, 1) over the rows and columns, 2 v - 1 over the bins), which
// is what the reference's 3D grid_sample computes.

#include <algorithm>

#include "common.h"

namespace {

constexpr int kMaxBins = 64;
constexpr int kMaxTiles = 64;

__device__ __forceinline__ int reflect_index(int i, int n) { return i < n ? i : 2 * (n - 1) - i; }

// torch.linspace(-1, 1, n)[i]: the lower half counts up from -1, the upper
// half down from 1.
__device__ __forceinline__ float linspace_pm1(int i, int n) {
  if (n == 1) return -1.0f;
  const float step = 2.0f / (float)(n - 1);
  return i < n / 2 ? -1.0f + step * (float)i : 1.0f - step * (float)(n - 1 - i);
}

// grid_sample's source index (align_corners=False) clipped to the border.
__device__ __forceinline__ float src_index(float coord, int size) {
  const float ix = ((coord + 1.0f) * (float)size - 1.0f) * 0.5f;
  return fminf(fmaxf(ix, 0.0f), (float)(size - 1));
}

__global__ __launch_bounds__(256) void clahe_kernel(const float* __restrict__ X, int H, int W, int gh, int gw,
                                                    int n_bins, float limit, float* __restrict__ Y) {
  __shared__ int hist[kMaxTiles * kMaxBins];
  __shared__ float cdf[kMaxTiles * kMaxBins];
  const int b = blockIdx.x;
  const float* __restrict__ x = X + (size_t)b * H * W;
  float* __restrict__ y = Y + (size_t)b * H * W;
  const int n_tiles = gh * gw;
  const int th = (H + gh - 1) / gh, tw = (W + gw - 1) / gw;
  const int Hp = th * gh, Wp = tw * gw;
  for (int e = threadIdx.x; e < n_tiles * n_bins; e += 256) hist[e] = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < Hp * Wp; e += 256) {
    const int r = e / Wp, c = e - (e / Wp) * Wp;
    const float v = x[reflect_index(r, H) * W + reflect_index(c, W)];
    // (tiles * (n_bins - 1)).to(int64): truncation; values outside [0, 1]
    // are clamped to the end bins
    const int bin = min(max((int)(v * (float)(n_bins - 1)), 0), n_bins - 1);
    atomicAdd(&hist[((r / th) * gw + c / tw) * n_bins + bin], 1);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float vox = (float)(th * tw);
  for (int t = wave; t < n_tiles; t += 4) {
    const bool on = lane < n_bins;
    const int cnt = on ? hist[t * n_bins + lane] : 0;
    // pdf = count / sum, in pixel counts: (count / vox) * vox
    float h = on ? ((float)cnt / vox) * vox : 0.0f;
    if (limit > 0.0f) {
      h = fminf(h, limit);
      float s = h;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      const float clipped = vox - s;
      const float residual = clipped - floorf(clipped / (float)n_bins) * (float)n_bins;
      const float redist = (clipped - residual) / (float)n_bins;
      h += redist;
      h += ((float)lane < residual) ? 1.0f : 0.0f;
      if (!on) h = 0.0f;
    }
    // inclusive scan over the bins
    float c = h;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const float o = __shfl_up(c, off, 64);
      if (lane >= off) c += o;
    }
    if (on) cdf[t * n_bins + lane] = fminf(fmaxf(c / vox, 0.0f), 1.0f);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < H * W; e += 256) {
    const int r = e / W, col = e - (e / W) * W;
    const float v = x[e];
    const float fz = src_index(2.0f * v - 1.0f, n_bins);
    const float fx = src_index(linspace_pm1(col, W), gw);
    const float fy = src_index(linspace_pm1(r, H), gh);
    const int z0 = (int)floorf(fz), x0 = (int)floorf(fx), y0 = (int)floorf(fy);
    const int z1 = min(z0 + 1, n_bins - 1), x1 = min(x0 + 1, gw - 1), y1 = min(y0 + 1, gh - 1);
    const float wz = fz - (float)z0, wx = fx - (float)x0, wy = fy - (float)y0;
    auto at = [&](int ty, int tx, int bz) { return cdf[(ty * gw + tx) * n_bins + bz]; };
    const float c00 = at(y0, x0, z0) + wz * (at(y0, x0, z1) - at(y0, x0, z0));
    const float c01 = at(y0, x1, z0) + wz * (at(y0, x1, z1) - at(y0, x1, z0));
    const float c10 = at(y1, x0, z0) + wz * (at(y1, x0, z1) - at(y1, x0, z0));
    const float c11 = at(y1, x1, z0) + wz * (at(y1, x1, z1) - at(y1, x1, z0));
    const float c0 = c00 + wx * (c01 - c00);
    const float c1 = c10 + wx * (c11 - c10);
    y[e] = c0 + wy * (c1 - c0);
  }
}

// mode 0: (x - min) / (1e-4 + max - min); 1: x - mean; 2: (x - mean) / std
// (unbiased, torch.std's default).
__global__ __launch_bounds__(256) void normalize_kernel(float* __restrict__ X, int P, int mode) {
  __shared__ float red[3][4];
  float* __restrict__ x = X + (size_t)blockIdx.x * P;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float a = mode == 0 ? INFINITY : 0.0f, m = -INFINITY;
  for (int i = threadIdx.x; i < P; i += 256) {
    const float v = x[i];
    if (mode == 0) {
      a = fminf(a, v);
      m = fmaxf(m, v);
    } else {
      a += v;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float oa = __shfl_xor(a, off, 64), om = __shfl_xor(m, off, 64);
    a = mode == 0 ? fminf(a, oa) : a + oa;
    m = fmaxf(m, om);
  }
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = m;
  }
  __syncthreads();
  float A = red[0][0], M = red[1][0];
  for (int w = 1; w < 4; ++w) {
    A = mode == 0 ? fminf(A, red[0][w]) : A + red[0][w];
    M = fmaxf(M, red[1][w]);
  }
  if (mode == 0) {
    const float den = 1e-4f + M - A;
    for (int i = threadIdx.x; i < P; i += 256) x[i] = (x[i] - A) / den;
    return;
  }
  const float mean = A / (float)P;
  if (mode == 1) {
    for (int i = threadIdx.x; i < P; i += 256) x[i] = x[i] - mean;
    return;
  }
  float s = 0.0f;
  for (int i = threadIdx.x; i < P; i += 256) {
    const float d = x[i] - mean;
    s = fmaf(d, d, s);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __syncthreads();  // every wave has read red[0..1]
  if (lane == 0) red[2][wave] = s;
  __syncthreads();
  const float var = (red[2][0] + red[2][1] + red[2][2] + red[2][3]) / (float)(P - 1);
  const float inv = 1.0f / sqrtf(var);
  for (int i = threadIdx.x; i < P; i += 256) x[i] = (x[i] - mean) * inv;
}

}  // namespace

extern "C" int ebsd_clahe(const float* x, int n, int h, int w, int grid_h, int grid_w, int n_bins, float clip_limit,
                          float* out, void* stream) {
  if (!x || !out) return EBSD_ERR_NULL;
  if (n < 0 || h < 2 || w < 2 || grid_h < 1 || grid_w < 1 || grid_h * grid_w > kMaxTiles || n_bins < 2 ||
      n_bins > kMaxBins)
    return EBSD_ERR_SHAPE;
  const int th = (h + grid_h - 1) / grid_h, tw = (w + grid_w - 1) / grid_w;
  // reflect padding needs pad < size (torch's rule)
  if (th * grid_h - h >= h || tw * grid_w - w >= w) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  // limit = max(clip_limit * tile_pixels // n_bins, 1) (clahe.py:86-87)
  float limit = 0.0f;
  if (clip_limit > 0.0f) limit = (float)std::max(std::floor((double)clip_limit * (th * tw) / n_bins), 1.0);
  hipLaunchKernelGGL(clahe_kernel, dim3((unsigned)n), dim3(256), 0, static_cast<hipStream_t>(stream), x, h, w,
                     grid_h, grid_w, n_bins, limit, out);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_normalize_patterns(float* x, int n, int n_pix, int mode, void* stream) {
  if (!x) return EBSD_ERR_NULL;
  if (n < 0 || n_pix < 2 || mode < 0 || mode > 2) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  hipLaunchKernelGGL(normalize_kernel, dim3((unsigned)n), dim3(256), 0, static_cast<hipStream_t>(stream), x, n_pix,
                     mode);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}
