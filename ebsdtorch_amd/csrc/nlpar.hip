// Non-local pattern averaging (NLPAR) of a 2D scan (gfx950): the two passes
// over the patterns of preprocessing/nlpar.py:20-192.
//
// The scan is X (H_s, W_s, P) float32 (P pixels per pattern, point-major).
// For every scan point and every offset o of the (2k+1)^2 window (in the
// caller's order, the reference's sort by L1 length), the distance pass
// writes D[p][o] = sum_pixels (X[p + o] - X[p])^2, with points outside the
// scan read as zero patterns (the reference zero-pads the scan); the average
// pass writes Y[p] = sum_o W[p][o] X[p + o] in the offset order, each term
// multiplied then added in FP32 as the reference's `denoised += x * w`.
//
// Both passes tile the scan: a workgroup owns 16 x 4 scan points and stages
// the (16 + 2k) x (4 + 2k) halo patterns in LDS, 32 pixels at a time, so each
// pattern chunk is read from HBM once per tile (1 + halo/tile times overall)
// instead of once per offset. The distance pass stores a chunk pixel-major
// (lanes = points read consecutive words); the average pass point-major
// (lanes = pixels, so both the LDS reads and the stores of a wave are
// consecutive). The weights between the passes (noise estimate,
// normalisation, the coefficient fit, softmax) are small (points x offsets)
// and are the caller's.

#include <algorithm>

#include "common.h"

namespace {

constexpr int kTx = 16;    // scan points per tile along x
constexpr int kTy = 4;     // along y
constexpr int kChunk = 32; // pixels per LDS stage
constexpr int kMaxRad = 5;

// Loads pixels [c0, c0 + 32) of the halo patterns of the tile at (x0, y0)
// into LDS: element (hp, j) at tile[hp * hstride + j * jstride].
__device__ __forceinline__ void load_halo(const float* __restrict__ X, int Hs, int Ws, int P, int k, int x0,
                                          int y0, int c0, float* tile, int hstride, int jstride, bool vec4) {
  const int hx = kTx + 2 * k, hy = kTy + 2 * k;
  const int n = hx * hy * (kChunk / 4);
  for (int idx = threadIdx.x; idx < n; idx += 256) {
    const int hp = idx >> 3, q = idx & 7;
    const int gy = y0 - k + hp / hx, gx = x0 - k + hp % hx;
    const bool in = gy >= 0 && gy < Hs && gx >= 0 && gx < Ws;
    const int c = c0 + 4 * q;
    const float* src = in ? X + ((size_t)gy * Ws + gx) * P + c : X;
    float v[4];
    if (in && vec4 && c + 4 <= P) {
      const float4 t = *reinterpret_cast<const float4*>(src);
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = (in && c + e < P) ? src[e] : 0.0f;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) tile[hp * hstride + (4 * q + e) * jstride] = v[e];
  }
}

template <int KR>
__global__ __launch_bounds__(256) void nlpar_dist_kernel(const float* __restrict__ X, int Hs, int Ws, int P,
                                                         const int* __restrict__ offs, float* __restrict__ D,
                                                         bool vec4) {
  constexpr int HX = kTx + 2 * KR, HY = kTy + 2 * KR, NH = HX * HY;
  constexpr int NOFF = (2 * KR + 1) * (2 * KR + 1);
  constexpr int NA = (NOFF + 3) / 4;
  __shared__ float tile[kChunk * NH];  // pixel-major: tile[j * NH + hp]
  const int x0 = blockIdx.x * kTx, y0 = blockIdx.y * kTy;
  const int p = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int px = p & (kTx - 1), py = p >> 4;
  const int hc = (py + KR) * HX + px + KR;
  int oh[NA];
  float acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int o = min(g + 4 * a, NOFF - 1);
    // offsets are clamped to the window so every LDS read stays in the tile
    oh[a] = hc + min(max(offs[2 * o], -KR), KR) * HX + min(max(offs[2 * o + 1], -KR), KR);
    acc[a] = 0.0f;
  }
  for (int c0 = 0; c0 < P; c0 += kChunk) {
    load_halo(X, Hs, Ws, P, KR, x0, y0, c0, tile, 1, NH, vec4);
    __syncthreads();
    const int nj = min(kChunk, P - c0);
    for (int j = 0; j < nj; ++j) {
      const float* row = tile + j * NH;
      const float cv = row[hc];
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        const float d = row[oh[a]] - cv;
        acc[a] = fmaf(d, d, acc[a]);
      }
    }
    __syncthreads();
  }
  const int gx = x0 + px, gy = y0 + py;
  if (gx < Ws && gy < Hs) {
    float* out = D + ((size_t)gy * Ws + gx) * NOFF;
#pragma unroll
    for (int a = 0; a < NA; ++a)
      if (g + 4 * a < NOFF) out[g + 4 * a] = acc[a];
  }
}

template <int KR>
__global__ __launch_bounds__(256) void nlpar_avg_kernel(const float* __restrict__ X, int Hs, int Ws, int P,
                                                        const int* __restrict__ offs, const float* __restrict__ Wt,
                                                        float* __restrict__ Y, bool vec4) {
  constexpr int HX = kTx + 2 * KR, HY = kTy + 2 * KR, NH = HX * HY;
  constexpr int NOFF = (2 * KR + 1) * (2 * KR + 1);
  __shared__ float tile[NH * kChunk];     // point-major: tile[hp * 32 + j]
  __shared__ float wts[kTx * kTy * NOFF]; // weights of the tile's points
  __shared__ int ofh[NOFF];               // halo offset of each window offset
  const int x0 = blockIdx.x * kTx, y0 = blockIdx.y * kTy;
  for (int e = threadIdx.x; e < kTx * kTy * NOFF; e += 256) {
    const int p = e / NOFF, o = e - p * NOFF;
    const int gx = x0 + (p & (kTx - 1)), gy = y0 + (p >> 4);
    wts[e] = (gx < Ws && gy < Hs) ? Wt[((size_t)gy * Ws + gx) * NOFF + o] : 0.0f;
  }
  for (int o = threadIdx.x; o < NOFF; o += 256)
    ofh[o] = min(max(offs[2 * o], -KR), KR) * HX + min(max(offs[2 * o + 1], -KR), KR);
  const int j = threadIdx.x & 31, pg = threadIdx.x >> 5;  // pixel in the chunk, point group
  for (int c0 = 0; c0 < P; c0 += kChunk) {
    __syncthreads();  // previous chunk's reads (and, first time, the weights) done
    load_halo(X, Hs, Ws, P, KR, x0, y0, c0, tile, kChunk, 1, vec4);
    __syncthreads();
    if (c0 + j < P) {
      for (int p = pg; p < kTx * kTy; p += 8) {
        const int px = p & (kTx - 1), py = p >> 4;
        const int gx = x0 + px, gy = y0 + py;
        if (gx >= Ws || gy >= Hs) continue;
        const int hc = (py + KR) * HX + px + KR;
        const float* w = wts + p * NOFF;
        float s = 0.0f;
        for (int o = 0; o < NOFF; ++o) s = __fadd_rn(s, __fmul_rn(tile[(hc + ofh[o]) * kChunk + j], w[o]));
        Y[((size_t)gy * Ws + gx) * P + c0 + j] = s;
      }
    }
  }
}

bool check_offsets_shape(int Hs, int Ws, int P, int k) {
  return Hs >= 1 && Ws >= 1 && P >= 1 && k >= 1 && k <= kMaxRad;
}

}  // namespace

extern "C" int ebsd_nlpar_distances(const float* x, int scan_h, int scan_w, int n_pix, int k_rad, const int* offsets,
                                    float* dists, void* stream) {
  if (!x || !offsets || !dists) return EBSD_ERR_NULL;
  if (!check_offsets_shape(scan_h, scan_w, n_pix, k_rad)) return EBSD_ERR_SHAPE;
  const dim3 grid((unsigned)((scan_w + kTx - 1) / kTx), (unsigned)((scan_h + kTy - 1) / kTy));
  const bool vec4 = (n_pix % 4) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (k_rad) {
#define EBSD_NLPAR_D(K) \
  case K: hipLaunchKernelGGL(nlpar_dist_kernel<K>, grid, dim3(256), 0, st, x, scan_h, scan_w, n_pix, offsets, dists, vec4); break;
    EBSD_NLPAR_D(1) EBSD_NLPAR_D(2) EBSD_NLPAR_D(3) EBSD_NLPAR_D(4) EBSD_NLPAR_D(5)
#undef EBSD_NLPAR_D
  }
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_nlpar_average(const float* x, int scan_h, int scan_w, int n_pix, int k_rad, const int* offsets,
                                  const float* weights, float* out, void* stream) {
  if (!x || !offsets || !weights || !out) return EBSD_ERR_NULL;
  if (!check_offsets_shape(scan_h, scan_w, n_pix, k_rad)) return EBSD_ERR_SHAPE;
  const dim3 grid((unsigned)((scan_w + kTx - 1) / kTx), (unsigned)((scan_h + kTy - 1) / kTy));
  const bool vec4 = (n_pix % 4) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (k_rad) {
#define EBSD_NLPAR_A(K) \
  case K: hipLaunchKernelGGL(nlpar_avg_kernel<K>, grid, dim3(256), 0, st, x, scan_h, scan_w, n_pix, offsets, weights, out, vec4); break;
    EBSD_NLPAR_A(1) EBSD_NLPAR_A(2) EBSD_NLPAR_A(3) EBSD_NLPAR_A(4) EBSD_NLPAR_A(5)
#undef EBSD_NLPAR_A
  }
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}
