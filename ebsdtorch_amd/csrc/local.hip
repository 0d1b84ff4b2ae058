// Local-dictionary refinement for the learned-proposal form (gfx950).
//
// For query pattern i with proposal quaternion c_i and a fixed set of small
// rotations g_j (a local orientation grid, ebsd/refining.py:13-68, composed
// as qu_prod(g_j, c_i) like ebsd/refining.py:114), every candidate pattern
// p_ij(d) = MP(lambert(R(g_j c_i) ray_d)) is generated on the fly and scored
// against the query x_i by the cosine of the zero-mean vectors
//   (sum x p - sum p * sum x / D) / (|x| sqrt(sum p^2 - (sum p)^2 / D)),
// with |x| taken as 1 (the caller passes unit-norm queries). Only the best
// `topk` (score, j) per query are written: candidate patterns never touch
// memory. The work is D bilinear samples per candidate, gather/VALU bound.
//
// Layout: one workgroup per query; each lane owns one candidate and walks
// all D pixels, so the per-pixel data (ray, query value) is wave-uniform and
// comes through scalar loads, and no reduction crosses lanes. A wave's 64
// candidates are neighbours in the grid (the host orders the grid along a
// space-filling curve), so for one pixel the 64 lanes sample a compact patch
// of the master pattern and the gathers share cache lines. The master
// pattern is read as a quad table: texel (y, x) holds the four bilinear taps
// (y, x), (y, x+1), (y+1, x), (y+1, x+1) (indices clamped to the image like
// the border lookup), so a sample is one 16-byte load.
//
// Sums are taken about the lane's first sample p0 (sum of p - p0 etc.; the
// cosine is unchanged by the shift) in four interleaved partial sums, which
// keeps the variance term free of cancellation.

#include <algorithm>

#include "common.h"
#include "lambert.h"

#include <hip/hip_fp16.h>

namespace {

using namespace ebsd_lambert;

constexpr int kLocalWaves = 4;
constexpr int kMaxGrid = 4096;

struct Rot {
  float m[9];
};

// Active rotation matrix of q = a * b (Hamilton product, w first), after
// normalising q; v' = R v equals the vector part of q (0, v) q*.
__device__ __forceinline__ Rot compose_matrix(const float* __restrict__ a, const float* __restrict__ b) {
  float w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  float x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  float y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  float z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
  const float inv = rsqrtf(w * w + x * x + y * y + z * z);
  w *= inv;
  x *= inv;
  y *= inv;
  z *= inv;
  Rot r;
  r.m[0] = 1.0f - 2.0f * (y * y + z * z);
  r.m[1] = 2.0f * (x * y - w * z);
  r.m[2] = 2.0f * (x * z + w * y);
  r.m[3] = 2.0f * (x * y + w * z);
  r.m[4] = 1.0f - 2.0f * (x * x + z * z);
  r.m[5] = 2.0f * (y * z - w * x);
  r.m[6] = 2.0f * (x * z - w * y);
  r.m[7] = 2.0f * (y * z + w * x);
  r.m[8] = 1.0f - 2.0f * (x * x + y * y);
  return r;
}

// Side-by-side Lambert map + bilinear border lookup (align_corners=False)
// through the quad table, for a unit direction v. The same mapping as
// lambert_sbs + sample_border, with the grid_sample source-index affine maps
// folded into the Lambert output (ix = u W/4 + W/4 - 1/2 + hemisphere W/2,
// iy = v H/2 + H/2 - 1/2), clamps as med3 and contraction allowed: the
// local scores only rank candidates and need no bit-level agreement with
// the reference's rounding sequence.
template <typename QuadT>
__device__ __forceinline__ float4 load_quad(const QuadT* __restrict__ quads, uint32_t texel);
template <>
__device__ __forceinline__ float4 load_quad<float4>(const float4* __restrict__ quads, uint32_t texel) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const unsigned char*>(quads) + (texel << 4));
}
template <>
__device__ __forceinline__ float4 load_quad<ushort4>(const ushort4* __restrict__ quads, uint32_t texel) {
  const ushort4 h = *reinterpret_cast<const ushort4*>(reinterpret_cast<const unsigned char*>(quads) + (texel << 3));
  return make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                     __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
}

// Source-index affine maps of the side-by-side Lambert coordinates (u, v) in
// [-1, 1]^2 of one hemisphere: ix = u qx + (x0, or xs in the southern
// hemisphere), iy = v qy + y0.
struct LambertAffine {
  float qx, qy, x0, xs, y0;
};

// grid_sample with align_corners=False on an (H, W) side-by-side image:
// ix = u W/4 + W/4 - 1/2 (+ W/2 south), iy = v H/2 + H/2 - 1/2.
__device__ __forceinline__ LambertAffine affine_unaligned(int H, int W) {
  LambertAffine a;
  a.qx = 0.25f * (float)W;
  a.qy = 0.5f * (float)H;
  a.x0 = a.qx - 0.5f;
  a.xs = a.x0 + 0.5f * (float)W;
  a.y0 = a.qy - 0.5f;
  return a;
}

// align_corners=True (master_pattern.py:197-206 as called by refining.py:200):
// ix = (u/4 + 1/4 (+ 1/2 south)) (W - 1), iy = (v/2 + 1/2) (H - 1).
__device__ __forceinline__ LambertAffine affine_aligned(int H, int W) {
  LambertAffine a;
  a.qx = 0.25f * (float)(W - 1);
  a.qy = 0.5f * (float)(H - 1);
  a.x0 = a.qx;
  a.xs = a.qx + 0.5f * (float)(W - 1);
  a.y0 = a.qy;
  return a;
}

template <typename QuadT>
__device__ __forceinline__ float sample_fast(const QuadT* __restrict__ quads, int H, int W, const LambertAffine& af,
                                             float vx, float vy, float vz) {
  const float ax = fabsf(vx), ay = fabsf(vy);
  const float factor = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-2.0f, fabsf(vz), 2.0f), 0.0f));
  const bool cond = ay <= ax;
  const float big = cond ? vx : vy;
  const float sml = cond ? vy : vx;
  const float abig = cond ? ax : ay;
  const float sf = big < 0.0f ? -factor : factor;
  const float ssml = big < 0.0f ? -sml : sml;
  const float ratio = abig > 0.0f ? ssml * __builtin_amdgcn_rcpf(abig) : 0.0f;
  const float arct = sf * (atan_unit(ratio) * 0.90031631615710606956f);  // 2 sqrt(2) / pi
  const float simpler = sf * 0.70710678118654752440f;                    // 2 / sqrt(8)
  const float u = cond ? simpler : arct;
  const float v = cond ? arct : simpler;
  float ix = fmaf(u, af.qx, (vz < 0.0f) ? af.xs : af.x0);
  float iy = fmaf(v, af.qy, af.y0);
  ix = __builtin_amdgcn_fmed3f(ix, 0.0f, (float)(W - 1));
  iy = __builtin_amdgcn_fmed3f(iy, 0.0f, (float)(H - 1));
  const float x0f = floorf(ix);
  const float y0f = floorf(iy);
  const float w = ix - x0f;
  const float n = iy - y0f;
  // texel index y0 W + x0 is exact in float32 (fewer than 2^24 texels, checked
  // at entry); 32-bit byte offsets
  const float4 t = load_quad<QuadT>(quads, (uint32_t)fmaf(y0f, (float)W, x0f));
  const float top = fmaf(w, t.y - t.x, t.x);
  const float bot = fmaf(w, t.w - t.z, t.z);
  return fmaf(n, bot - top, top);
}

template <typename QuadT>
__device__ __forceinline__ float sample_rotated(const QuadT* __restrict__ quads, int H, int W,
                                                const LambertAffine& af, const Rot& R, float rx, float ry,
                                                float rz) {
  const float vx = fmaf(R.m[0], rx, fmaf(R.m[1], ry, R.m[2] * rz));
  const float vy = fmaf(R.m[3], rx, fmaf(R.m[4], ry, R.m[5] * rz));
  const float vz = fmaf(R.m[6], rx, fmaf(R.m[7], ry, R.m[8] * rz));
  return sample_fast(quads, H, W, af, vx, vy, vz);
}

template <typename QuadT>
__device__ __forceinline__ float sample_candidate(const QuadT* __restrict__ quads, int H, int W, const Rot& R,
                                                  float rx, float ry, float rz) {
  return sample_rotated(quads, H, W, affine_unaligned(H, W), R, rx, ry, rz);
}

template <typename QuadT>
__device__ __forceinline__ QuadT make_quad(float a, float b, float c, float d);
template <>
__device__ __forceinline__ float4 make_quad<float4>(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
template <>
__device__ __forceinline__ ushort4 make_quad<ushort4>(float a, float b, float c, float d) {
  return make_ushort4(__half_as_ushort(__float2half(a)), __half_as_ushort(__float2half(b)),
                      __half_as_ushort(__float2half(c)), __half_as_ushort(__float2half(d)));
}

template <typename QuadT>
__global__ __launch_bounds__(256) void quad_table_kernel(const float* __restrict__ mp, int H, int W,
                                                         QuadT* __restrict__ quads) {
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < (int64_t)H * W;
       idx += (int64_t)gridDim.x * 256) {
    const int y = (int)(idx / W), x = (int)(idx % W);
    const int x1 = min(x + 1, W - 1), y1 = min(y + 1, H - 1);
    quads[idx] = make_quad<QuadT>(mp[y * W + x], mp[y * W + x1], mp[y1 * W + x], mp[y1 * W + x1]);
  }
}

template <typename QuadT>
__global__ __launch_bounds__(kLocalWaves * 64) void local_refine_kernel(
    const QuadT* __restrict__ quads, int H, int W, const float* __restrict__ rays, int n_pix,
    const float* __restrict__ Q, int ld, const float* __restrict__ centers, const float* __restrict__ grid,
    int n_grid, int topk, float* __restrict__ out_val, int* __restrict__ out_idx) {
  __shared__ float scores[kMaxGrid];
  __shared__ float red[4];
  const int i = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const float* __restrict__ xrow = Q + (size_t)i * ld;

  float sx = 0.0f;
  for (int d = tid; d < n_pix; d += kLocalWaves * 64) sx += xrow[d];
  const float Sx = block_sum_256(sx, red);
  const float inv_d = 1.0f / (float)n_pix;

  float c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] = centers[(size_t)i * 4 + k];

  const int n_chunks = (n_grid + 63) / 64;
  for (int ch = wave; ch < n_chunks; ch += kLocalWaves) {
    const int j = ch * 64 + lane;
    const Rot R = compose_matrix(grid + (size_t)min(j, n_grid - 1) * 4, c);
    const float p0 = sample_candidate(quads, H, W, R, rays[0], rays[1], rays[2]);
    float sp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float spp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float sxp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int d = 0;
    for (; d + 4 <= n_pix; d += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* r = rays + 3 * (d + u);
        const float p = sample_candidate(quads, H, W, R, r[0], r[1], r[2]) - p0;
        sp[u] += p;
        spp[u] = fmaf(p, p, spp[u]);
        sxp[u] = fmaf(xrow[d + u], p, sxp[u]);
      }
    }
    for (; d < n_pix; ++d) {
      const float* r = rays + 3 * d;
      const float p = sample_candidate(quads, H, W, R, r[0], r[1], r[2]) - p0;
      sp[0] += p;
      spp[0] = fmaf(p, p, spp[0]);
      sxp[0] = fmaf(xrow[d], p, sxp[0]);
    }
    const float Sp = (sp[0] + sp[1]) + (sp[2] + sp[3]);
    const float Spp = (spp[0] + spp[1]) + (spp[2] + spp[3]);
    const float Sxp = (sxp[0] + sxp[1]) + (sxp[2] + sxp[3]);
    const float var = Spp - Sp * Sp * inv_d;
    const float num = Sxp - Sp * Sx * inv_d;
    if (j < n_grid) scores[j] = var > 0.0f ? num * rsqrtf(var) : 0.0f;
  }
  __syncthreads();

  // top-k by repeated wave arg-max over the LDS scores (higher score, then
  // lower index); a chosen entry is retired by setting it to -inf.
  if (wave == 0) {
    for (int k = 0; k < topk; ++k) {
      float bv = -INFINITY;
      int bi = INT_MAX;
      for (int j = lane; j < n_grid; j += 64) {
        const float v = scores[j];
        if (topk_better(v, j, bv, bi)) {
          bv = v;
          bi = j;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (topk_better(ov, oi, bv, bi)) {
          bv = ov;
          bi = oi;
        }
      }
      const bool found = bi != INT_MAX && bv != -INFINITY;
      if (lane == 0) {
        out_val[(size_t)i * topk + k] = found ? bv : -INFINITY;
        out_idx[(size_t)i * topk + k] = found ? bi : -1;
      }
      if (found && (bi & 63) == lane) scores[bi] = -INFINITY;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
  }
}

// Rescoring of k listed candidates per query (k <= 64), typically the top-k
// of a local search on binned patterns, against the full-resolution query.
// One wave per query: lane = slice * kp + c with kp the power of two >= k,
// so candidate c (lanes c >= k idle) is scored by 64 / kp lanes, slice s
// taking pixels s, s + 64/kp, ...; a pixel's query value and ray are then
// shared by the kp lanes of a slice and consecutive slices read consecutive
// pixels. The sums (about the candidate's first sample, as in the search)
// are combined over the slices by butterflies, and the best candidate of the
// query by a butterfly over c (higher score, then lower list position).
template <typename QuadT>
__global__ __launch_bounds__(kLocalWaves * 64) void local_rescore_kernel(
    const QuadT* __restrict__ quads, int H, int W, const float* __restrict__ rays, int n_pix,
    const float* __restrict__ Q, int ld, int m, const float* __restrict__ centers, const float* __restrict__ grid,
    int n_grid, const int* __restrict__ cand, int k, int kp, float* __restrict__ out_val,
    float* __restrict__ best_val, int* __restrict__ best_idx) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kLocalWaves + (threadIdx.x >> 6);
  if (i >= m) return;  // whole waves only; the kernel has no block barrier
  const int c = lane & (kp - 1);
  const int n_slices = 64 / kp;
  const int slice = lane / kp;
  const int j = cand[(size_t)i * k + min(c, k - 1)];
  const bool valid = c < k && j >= 0 && j < n_grid;
  const float* __restrict__ xrow = Q + (size_t)i * ld;

  const Rot R = compose_matrix(grid + (size_t)(valid ? j : 0) * 4, centers + (size_t)i * 4);
  const float p0 = sample_candidate(quads, H, W, R, rays[0], rays[1], rays[2]);
  float sx[2] = {0.0f, 0.0f}, sp[2] = {0.0f, 0.0f}, spp[2] = {0.0f, 0.0f}, sxp[2] = {0.0f, 0.0f};
  int d = slice;
  for (; d + n_slices < n_pix; d += 2 * n_slices) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = d + u * n_slices;
      const float* r = rays + 3 * e;
      const float xv = xrow[e];
      const float p = sample_candidate(quads, H, W, R, r[0], r[1], r[2]) - p0;
      sx[u] += xv;
      sp[u] += p;
      spp[u] = fmaf(p, p, spp[u]);
      sxp[u] = fmaf(xv, p, sxp[u]);
    }
  }
  if (d < n_pix) {
    const float* r = rays + 3 * d;
    const float xv = xrow[d];
    const float p = sample_candidate(quads, H, W, R, r[0], r[1], r[2]) - p0;
    sx[0] += xv;
    sp[0] += p;
    spp[0] = fmaf(p, p, spp[0]);
    sxp[0] = fmaf(xv, p, sxp[0]);
  }
  float Sx = sx[0] + sx[1], Sp = sp[0] + sp[1], Spp = spp[0] + spp[1], Sxp = sxp[0] + sxp[1];
  for (int off = kp; off < 64; off <<= 1) {
    Sx += __shfl_xor(Sx, off, 64);
    Sp += __shfl_xor(Sp, off, 64);
    Spp += __shfl_xor(Spp, off, 64);
    Sxp += __shfl_xor(Sxp, off, 64);
  }
  const float inv_d = 1.0f / (float)n_pix;
  const float var = Spp - Sp * Sp * inv_d;
  const float num = Sxp - Sp * Sx * inv_d;
  const float score = valid ? (var > 0.0f ? num * rsqrtf(var) : 0.0f) : -INFINITY;
  if (out_val && slice == 0 && c < k) out_val[(size_t)i * k + c] = score;

  float bv = score;
  int bc = c < k ? c : INT_MAX;
  for (int off = 1; off < kp; off <<= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oc = __shfl_xor(bc, off, 64);
    if (topk_better(ov, oc, bv, bc)) {
      bv = ov;
      bc = oc;
    }
  }
  if (lane == 0) {
    const bool found = bc != INT_MAX && bv != -INFINITY;
    best_val[i] = found ? bv : -INFINITY;
    best_idx[i] = found ? cand[(size_t)i * k + bc] : -1;
  }
}

// q = standardised (w >= 0) Hamilton product a * b (quaternions.py:78-91).
__device__ __forceinline__ void qu_prod_std(const float* a, const float* b, float* q) {
  float w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  float x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  float y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  float z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
  const float s = w < 0.0f ? -1.0f : 1.0f;
  q[0] = s * w;
  q[1] = s * x;
  q[2] = s * y;
  q[3] = s * z;
}

// One iteration of the grid refinement of refining.py:172-237 for query
// pattern i: the candidates qu_std(g_j * c_i) (g_j the local grid at the
// pattern's shrink level), each moved to the fundamental zone as
// ori_to_fz_laue does (laue_fz_ori.py:35-81: the right product with the
// symmetry operator giving the largest |w|, first such operator), projected
// through the align_corners=True lookup of the (binned) master pattern and
// scored by the dot product of the zero-mean candidate with the zero-mean
// query (not normalised). Writes the best (score, j, reduced quaternion),
// ties to the lower j (torch.max). Layout of local_rescore_kernel: one wave
// per query, lane = slice * kp + c; candidates in chunks of kp.
template <typename QuadT>
__global__ __launch_bounds__(kLocalWaves * 64) void grid_refine_kernel(
    const QuadT* __restrict__ quads, int H, int W, const float* __restrict__ rays, int n_pix,
    const float* __restrict__ Q, int ld, int m, const float* __restrict__ centers, const float* __restrict__ grids,
    int n_grid, int n_levels, const int* __restrict__ level, const float* __restrict__ ops, int n_ops, int kp,
    float* __restrict__ best_val, int* __restrict__ best_idx, float* __restrict__ best_q) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kLocalWaves + (threadIdx.x >> 6);
  if (i >= m) return;  // whole waves only; the kernel has no block barrier
  const int c = lane & (kp - 1);
  const int n_slices = 64 / kp;
  const int slice = lane / kp;
  const float* __restrict__ xrow = Q + (size_t)i * ld;
  // levels outside [0, n_levels) are clamped (the grids buffer has n_levels)
  const float* __restrict__ grid = grids + (size_t)min(max(level[i], 0), n_levels - 1) * n_grid * 4;
  float ctr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) ctr[k] = centers[(size_t)i * 4 + k];
  const LambertAffine af = affine_aligned(H, W);
  const float inv_d = 1.0f / (float)n_pix;
  const float ident[4] = {1.0f, 0.0f, 0.0f, 0.0f};

  float bv = -INFINITY;
  int bj = INT_MAX;
  float bq[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int j0 = 0; j0 < n_grid; j0 += kp) {
    const int j = j0 + c;
    const bool valid = j < n_grid;
    float q[4], qf[4];
    qu_prod_std(grid + (size_t)min(j, n_grid - 1) * 4, ctr, q);
    int og = 0;
    float amax = -1.0f;
    for (int o = 0; o < n_ops; ++o) {
      const float* g = ops + 4 * o;
      const float a = fabsf(q[0] * g[0] - q[1] * g[1] - q[2] * g[2] - q[3] * g[3]);
      if (a > amax) {
        amax = a;
        og = o;
      }
    }
    qu_prod_std(q, ops + 4 * og, qf);
    const Rot R = compose_matrix(ident, qf);
    const float p0 = sample_rotated(quads, H, W, af, R, rays[0], rays[1], rays[2]);
    float sx = 0.0f, sp = 0.0f, sxp = 0.0f;
    for (int d = slice; d < n_pix; d += n_slices) {
      const float* r = rays + 3 * d;
      const float xv = xrow[d];
      const float p = sample_rotated(quads, H, W, af, R, r[0], r[1], r[2]) - p0;
      sx += xv;
      sp += p;
      sxp = fmaf(xv, p, sxp);
    }
    for (int off = kp; off < 64; off <<= 1) {
      sx += __shfl_xor(sx, off, 64);
      sp += __shfl_xor(sp, off, 64);
      sxp += __shfl_xor(sxp, off, 64);
    }
    // sum (x - mean x)(p - mean p) = Sxp - Sp Sx / D, unchanged by the p0 shift
    const float score = valid ? sxp - sp * sx * inv_d : -INFINITY;
    if (valid && topk_better(score, j, bv, bj)) {
      bv = score;
      bj = j;
#pragma unroll
      for (int k = 0; k < 4; ++k) bq[k] = qf[k];
    }
  }
  for (int off = 1; off < kp; off <<= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oj = __shfl_xor(bj, off, 64);
    float oq[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) oq[k] = __shfl_xor(bq[k], off, 64);
    if (topk_better(ov, oj, bv, bj)) {
      bv = ov;
      bj = oj;
#pragma unroll
      for (int k = 0; k < 4; ++k) bq[k] = oq[k];
    }
  }
  if (lane == 0) {
    const bool found = bj != INT_MAX;
    best_val[i] = found ? bv : -INFINITY;
    best_idx[i] = found ? bj : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) best_q[(size_t)i * 4 + k] = found ? bq[k] : ctr[k];
  }
}

}  // namespace

extern "C" int ebsd_mp_quad_table(const float* mp, int mp_rows, int mp_cols, void* quads, int dtype,
                                  void* stream) {
  if (!mp || !quads) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || (long long)mp_rows * mp_cols > INT_MAX) return EBSD_ERR_SHAPE;
  if ((reinterpret_cast<uintptr_t>(quads) & 15) != 0) return EBSD_ERR_ALIGN;
  const int64_t n = (int64_t)mp_rows * mp_cols;
  const dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(quad_table_kernel<float4>, grid, dim3(256), 0, st, mp, mp_rows, mp_cols,
                       static_cast<float4*>(quads));
  else if (dtype == EBSD_DTYPE_F16)
    hipLaunchKernelGGL(quad_table_kernel<ushort4>, grid, dim3(256), 0, st, mp, mp_rows, mp_cols,
                       static_cast<ushort4*>(quads));
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_local_refine(const void* quads, int quad_dtype, int mp_rows, int mp_cols, const float* rays,
                                 int n_pix, const float* q, int m, int ld, const float* centers, const float* grid,
                                 int n_grid, int topk, float* out_val, int* out_idx, void* stream) {
  if (!quads || !rays || !q || !centers || !grid || !out_val || !out_idx) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || (long long)mp_rows * mp_cols >= (1LL << 24)) return EBSD_ERR_SHAPE;
  if (n_pix < 1 || ld < n_pix || m < 0 || n_grid < 1 || n_grid > kMaxGrid) return EBSD_ERR_SHAPE;
  if (topk < 1 || topk > EBSD_MAX_TOPK) return EBSD_ERR_TOPK;
  if ((reinterpret_cast<uintptr_t>(quads) & 15) != 0) return EBSD_ERR_ALIGN;
  if (m == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(m), b(kLocalWaves * 64);
  if (quad_dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(local_refine_kernel<float4>, g, b, 0, st, static_cast<const float4*>(quads), mp_rows,
                       mp_cols, rays, n_pix, q, ld, centers, grid, n_grid, topk, out_val, out_idx);
  else if (quad_dtype == EBSD_DTYPE_F16)
    hipLaunchKernelGGL(local_refine_kernel<ushort4>, g, b, 0, st, static_cast<const ushort4*>(quads), mp_rows,
                       mp_cols, rays, n_pix, q, ld, centers, grid, n_grid, topk, out_val, out_idx);
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_local_rescore(const void* quads, int quad_dtype, int mp_rows, int mp_cols, const float* rays,
                                  int n_pix, const float* q, int m, int ld, const float* centers, const float* grid,
                                  int n_grid, const int* cand, int k, float* out_val, float* best_val,
                                  int* best_idx, void* stream) {
  if (!quads || !rays || !q || !centers || !grid || !cand || !best_val || !best_idx) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || (long long)mp_rows * mp_cols >= (1LL << 24)) return EBSD_ERR_SHAPE;
  if (n_pix < 1 || ld < n_pix || m < 0 || n_grid < 1) return EBSD_ERR_SHAPE;
  if (k < 1 || k > 64) return EBSD_ERR_TOPK;
  if ((reinterpret_cast<uintptr_t>(quads) & 15) != 0) return EBSD_ERR_ALIGN;
  if (m == 0) return EBSD_OK;
  int kp = 1;
  while (kp < k) kp <<= 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g((unsigned)((m + kLocalWaves - 1) / kLocalWaves)), b(kLocalWaves * 64);
  if (quad_dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(local_rescore_kernel<float4>, g, b, 0, st, static_cast<const float4*>(quads), mp_rows,
                       mp_cols, rays, n_pix, q, ld, m, centers, grid, n_grid, cand, k, kp, out_val, best_val,
                       best_idx);
  else if (quad_dtype == EBSD_DTYPE_F16)
    hipLaunchKernelGGL(local_rescore_kernel<ushort4>, g, b, 0, st, static_cast<const ushort4*>(quads), mp_rows,
                       mp_cols, rays, n_pix, q, ld, m, centers, grid, n_grid, cand, k, kp, out_val, best_val,
                       best_idx);
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_grid_refine_step(const void* quads, int quad_dtype, int mp_rows, int mp_cols, const float* rays,
                                     int n_pix, const float* q, int m, int ld, const float* centers,
                                     const float* grids, int n_grid, int n_levels, const int* level,
                                     const float* ops, int n_ops, float* best_val, int* best_idx, float* best_q,
                                     void* stream) {
  if (!quads || !rays || !q || !centers || !grids || !level || !ops || !best_val || !best_idx || !best_q)
    return EBSD_ERR_NULL;
  if (mp_rows < 2 || mp_cols < 2 || (long long)mp_rows * mp_cols >= (1LL << 24)) return EBSD_ERR_SHAPE;
  if (n_pix < 1 || ld < n_pix || m < 0 || n_grid < 1 || n_levels < 1 || n_ops < 1 || n_ops > 64)
    return EBSD_ERR_SHAPE;
  if ((reinterpret_cast<uintptr_t>(quads) & 15) != 0) return EBSD_ERR_ALIGN;
  if (m == 0) return EBSD_OK;
  int kp = 1;
  while (kp < n_grid && kp < 64) kp <<= 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g((unsigned)((m + kLocalWaves - 1) / kLocalWaves)), b(kLocalWaves * 64);
  if (quad_dtype == EBSD_DTYPE_F32)
    hipLaunchKernelGGL(grid_refine_kernel<float4>, g, b, 0, st, static_cast<const float4*>(quads), mp_rows, mp_cols,
                       rays, n_pix, q, ld, m, centers, grids, n_grid, n_levels, level, ops, n_ops, kp, best_val,
                       best_idx, best_q);
  else if (quad_dtype == EBSD_DTYPE_F16)
    hipLaunchKernelGGL(grid_refine_kernel<ushort4>, g, b, 0, st, static_cast<const ushort4*>(quads), mp_rows,
                       mp_cols, rays, n_pix, q, ld, m, centers, grids, n_grid, n_levels, level, ops, n_ops, kp,
                       best_val, best_idx, best_q);
  else
    return EBSD_ERR_DTYPE;
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}
