// Device helpers shared by the projection kernels: quaternion ray rotation,
// the side-by-side Rosca-Lambert map (sphere.py:145-185) and the bilinear,
// border-padded, align_corners=False master-pattern lookup
// (master_pattern.py:197-206).
#pragma once

#include "common.h"

namespace ebsd_lambert {

// Rotate p by unit quaternion (w, x, y, z) with the exact operation order of
// qu_apply (quaternions.py:203-211): q * (1, p) * conj(q), vector part. Every
// product and sum rounds separately, as in the reference's elementwise torch
// ops; a rotation-matrix form is 4x cheaper but moves sampling coordinates by
// ~3e-7, enough to change values on steep master-pattern bands by ~1e-4.
__device__ __forceinline__ void rotate_ray(float aw, float ax, float ay, float az,
                                           float bx, float by, float bz,
                                           float& ox, float& oy, float& oz) {
#pragma clang fp contract(off)
  const float iw = ((aw - ax * bx) - ay * by) - az * bz;
  const float ix = ((aw * bx + ax) + ay * bz) - az * by;
  const float iy = ((aw * by - ax * bz) + ay) + az * bx;
  const float iz = ((aw * bz + ax * by) - ay * bx) + az;
  ox = (((-iw) * ax + ix * aw) - iy * az) + iz * ay;
  oy = (((-iw) * ay + ix * az) + iy * aw) - iz * ax;
  oz = (((-iw) * az - ix * ay) + iy * ax) + iz * aw;
}

// atan(t) for |t| <= 1: t * P(t^2), P the degree-10 minimax (least maximum
// absolute error) polynomial for atan(sqrt(s)) / sqrt(s) on s in [0, 1],
// rounded to float32; fit_atan_unit.py beside this file writes these lines
// and measures the error over every float32 in [0, 1]: at most 1.415 ulp of
// float32 with this Horner evaluation by fused multiply-adds.
__device__ __forceinline__ float atan_unit(float t) {
  const float s = t * t;
  float p = 1.11207587e-03f;
  p = fmaf(p, s, -7.30408728e-03f);
  p = fmaf(p, s, 2.24965569e-02f);
  p = fmaf(p, s, -4.46185172e-02f);
  p = fmaf(p, s, 6.73463270e-02f);
  p = fmaf(p, s, -8.80690664e-02f);
  p = fmaf(p, s, 1.10566750e-01f);
  p = fmaf(p, s, -1.42794937e-01f);
  p = fmaf(p, s, 1.99996307e-01f);
  p = fmaf(p, s, -3.33333254e-01f);
  p = fmaf(p, s, 1.00000000e+00f);
  return p * t;
}

// Side-by-side Rosca-Lambert coordinates in grid_sample's [-1, 1] frame
// (sphere.py:161-185): gx spans the 2S-wide image (north half left), gy the
// S rows. The reference's atan2(sml * sgn, |big|) always has
// |sml| <= |big|, so it is atan of a ratio in [-1, 1].
__device__ __forceinline__ void lambert_sbs(float x, float y, float z, float& gx, float& gy) {
#pragma clang fp contract(off)
  const float factor = __builtin_amdgcn_sqrtf(fmaxf(2.0f * (1.0f - fabsf(z)), 0.0f));
  const bool cond = fabsf(y) <= fabsf(x);
  const float big = cond ? x : y;
  const float sml = cond ? y : x;
  const float sgn = (big < 0.0f) ? -1.0f : 1.0f;
  const float sf = sgn * factor;
  const float abig = fabsf(big);
  const float ratio = abig > 0.0f ? (sml * sgn) * __builtin_amdgcn_rcpf(abig) : 0.0f;
  const float simpler = sf * 0.70710678118654752440f;                         // 2 / sqrt(8)
  const float arct = (sf * atan_unit(ratio)) * 0.90031631615710606956f;      // 2 sqrt(2) / pi
  const float u = cond ? simpler : arct;
  gy = cond ? arct : simpler;
  gx = ((u * 0.5f) - 0.5f) + ((z < 0.0f) ? 1.0f : 0.0f);
}

// lambert_sbs with its Jacobian: j[0..2] = d gx / d(x, y, z), j[3..5] =
// d gy / d(x, y, z), the derivatives autograd takes through the reference's
// torch expressions (sphere.py:161-185): factor = sqrt(2 (1 - |z|)),
// atan2(sml sgn, |big|) with d/d sml = big / (big^2 + sml^2) and
// d/d big = -sml / (big^2 + sml^2). At the poles (factor 0, or big = 0) the
// derivatives are taken as 0 where autograd would give inf or NaN.
__device__ __forceinline__ void lambert_sbs_jac(float x, float y, float z, float& gx, float& gy, float (&j)[6]) {
  lambert_sbs(x, y, z, gx, gy);
  const float factor = __builtin_amdgcn_sqrtf(fmaxf(2.0f * (1.0f - fabsf(z)), 0.0f));
  const bool cond = fabsf(y) <= fabsf(x);
  const float big = cond ? x : y;
  const float sml = cond ? y : x;
  const float sgn = (big < 0.0f) ? -1.0f : 1.0f;
  const float abig = fabsf(big);
  const float ratio = abig > 0.0f ? (sml * sgn) / abig : 0.0f;
  const float c1 = 0.70710678118654752440f, c2 = 0.90031631615710606956f;
  const float zsign = z > 0.0f ? 1.0f : (z < 0.0f ? -1.0f : 0.0f);
  const float dfdz = factor > 0.0f ? -zsign / factor : 0.0f;
  const float r2 = big * big + sml * sml;
  const float d_sml = r2 > 0.0f ? sgn * factor * c2 * big / r2 : 0.0f;
  const float d_big = r2 > 0.0f ? -sgn * factor * c2 * sml / r2 : 0.0f;
  const float simpler_z = sgn * c1 * dfdz;
  const float arct_z = sgn * c2 * atan_unit(ratio) * dfdz;
  float du[3], dv[3];
  if (cond) {  // big = x, sml = y: u = simpler, v = arct
    du[0] = 0.0f; du[1] = 0.0f; du[2] = simpler_z;
    dv[0] = d_big; dv[1] = d_sml; dv[2] = arct_z;
  } else {     // big = y, sml = x: u = arct, v = simpler
    du[0] = d_sml; du[1] = d_big; du[2] = arct_z;
    dv[0] = 0.0f; dv[1] = 0.0f; dv[2] = simpler_z;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    j[i] = 0.5f * du[i];
    j[3 + i] = dv[i];
  }
}

// Bilinear sample of an (H, W) image at normalised (gx, gy) with border
// padding and align_corners=False: source index ((g + 1) * size - 1) / 2,
// clipped to [0, size - 1]; the four taps are blended as
// nw*v_nw + ne*v_ne + sw*v_sw + se*v_se (the order of torch's kernel).
__device__ __forceinline__ float sample_border(const float* __restrict__ img, int H, int W,
                                               float gx, float gy) {
#pragma clang fp contract(off)
  float ix = (gx + 1.0f) * (0.5f * (float)W) - 0.5f;
  float iy = (gy + 1.0f) * (0.5f * (float)H) - 0.5f;
  ix = fminf(fmaxf(ix, 0.0f), (float)(W - 1));
  iy = fminf(fmaxf(iy, 0.0f), (float)(H - 1));
  const float x0f = floorf(ix);
  const float y0f = floorf(iy);
  const int x0 = (int)x0f;
  const int y0 = (int)y0f;
  // A tap past the last column/row carries weight exactly 0; clamp its
  // address so it stays in bounds.
  const int x1 = min(x0 + 1, W - 1);
  const int y1 = min(y0 + 1, H - 1);
  const float w = ix - x0f;
  const float e = 1.0f - w;
  const float n = iy - y0f;
  const float s = 1.0f - n;
  // 32-bit offsets: the image holds fewer than 2^31 pixels (checked at entry)
  const float v_nw = __ldg(img + (y0 * W + x0));
  const float v_ne = __ldg(img + (y0 * W + x1));
  const float v_sw = __ldg(img + (y1 * W + x0));
  const float v_se = __ldg(img + (y1 * W + x1));
  return ((v_nw * (s * e) + v_ne * (s * w)) + v_sw * (n * e)) + v_se * (n * w);
}

// sample_border through a quad table (texel (y, x) holds the taps at (y, x),
// (y, x1), (y1, x), (y1, x1) with x1, y1 clamped like the scalar path): the
// same source index, weights, taps and blend order, so the same float32
// result bit for bit, from one 16-byte load instead of four.
__device__ __forceinline__ float sample_border_quad(const float4* __restrict__ quads, int H, int W,
                                                    float gx, float gy) {
#pragma clang fp contract(off)
  float ix = (gx + 1.0f) * (0.5f * (float)W) - 0.5f;
  float iy = (gy + 1.0f) * (0.5f * (float)H) - 0.5f;
  ix = fminf(fmaxf(ix, 0.0f), (float)(W - 1));
  iy = fminf(fmaxf(iy, 0.0f), (float)(H - 1));
  const float x0f = floorf(ix);
  const float y0f = floorf(iy);
  const int x0 = (int)x0f;
  const int y0 = (int)y0f;
  const float w = ix - x0f;
  const float e = 1.0f - w;
  const float n = iy - y0f;
  const float s = 1.0f - n;
  const float4 t = quads[y0 * W + x0];
  return ((t.x * (s * e) + t.y * (s * w)) + t.z * (n * e)) + t.w * (n * w);
}

// Source position of a bilinear, border, align_corners=False sample split
// into the quad-table texel (y0 * W + x0) and the two fractional offsets:
// the arithmetic of sample_border_quad, so quad_blend(quads[texel], w, n)
// equals sample_border_quad(quads, H, W, gx, gy) bit for bit. The clip to
// [0, size - 1] is one v_med3 (the same value as fmin(fmax(., 0), size - 1)
// for the finite coordinates the Lambert map produces).
__device__ __forceinline__ int quad_source(int H, int W, float gx, float gy, float& w, float& n) {
#pragma clang fp contract(off)
  float ix = (gx + 1.0f) * (0.5f * (float)W) - 0.5f;
  float iy = (gy + 1.0f) * (0.5f * (float)H) - 0.5f;
  ix = __builtin_amdgcn_fmed3f(ix, 0.0f, (float)(W - 1));
  iy = __builtin_amdgcn_fmed3f(iy, 0.0f, (float)(H - 1));
  const float x0f = floorf(ix);
  const float y0f = floorf(iy);
  w = ix - x0f;
  n = iy - y0f;
  return (int)y0f * W + (int)x0f;
}

// The blend of sample_border_quad for a fetched texel t.
__device__ __forceinline__ float quad_blend(const float4 t, float w, float n) {
#pragma clang fp contract(off)
  const float e = 1.0f - w;
  const float s = 1.0f - n;
  return ((t.x * (s * e) + t.y * (s * w)) + t.z * (n * e)) + t.w * (n * w);
}

// x / |x| with the squared norm accumulated as fma(z, z, fma(y, y, x*x)),
// the rounding sequence of torch's float32 CPU vector norm over a length-3
// row (master_pattern.py:166-167).
__device__ __forceinline__ void renormalise(float& x, float& y, float& z) {
  const float nr = sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
  x = x / nr;
  y = y / nr;
  z = z / nr;
}

}  // namespace ebsd_lambert
