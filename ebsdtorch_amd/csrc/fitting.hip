// Mutual-information maps of the pattern-centre fitter's "mi" metric
// (ebsd/fitting.py:486-496, utils/mutual_information.py:304-673, cmif with 4
// + 4 bins and per-shift marginals) for gfx950.
//
// The fitter correlates level-indicator images: channel 0 of a destination
// (dictionary) pattern is its foreground indicator, channels 1-4 mark the
// pixels whose value falls in each of the 4 bins, and the same for a source
// (experimental) pattern. Circular cross-correlation of every destination
// channel i with every source channel j (FFT products, done by rocFFT
// outside) gives, at each shift, the pixel-pair counts c[i][j]. Per shift
// this kernel turns the 25 counts into the mutual information:
//   counts below 1 are zeroed (FFT round-off);
//   joint  = c[1..4][1..4] / its sum,   H_joint = -sum p log p (p > 0);
//   m_src  = c[1..4][0]    / its sum,   H_src   (the destination levels
//            within the source foreground: the reference's "marg_hists_src");
//   m_dst  = c[0][1..4]    / its sum,   H_dst;
//   mi = H_dst + H_src - H_joint, and 0 where that is NaN, infinite or < 0.
// A histogram whose counts are all zero contributes entropy 0 (the
// reference's 0 / 0 entries fail its p > 0 test).

#include "common.h"

namespace {

constexpr int kLevels = 5;  // indicator + 4 bins

__device__ __forceinline__ float entropy4(const float (&h)[4]) {
  const float s = ((h[0] + h[1]) + h[2]) + h[3];
  float e = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float p = h[k] / s;
    if (p > 0.0f) e += p * logf(p);
  }
  return -e;
}

__global__ __launch_bounds__(256) void cmif_kernel(const float* __restrict__ corr, long long n_maps, int n_pix,
                                                   float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n_maps * (long long)n_pix) return;
  const long long m = t / n_pix;
  const int p = (int)(t - m * n_pix);
  const float* c = corr + m * (long long)(kLevels * kLevels) * n_pix + p;
  float v[kLevels][kLevels];
#pragma unroll
  for (int i = 0; i < kLevels; ++i)
#pragma unroll
    for (int j = 0; j < kLevels; ++j) {
      const float x = __builtin_nontemporal_load(c + (size_t)(i * kLevels + j) * n_pix);
      v[i][j] = x < 1.0f ? 0.0f : x;
    }
  float js = 0.0f;
#pragma unroll
  for (int i = 1; i < kLevels; ++i)
#pragma unroll
    for (int j = 1; j < kLevels; ++j) js += v[i][j];
  float hj = 0.0f;
#pragma unroll
  for (int i = 1; i < kLevels; ++i)
#pragma unroll
    for (int j = 1; j < kLevels; ++j) {
      const float q = v[i][j] / js;
      if (q > 0.0f) hj += q * logf(q);
    }
  const float src[4] = {v[1][0], v[2][0], v[3][0], v[4][0]};
  const float dst[4] = {v[0][1], v[0][2], v[0][3], v[0][4]};
  const float mi = (entropy4(dst) + entropy4(src)) - (-hj);
  out[t] = (isnan(mi) || isinf(mi) || mi < 0.0f) ? 0.0f : mi;
}

}  // namespace

extern "C" int ebsd_cmif_maps(const float* corr, long long n_maps, int n_pix, float* out, void* stream) {
  if (!corr || !out) return EBSD_ERR_NULL;
  if (n_maps < 0 || n_pix < 1) return EBSD_ERR_SHAPE;
  const long long total = n_maps * (long long)n_pix;
  if (total == 0) return EBSD_OK;
  const long long blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return EBSD_ERR_SHAPE;
  hipLaunchKernelGGL(cmif_kernel, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream), corr,
                     n_maps, n_pix, out);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}
