// Counter-based noise for on-the-fly pattern simulation (gfx950).
//
// Detector noise model of the synthetic workloads: for a clean pattern p
// (n_pix values) and a per-pattern mean count c > 0,
//   counts_d ~ Poisson(c * p_d / mean(p)),   out_d = counts_d / c + sigma * N(0, 1);
// c <= 0 leaves out Poisson noise: out_d = p_d + sigma * N(0, 1).
// Random numbers come from Philox4x32-10 (Salmon, Moraes, Dror & Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by the seed
// and counted by (pattern, pixel, attempt), so the output is a pure function
// of (inputs, seed, offset) whatever the launch shape. Poisson variates use
// multiplication of uniforms for rates below 10 and Hormann's transformed
// rejection with squeeze (PTRS; W. Hormann, "The transformed rejection method
// for generating Poisson random variables", Insurance: Mathematics and
// Economics 12 (1993)) above; normals use the Box-Muller transform.

#include "common.h"

namespace {

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// (0, 1] and [0, 1) uniforms from 24 random bits.
__device__ __forceinline__ float u01_open0(uint32_t v) { return ((v >> 8) + 1u) * (1.0f / 16777216.0f); }
__device__ __forceinline__ float u01(uint32_t v) { return (v >> 8) * (1.0f / 16777216.0f); }

struct Rng {
  uint32_t k0, k1;       // key (seed)
  uint32_t c_hi;         // stream offset
  uint32_t c_pat, c_pix;  // pattern and pixel counters
  uint32_t draw;         // attempt counter
  __device__ __forceinline__ U4 next() {
    const U4 c{c_pix, c_pat, draw++, c_hi};
    return philox4x32_10(c, k0, k1);
  }
};

__device__ int poisson(float lam, Rng& rng) {
  if (!(lam > 0.0f)) return 0;
  if (lam < 10.0f) {
    // count uniforms until their product drops below exp(-lam)
    const float lim = expf(-lam);
    float prod = 1.0f;
    int k = -1;
    while (true) {
      const U4 r = rng.next();
      const uint32_t v[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        prod *= u01_open0(v[i]);
        ++k;
        if (prod <= lim) return k;
      }
    }
  }
  const float slam = sqrtf(lam);
  const float loglam = logf(lam);
  const float b = 0.931f + 2.53f * slam;
  const float a = -0.059f + 0.02483f * b;
  const float inv_alpha = 1.1239f + 1.1328f / (b - 3.4f);
  const float vr = 0.9277f - 3.6224f / (b - 2.0f);
  for (int it = 0; it < 64; ++it) {
    const U4 r = rng.next();
    const float U = u01(r.x) - 0.5f;
    const float V = u01_open0(r.y);
    const float us = 0.5f - fabsf(U);
    const float kf = floorf((2.0f * a / us + b) * U + lam + 0.43f);
    if (us >= 0.07f && V <= vr) return (int)kf;
    if (kf < 0.0f || (us < 0.013f && V > us)) continue;
    if (logf(V) + logf(inv_alpha) - logf(a / (us * us) + b) <= -lam + kf * loglam - lgammaf(kf + 1.0f))
      return (int)kf;
  }
  return (int)floorf(lam + 0.5f);  // not reached in practice (acceptance > 0.88)
}

template <int ITEMS>
__global__ __launch_bounds__(256) void noise_kernel(const float* __restrict__ clean, int n_pix, int ld_in,
                                                    float* __restrict__ out, int ld_out,
                                                    const float* __restrict__ mean_counts, float sigma,
                                                    uint32_t k0, uint32_t k1, uint32_t offset) {
  __shared__ float red[4];
  const int n = blockIdx.x;
  const float* src = clean + (size_t)n * ld_in;
  float vals[ITEMS];
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const int d = threadIdx.x + i * 256;
    vals[i] = d < n_pix ? src[d] : 0.0f;
    s += vals[i];
  }
  const float mean = block_sum_256(s, red) / (float)n_pix;
  // a non-positive mean count switches the Poisson stage off (out = clean +
  // Gaussian); a pattern with non-positive mean gets rate 0
  const float c = mean_counts[n];
  const bool poisson_on = c > 0.0f;
  const float scale = (poisson_on && mean > 0.0f) ? c / mean : 0.0f;
  float* dst = out + (size_t)n * ld_out;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const int d = threadIdx.x + i * 256;
    if (d < n_pix) {
      Rng rng{k0, k1, offset, (uint32_t)n, (uint32_t)d, 1u};
      const float signal = poisson_on ? (float)poisson(vals[i] * scale, rng) / c : vals[i];
      // the Gaussian uses its own counter slot (draw 0)
      const U4 g = philox4x32_10(U4{(uint32_t)d, (uint32_t)n, 0u, offset}, k0, k1);
      const float r = sqrtf(-2.0f * logf(u01_open0(g.x)));
      const float z = r * cospif(2.0f * u01(g.y));
      dst[d] = signal + sigma * z;
    }
  }
}

}  // namespace

extern "C" int ebsd_poisson_gaussian_noise(const float* clean, int n, int n_pix, int ld_in, float* out,
                                           int ld_out, const float* mean_counts, float sigma,
                                           unsigned long long seed, unsigned long long offset, void* stream) {
  if (!clean || !out || !mean_counts) return EBSD_ERR_NULL;
  if (n < 0 || n_pix < 1 || n_pix > 64 * 256 || ld_in < n_pix || ld_out < n_pix) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32), off = (uint32_t)offset;
  const int items = (n_pix + 255) / 256;
#define EBSD_NOISE_CASE(I)                                                                              \
  if (items <= I) {                                                                                     \
    hipLaunchKernelGGL(noise_kernel<I>, dim3(n), dim3(256), 0, st, clean, n_pix, ld_in, out, ld_out,    \
                       mean_counts, sigma, k0, k1, off);                                                \
    EBSD_LAUNCH_CHECK();                                                                                \
    return EBSD_OK;                                                                                     \
  }
  EBSD_NOISE_CASE(4)
  EBSD_NOISE_CASE(16)
  EBSD_NOISE_CASE(64)
#undef EBSD_NOISE_CASE
  return EBSD_ERR_SHAPE;
}
