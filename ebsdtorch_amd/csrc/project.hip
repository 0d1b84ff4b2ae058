// Dictionary projection and experimental-pattern preparation for gfx950.
//
// ebsd_project fuses, per (orientation, detector pixel):
//   quaternion rotation of the detector ray       (quaternions.py:186-213)
//   side-by-side Rosca-Lambert mapping            (sphere.py:145-185)
//   bilinear grid_sample, border, corners=False   (master_pattern.py:197-206)
// and, per orientation, the mean subtraction of indexing.py:139-143 plus the
// optional unit norm / low-precision cast used by the precision forms.
//
// One workgroup (256 threads up to 4,096 pixels, 1,024 above) owns one
// pattern. Its pixels live in registers (at most 16 per thread) between the
// sampling pass and the normalisation pass, so every output byte is written
// exactly once and coalesced. The master pattern
// (2 x S^2 floats, 2 MB at S = 501) stays resident in each XCD's L2.
//
// Numerics: rotation, Lambert arithmetic and the bilinear blend round step by
// step like the reference's float32 torch expressions (contraction is off
// inside them); the atan of the Lambert map is this library's own
// polynomial (<= 1.415 ulp) evaluated on a reciprocal-based ratio.

#include "common.h"
#include "lambert.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_fp8.h>

namespace {

using namespace ebsd_lambert;

constexpr float kFp8Max = 448.0f;  // largest finite OCP e4m3fn value
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T to_out(float v);
template <>
__device__ __forceinline__ float to_out<float>(float v) { return v; }
template <>
__device__ __forceinline__ __hip_bfloat16 to_out<__hip_bfloat16>(float v) { return __float2bfloat16(v); }
template <>
__device__ __forceinline__ __half to_out<__half>(float v) { return __float2half(v); }
template <>
__device__ __forceinline__ uint8_t to_out<uint8_t>(float v) {
  return (uint8_t)__hip_cvt_float_to_fp8(v, __HIP_SATFINITE, __HIP_E4M3);
}

template <typename T> struct is_fp8 { static constexpr bool value = false; };
template <> struct is_fp8<uint8_t> { static constexpr bool value = true; };

// Register slot layouts of a row: V = 1, slot i of thread t is column
// t + i BLOCK (one value per thread and load); V = 4, slot i is column
// 4 (t + (i / 4) BLOCK) + i % 4 (four consecutive columns per thread and
// 16-byte load or store). Both cover columns [0, ITEMS BLOCK).
template <int V, int BLOCK>
__device__ __forceinline__ int slot_col(int i) {
  if constexpr (V == 1)
    return (int)threadIdx.x + i * BLOCK;
  else
    return 4 * ((int)threadIdx.x + (i / 4) * BLOCK) + (i & 3);
}

// Four consecutive outputs of type T as one store.
template <typename T>
__device__ __forceinline__ void store4(T* p, float a, float b, float c, float d) {
  if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
  } else if constexpr (sizeof(T) == 2) {
    const T v[4] = {to_out<T>(a), to_out<T>(b), to_out<T>(c), to_out<T>(d)};
    *reinterpret_cast<uint2*>(p) = *reinterpret_cast<const uint2*>(v);
  } else {
    const T v[4] = {to_out<T>(a), to_out<T>(b), to_out<T>(c), to_out<T>(d)};
    *reinterpret_cast<unsigned*>(p) = *reinterpret_cast<const unsigned*>(v);
  }
}

// Normalise the ITEMS register values of one pattern (n_valid of them are
// real pixels) and store row `row` of `out`, zero-padding to ld.
template <int ITEMS, typename OutT, int BLOCK = 256, int V = 1>
__device__ __forceinline__ void finish_row(float (&vals)[ITEMS], int n_valid, int flags,
                                           OutT* __restrict__ out, int ld, float* __restrict__ out_scale,
                                           size_t row, float* red) {
  static_assert(V == 1 || (V == 4 && ITEMS % 4 == 0), "slot layout");
  const int tid = threadIdx.x;
  if (flags & EBSD_FLAG_ZERO_MEAN) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) s += vals[i];  // padding slots hold 0
    const float mean = block_sum<BLOCK>(s, red) / (float)n_valid;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int d = slot_col<V, BLOCK>(i);
      vals[i] = (d < n_valid) ? (vals[i] - mean) : 0.0f;
    }
  }
  if (flags & EBSD_FLAG_UNIT_NORM) {
    float ss = 0.0f;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) ss += vals[i] * vals[i];
    const float tot = block_sum<BLOCK>(ss, red);
    const float inv = tot > 0.0f ? 1.0f / sqrtf(tot) : 0.0f;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) vals[i] *= inv;
  }
  float mul = 1.0f;
  if constexpr (is_fp8<OutT>::value) {
    float am = 0.0f;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) am = fmaxf(am, fabsf(vals[i]));
    am = block_max<BLOCK>(am, red);
    // the smallest power of two 2^k with am / 2^k <= 448 (the row's largest
    // magnitude lands in (224, 448]): an E8M0 block scale for the FP8
    // matrix instructions, and an exact multiplier
    int k = 0;
    if (am > 0.0f) {
      int e;
      const float m = frexpf(am / kFp8Max, &e);  // am / 448 = m 2^e, m in [0.5, 1)
      k = min(max(m == 0.5f ? e - 1 : e, -126), 127);
    }
    mul = ldexpf(1.0f, -k);
    if (tid == 0 && out_scale) out_scale[row] = ldexpf(1.0f, k);
  } else {
    if (tid == 0 && out_scale) out_scale[row] = 1.0f;
  }
  OutT* orow = out + row * (size_t)ld;
  if constexpr (V == 1) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int d = tid + i * BLOCK;
      if (d < ld) orow[d] = to_out<OutT>(d < n_valid ? vals[i] * mul : 0.0f);
    }
  } else {
    // ld and n_valid are multiples of 4 here (see launch_prep)
#pragma unroll
    for (int i = 0; i < ITEMS; i += 4) {
      const int d = slot_col<V, BLOCK>(i);
      if (d < ld) {
        const bool in = d < n_valid;
        store4<OutT>(orow + d, in ? vals[i] * mul : 0.0f, in ? vals[i + 1] * mul : 0.0f,
                     in ? vals[i + 2] * mul : 0.0f, in ? vals[i + 3] * mul : 0.0f);
      }
    }
  }
  // Columns beyond ITEMS*BLOCK (only when ld exceeds it) are zero padding.
  for (int d = ITEMS * BLOCK + tid; d < ld; d += BLOCK) orow[d] = to_out<OutT>(0.0f);
}

// Sample-frame direction of pixel slot d of pattern n (rotation by the
// pattern's quaternion q; EXTRA: the per-pattern pixel shift o with the
// re-normalisation of the shifted ray, and/or the re-normalisation of the
// rotated one), then its side-by-side Lambert position.
template <bool EXTRA>
__device__ __forceinline__ void pixel_position(float4 q, float bx, float by, float bz, bool shift, float ox,
                                               float oy, float oz, bool renorm, float& gx, float& gy) {
  if constexpr (EXTRA) {
    if (shift) {
      bx += ox;
      by += oy;
      bz += oz;
      renormalise(bx, by, bz);
    }
  }
  float x, y, z;
  rotate_ray(q.x, q.y, q.z, q.w, bx, by, bz, x, y, z);
  if constexpr (EXTRA) {
    if (renorm) renormalise(x, y, z);
  }
  lambert_sbs(x, y, z, gx, gy);
}

// The ITEMS values of one pattern held by this thread (slots d = tid + i *
// BLOCK; slots past n_pix read 0). Per group of kProjGroup slots, the
// positions and texel addresses are computed first, then the gathers are
// issued, then the blends, with no branch between slots, so the gathers of
// consecutive slots overlap. Measured on the 995,277-orientation
// dictionaries (FP32 rows): groups of 1 / 2 / 4 / 8 / 16 slots take 16.7 /
// 15.0 / 15.2 / 16.6 / 21.7 ms at 60 x 60 and 85.9 / 70.3 / 74.7 / 100.5 /
// 134.8 ms at 120 x 120 (larger groups hold more registers and lower the
// occupancy); one branchy slot at a time took 14.9 / 121 ms.
constexpr int kProjGroup = 2;

template <int ITEMS, int BLOCK, bool QUADS, bool EXTRA>
__device__ __forceinline__ void pattern_values(float (&vals)[ITEMS], const float* __restrict__ mp,
                                               const float4* __restrict__ quads, int H, int W,
                                               const float (&rx)[ITEMS], const float (&ry)[ITEMS],
                                               const float (&rz)[ITEMS], int n_pix, float4 q, bool shift,
                                               float ox, float oy, float oz, bool renorm) {
  constexpr int G = ITEMS < kProjGroup ? ITEMS : kProjGroup;
#pragma unroll
  for (int g0 = 0; g0 < ITEMS; g0 += G) {
    // the last group may be short (ITEMS not a multiple of G)
    constexpr int kFull = G;
    const int n_in = ITEMS - g0 < kFull ? ITEMS - g0 : kFull;  // compile-time after unrolling
    float gx[G], gy[G];
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (j < n_in)
        pixel_position<EXTRA>(q, rx[g0 + j], ry[g0 + j], rz[g0 + j], shift, ox, oy, oz, renorm, gx[j], gy[j]);
    if constexpr (QUADS) {
      int t[G];
      float fw[G], fn[G];
#pragma unroll
      for (int j = 0; j < G; ++j)
        if (j < n_in) t[j] = quad_source(H, W, gx[j], gy[j], fw[j], fn[j]);
      float4 tap[G];
#pragma unroll
      for (int j = 0; j < G; ++j)
        if (j < n_in) tap[j] = quads[t[j]];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        if (j < n_in) {
          const int d = threadIdx.x + (g0 + j) * BLOCK;
          vals[g0 + j] = d < n_pix ? quad_blend(tap[j], fw[j], fn[j]) : 0.0f;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        if (j < n_in) {
          const int d = threadIdx.x + (g0 + j) * BLOCK;
          vals[g0 + j] = d < n_pix ? sample_border(mp, H, W, gx[j], gy[j]) : 0.0f;
        }
      }
    }
  }
}

// Persistent: workgroup g projects patterns g, g + gridDim.x, ... One
// workgroup of BLOCK threads per pattern, ITEMS (<= 16) pixel slots per
// thread; the thread's rays stay in registers across patterns. The optional
// per-pattern shift and re-normalisations (uniform flags) select a second
// copy of the pixel loop, so the plain path carries no branch per pixel.
template <int ITEMS, int BLOCK, typename OutT>
__global__ __launch_bounds__(BLOCK) void project_kernel(
    const float* __restrict__ mp, const float4* __restrict__ quads, int H, int W,
    const float* __restrict__ rays, int n_pix, const float* __restrict__ quats, int n_ori,
    OutT* __restrict__ out, int ld, float* __restrict__ out_scale, int flags,
    const float* __restrict__ offsets, const int* __restrict__ order) {
  __shared__ float red[BLOCK / 64];
  const bool renorm = (flags & EBSD_FLAG_RENORM_DIRS) != 0;
  const bool extra = renorm || offsets != nullptr;
  float rx[ITEMS], ry[ITEMS], rz[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const int d = threadIdx.x + i * BLOCK;
    const int dd = d < n_pix ? d : n_pix - 1;  // padding slots repeat a real ray (value dropped)
    rx[i] = __ldg(rays + 3 * dd + 0);
    ry[i] = __ldg(rays + 3 * dd + 1);
    rz[i] = __ldg(rays + 3 * dd + 2);
  }
  for (int k = blockIdx.x; k < n_ori; k += gridDim.x) {
    // processing slot k projects row `order[k]` (a permutation of the rows),
    // or row k: the order only decides which patterns run together
    const int n = order ? order[k] : k;
    const float4 q = *reinterpret_cast<const float4*>(quats + 4 * (size_t)n);
    // per-pattern shift of the (un-normalised) pixel positions: a moved
    // pattern centre translates every detector pixel by the same vector
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if (offsets) {
      ox = offsets[3 * (size_t)n];
      oy = offsets[3 * (size_t)n + 1];
      oz = offsets[3 * (size_t)n + 2];
    }
    float vals[ITEMS];
    if (quads) {
      if (extra)
        pattern_values<ITEMS, BLOCK, true, true>(vals, mp, quads, H, W, rx, ry, rz, n_pix, q, offsets != nullptr,
                                                 ox, oy, oz, renorm);
      else
        pattern_values<ITEMS, BLOCK, true, false>(vals, mp, quads, H, W, rx, ry, rz, n_pix, q, false, 0.0f, 0.0f,
                                                  0.0f, false);
    } else {
      if (extra)
        pattern_values<ITEMS, BLOCK, false, true>(vals, mp, quads, H, W, rx, ry, rz, n_pix, q, offsets != nullptr,
                                                  ox, oy, oz, renorm);
      else
        pattern_values<ITEMS, BLOCK, false, false>(vals, mp, quads, H, W, rx, ry, rz, n_pix, q, false, 0.0f,
                                                   0.0f, 0.0f, false);
    }
    finish_row<ITEMS, OutT, BLOCK>(vals, n_pix, flags, out, ld, out_scale, (size_t)n, red);
  }
}

// One workgroup of BLOCK threads per row, ITEMS values per thread (rows of
// up to 4,096 pixels: 256 threads; longer rows: 1,024 threads, so that a
// thread holds at most 16 values and occupancy stays up).
// V = 4: rows read and written 16 bytes per thread and access (unmasked rows
// whose width, leading dimensions and base addresses are multiples of 4
// values / 16 bytes; see launch_prep), V = 1 otherwise.
template <int ITEMS, int BLOCK, int V>
__device__ __forceinline__ void prep_load(float (&vals)[ITEMS], const float* __restrict__ row,
                                          const int* __restrict__ pix_idx, int n_pix) {
  if constexpr (V == 1) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int d = threadIdx.x + i * BLOCK;
      float v = 0.0f;
      if (d < n_pix) v = __builtin_nontemporal_load(&row[pix_idx ? pix_idx[d] : d]);  // read once
      vals[i] = v;
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; i += 4) {
      const int d = slot_col<V, BLOCK>(i);
      f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
      if (d < n_pix) v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(row + d));  // read once
      vals[i] = v[0];
      vals[i + 1] = v[1];
      vals[i + 2] = v[2];
      vals[i + 3] = v[3];
    }
  }
}

template <int ITEMS, int BLOCK, typename OutT, int V>
__global__ __launch_bounds__(BLOCK) void prep_kernel(
    const float* __restrict__ pats, int ld_in, const int* __restrict__ pix_idx, int n_pix,
    OutT* __restrict__ out, int ld, float* __restrict__ out_scale, int flags) {
  __shared__ float red[BLOCK / 64];
  const size_t n = blockIdx.x;
  float vals[ITEMS];
  prep_load<ITEMS, BLOCK, V>(vals, pats + n * (size_t)ld_in, pix_idx, n_pix);
  finish_row<ITEMS, OutT, BLOCK, V>(vals, n_pix, flags, out, ld, out_scale, n, red);
}

// prep_kernel with a second output: the finished FP32 row and the same
// values converted to Out2 (BF16 / FP16), from one read of the patterns.
template <int ITEMS, int BLOCK, typename Out2, int V>
__global__ __launch_bounds__(BLOCK) void prep_pair_kernel(
    const float* __restrict__ pats, int ld_in, const int* __restrict__ pix_idx, int n_pix,
    float* __restrict__ out, int ld, Out2* __restrict__ out2, int ld2, int flags) {
  __shared__ float red[BLOCK / 64];
  const size_t n = blockIdx.x;
  float vals[ITEMS];
  prep_load<ITEMS, BLOCK, V>(vals, pats + n * (size_t)ld_in, pix_idx, n_pix);
  // finish_row leaves the finished values in vals (FP32 output: no scale)
  finish_row<ITEMS, float, BLOCK, V>(vals, n_pix, flags, out, ld, nullptr, n, red);
  Out2* orow = out2 + n * (size_t)ld2;
  if constexpr (V == 1) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int d = threadIdx.x + i * BLOCK;
      if (d < ld2) orow[d] = to_out<Out2>(d < n_pix ? vals[i] : 0.0f);
    }
  } else if ((ld2 & 3) == 0 && (reinterpret_cast<uintptr_t>(orow) & (4 * sizeof(Out2) - 1)) == 0) {
#pragma unroll
    for (int i = 0; i < ITEMS; i += 4) {
      const int d = slot_col<V, BLOCK>(i);
      if (d < ld2) {
        const bool in = d < n_pix;
        store4<Out2>(orow + d, in ? vals[i] : 0.0f, in ? vals[i + 1] : 0.0f, in ? vals[i + 2] : 0.0f,
                     in ? vals[i + 3] : 0.0f);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int d = slot_col<V, BLOCK>(i);
      if (d < ld2) orow[d] = to_out<Out2>(d < n_pix ? vals[i] : 0.0f);
    }
  }
  for (int d = ITEMS * BLOCK + threadIdx.x; d < ld2; d += BLOCK) orow[d] = to_out<Out2>(0.0f);
}

// Sample the master pattern at arbitrary sphere directions (one per thread).
__global__ __launch_bounds__(256) void sample_dirs_kernel(const float* __restrict__ mp, int H, int W,
                                                          const float* __restrict__ dirs, int64_t n,
                                                          float* __restrict__ out, int renorm) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float x = dirs[3 * i], y = dirs[3 * i + 1], z = dirs[3 * i + 2];
  if (renorm) renormalise(x, y, z);
  float gx, gy;
  lambert_sbs(x, y, z, gx, gy);
  out[i] = sample_border(mp, H, W, gx, gy);
}

// grid_sample's other modes (torch.nn.functional.grid_sample, as
// MasterPattern.interpolate passes them through, master_pattern.py:197-206),
// with the float32 operation sequence of torch's CPU kernel: source index
// (g + 1) * size / 2 - 0.5 (align_corners=False) or (g + 1) * (size - 1) / 2;
// border clips the coordinate to [0, size - 1]; reflection folds it about
// [-0.5, size - 0.5] (or [0, size - 1]) and then clips; zeros reads taps
// outside the image as 0. Bilinear and nearest pad the coordinate, bicubic
// pads each of its 16 integer taps (Keys' cubic convolution, A = -0.75).
template <int PAD, bool ALIGN>
struct SampleAxis {
  int size;
  float max_val, scale, low, twice_span;
  __device__ explicit SampleAxis(int n)
      : size(n),
        max_val((float)(n - 1)),
        scale(ALIGN ? (float)(n - 1) / 2.0f : (float)n / 2.0f),
        low(ALIGN ? 0.0f : -0.5f),
        twice_span(ALIGN ? (float)(n - 1) * 2.0f : (float)n * 2.0f) {}
  __device__ float unnormalize(float g) const {
#pragma clang fp contract(off)
    return ALIGN ? (g + 1.0f) * scale : (g + 1.0f) * scale - 0.5f;
  }
  __device__ float clip(float v) const { return fminf(max_val, fmaxf(0.0f, v)); }
  __device__ float reflect(float v) const {
#pragma clang fp contract(off)
    if (twice_span == 0.0f) return 0.0f;
    const float a = fabsf(v - low);
    const float flips = truncf(a / twice_span);
    const float extra = a - flips * twice_span;
    return fminf(extra, twice_span - extra) + low;
  }
  // padding applied to a continuous (or integer-valued) coordinate
  __device__ float pad(float v) const {
    if constexpr (PAD == EBSD_PAD_BORDER) return clip(v);
    else if constexpr (PAD == EBSD_PAD_REFLECTION) return clip(reflect(v));
    else return v;
  }
  __device__ bool inside(int i) const { return i > -1 && i < size; }
  // pad() with the derivative of its result by its argument, as torch's
  // grid_sample backward takes it: clipping gives 0 on and beyond the
  // borders, reflection -1 where the coordinate was folded an odd number of
  // times
  __device__ float pad_grad(float v, float& d) const {
#pragma clang fp contract(off)
    d = 1.0f;
    if constexpr (PAD == EBSD_PAD_REFLECTION) {
      if (twice_span == 0.0f) {
        d = 0.0f;
        return 0.0f;
      }
      const float rel = v - low;
      const float a = fabsf(rel);
      const float flips = truncf(a / twice_span);
      const float extra = a - flips * twice_span;
      const float back = twice_span - extra;
      const bool more = extra > back;
      d = (more != (rel < 0.0f)) ? -1.0f : 1.0f;
      v = (more ? back : extra) + low;
    }
    if constexpr (PAD != EBSD_PAD_ZEROS) {
      if (!(v > 0.0f && v < max_val)) d = 0.0f;
      v = clip(v);
    }
    return v;
  }
};

template <int PAD, bool ALIGN>
__device__ __forceinline__ float sample_tap(const float* __restrict__ img, const SampleAxis<PAD, ALIGN>& ax,
                                            const SampleAxis<PAD, ALIGN>& ay, int xi, int yi) {
  if (!ax.inside(xi) || !ay.inside(yi)) return 0.0f;
  return __ldg(img + (yi * ax.size + xi));
}

__device__ __forceinline__ void cubic_coefficients(float t, float (&c)[4]) {
#pragma clang fp contract(off)
  const float A = -0.75f;
  float x = t + 1.0f;
  c[0] = ((A * x - 5.0f * A) * x + 8.0f * A) * x - 4.0f * A;
  x = t;
  c[1] = ((A + 2.0f) * x - (A + 3.0f)) * x * x + 1.0f;
  x = 1.0f - t;
  c[2] = ((A + 2.0f) * x - (A + 3.0f)) * x * x + 1.0f;
  x = 2.0f - t;
  c[3] = ((A * x - 5.0f * A) * x + 8.0f * A) * x - 4.0f * A;
}

template <int MODE, int PAD, bool ALIGN>
__device__ float sample_mode(const float* __restrict__ img, int H, int W, float gx, float gy) {
#pragma clang fp contract(off)
  const SampleAxis<PAD, ALIGN> ax(W), ay(H);
  if constexpr (MODE == EBSD_SAMPLE_BICUBIC) {
    const float x = ax.unnormalize(gx), y = ay.unnormalize(gy);
    const float fx = floorf(x), fy = floorf(y);
    float cx[4], cy[4];
    cubic_coefficients(x - fx, cx);
    cubic_coefficients(y - fy, cy);
    int xi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) xi[i] = (int)ax.pad(fx + (float)(i - 1));
    float out = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int yi = (int)ay.pad(fy + (float)(j - 1));
      const float row = ((cx[0] * sample_tap(img, ax, ay, xi[0], yi) + cx[1] * sample_tap(img, ax, ay, xi[1], yi)) +
                         cx[2] * sample_tap(img, ax, ay, xi[2], yi)) +
                        cx[3] * sample_tap(img, ax, ay, xi[3], yi);
      out = j == 0 ? cy[0] * row : out + cy[j] * row;
    }
    return out;
  } else {
    const float x = ax.pad(ax.unnormalize(gx)), y = ay.pad(ay.unnormalize(gy));
    if constexpr (MODE == EBSD_SAMPLE_NEAREST) {
      return sample_tap(img, ax, ay, (int)rintf(x), (int)rintf(y));
    } else {
      const float x_w = floorf(x), y_n = floorf(y);
      const float w = x - x_w, e = 1.0f - w, n = y - y_n, s = 1.0f - n;
      const int ix = (int)x_w, iy = (int)y_n;
      return ((sample_tap(img, ax, ay, ix, iy) * (s * e) + sample_tap(img, ax, ay, ix + 1, iy) * (s * w)) +
              sample_tap(img, ax, ay, ix, iy + 1) * (n * e)) +
             sample_tap(img, ax, ay, ix + 1, iy + 1) * (n * w);
    }
  }
}

// d/dt of the four cubic-convolution weights of cubic_coefficients(t)
__device__ __forceinline__ void cubic_coefficients_grad(float t, float (&c)[4]) {
  const float A = -0.75f;
  auto d_near = [&](float x) { return (3.0f * (A + 2.0f) * x - 2.0f * (A + 3.0f)) * x; };  // |x| <= 1
  auto d_far = [&](float x) { return (3.0f * A * x - 10.0f * A) * x + 8.0f * A; };         // 1 < |x| < 2
  c[0] = d_far(t + 1.0f);
  c[1] = d_near(t);
  c[2] = -d_near(1.0f - t);
  c[3] = -d_far(2.0f - t);
}

// Derivatives of sample_mode's value by gx and gy (0 for nearest).
template <int MODE, int PAD, bool ALIGN>
__device__ void sample_mode_grad(const float* __restrict__ img, int H, int W, float gx, float gy, float& dgx,
                                 float& dgy) {
  const SampleAxis<PAD, ALIGN> ax(W), ay(H);
  dgx = 0.0f;
  dgy = 0.0f;
  if constexpr (MODE == EBSD_SAMPLE_BICUBIC) {
    const float x = ax.unnormalize(gx), y = ay.unnormalize(gy);
    const float fx = floorf(x), fy = floorf(y);
    float cx[4], cy[4], dx[4], dy[4];
    cubic_coefficients(x - fx, cx);
    cubic_coefficients(y - fy, cy);
    cubic_coefficients_grad(x - fx, dx);
    cubic_coefficients_grad(y - fy, dy);
    int xi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) xi[i] = (int)ax.pad(fx + (float)(i - 1));
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int yi = (int)ay.pad(fy + (float)(j - 1));
      float row = 0.0f, drow = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float v = sample_tap(img, ax, ay, xi[i], yi);
        row += cx[i] * v;
        drow += dx[i] * v;
      }
      dgx += cy[j] * drow;
      dgy += dy[j] * row;
    }
    dgx *= ax.scale;
    dgy *= ay.scale;
  } else if constexpr (MODE == EBSD_SAMPLE_BILINEAR) {
    float mx, my;
    const float x = ax.pad_grad(ax.unnormalize(gx), mx), y = ay.pad_grad(ay.unnormalize(gy), my);
    const float x_w = floorf(x), y_n = floorf(y);
    const float w = x - x_w, e = 1.0f - w, n = y - y_n, s = 1.0f - n;
    const int ix = (int)x_w, iy = (int)y_n;
    const float v_nw = sample_tap(img, ax, ay, ix, iy), v_ne = sample_tap(img, ax, ay, ix + 1, iy);
    const float v_sw = sample_tap(img, ax, ay, ix, iy + 1), v_se = sample_tap(img, ax, ay, ix + 1, iy + 1);
    dgx = (s * (v_ne - v_nw) + n * (v_se - v_sw)) * (mx * ax.scale);
    dgy = (e * (v_sw - v_nw) + w * (v_se - v_ne)) * (my * ay.scale);
  }
}

// Value and gradient by the direction (normalize_coords=False semantics: the
// direction is used as given) of the master-pattern lookup.
template <int MODE, int PAD, bool ALIGN>
__global__ __launch_bounds__(256) void sample_dirs_grad_kernel(const float* __restrict__ mp, int H, int W,
                                                               const float* __restrict__ dirs, int64_t n,
                                                               float* __restrict__ out,
                                                               float* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float x = dirs[3 * i], y = dirs[3 * i + 1], z = dirs[3 * i + 2];
  float gx, gy, jac[6];
  lambert_sbs_jac(x, y, z, gx, gy, jac);
  out[i] = sample_mode<MODE, PAD, ALIGN>(mp, H, W, gx, gy);
  float dgx, dgy;
  sample_mode_grad<MODE, PAD, ALIGN>(mp, H, W, gx, gy, dgx, dgy);
#pragma unroll
  for (int c = 0; c < 3; ++c) grad[3 * i + c] = dgx * jac[c] + dgy * jac[3 + c];
}

template <int MODE, int PAD, bool ALIGN>
__global__ __launch_bounds__(256) void sample_dirs_mode_kernel(const float* __restrict__ mp, int H, int W,
                                                               const float* __restrict__ dirs, int64_t n,
                                                               float* __restrict__ out, int renorm) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float x = dirs[3 * i], y = dirs[3 * i + 1], z = dirs[3 * i + 2];
  if (renorm) renormalise(x, y, z);
  float gx, gy;
  lambert_sbs(x, y, z, gx, gy);
  out[i] = sample_mode<MODE, PAD, ALIGN>(mp, H, W, gx, gy);
}

template <typename OutT>
int launch_project(const float* mp, int H, int W, const float* rays, int n_pix, const float* quats,
                   int n_ori, void* out, int ld, float* out_scale, int flags, hipStream_t st,
                   const float* offsets = nullptr, const float4* quads = nullptr, const int* order = nullptr) {
  OutT* o = static_cast<OutT*>(out);
  // the first (slots per thread, threads) pair that covers n_pix. Pairs that
  // fit the detector more tightly (60 x 60 on 19 x 192 slots, 80 x 80 on
  // 20 x 320, 120 x 120 on 15 x 960) spend less VALU work on padding slots
  // but hold more registers per thread and ran slower (15.5 / 35.5 / 78.8 ms
  // against 15.0 / 30.6 / 70.3 ms for the FP32 rows of the 1M dictionaries)
  // persistent: about 4,096 x 256 threads (16 workgroups of 256 per CU on
  // 256 CUs); each workgroup walks n_ori / grid patterns
#define EBSD_PROJ_CASE(I, B)                                                                            \
  if (n_pix <= (I) * (B)) {                                                                             \
    const int g = (4096 * 256) / (B);                                                                   \
    hipLaunchKernelGGL((project_kernel<I, B, OutT>), dim3(n_ori < g ? n_ori : g), dim3(B), 0, st, mp,   \
                       quads, H, W, rays, n_pix, quats, n_ori, o, ld, out_scale, flags, offsets, order); \
    EBSD_LAUNCH_CHECK();                                                                                \
    return EBSD_OK;                                                                                     \
  }
  EBSD_PROJ_CASE(4, 256)
  EBSD_PROJ_CASE(8, 256)
  EBSD_PROJ_CASE(16, 256)
  EBSD_PROJ_CASE(8, 1024)
  EBSD_PROJ_CASE(16, 1024)
#undef EBSD_PROJ_CASE
  return EBSD_ERR_SHAPE;
}

// Rows take the 16-byte layout (V = 4) when they are read whole (no pixel
// mask) and their width, row stride and base address are multiples of 4
// values / 16 bytes. The layout only changes which thread holds which
// values (and so the order of the row sums): prep_kernel and
// prep_pair_kernel decide it alike, so their FP32 rows stay identical.
inline bool prep_vec4(const float* pats, int ld_in, const int* pix_idx, int n_pix) {
  return pix_idx == nullptr && n_pix % 4 == 0 && ld_in % 4 == 0 && (reinterpret_cast<uintptr_t>(pats) & 15) == 0;
}

template <typename Out2>
int launch_prep_pair(const float* pats, int n, int ld_in, const int* pix_idx, int n_pix, float* out, int ld,
                     void* out2, int ld2, int flags, hipStream_t st) {
  Out2* o2 = static_cast<Out2*>(out2);
  // the FP32 rows' layout is decided as launch_prep decides it for FP32
  // output; the second output is stored 4 at a time where aligned
  const bool v4 = prep_vec4(pats, ld_in, pix_idx, n_pix) && ld % 4 == 0 &&
                  (reinterpret_cast<uintptr_t>(out) & 15) == 0;
#define EBSD_PREP_PAIR_CASE(I, B)                                                                            \
  if (n_pix <= (I) * (B)) {                                                                                  \
    if (v4)                                                                                                  \
      hipLaunchKernelGGL((prep_pair_kernel<I, B, Out2, 4>), dim3(n), dim3(B), 0, st, pats, ld_in, pix_idx,    \
                         n_pix, out, ld, o2, ld2, flags);                                                    \
    else                                                                                                     \
      hipLaunchKernelGGL((prep_pair_kernel<I, B, Out2, 1>), dim3(n), dim3(B), 0, st, pats, ld_in, pix_idx,    \
                         n_pix, out, ld, o2, ld2, flags);                                                    \
    EBSD_LAUNCH_CHECK();                                                                                     \
    return EBSD_OK;                                                                                          \
  }
  EBSD_PREP_PAIR_CASE(4, 256)
  EBSD_PREP_PAIR_CASE(8, 256)
  EBSD_PREP_PAIR_CASE(16, 256)
  EBSD_PREP_PAIR_CASE(8, 1024)
  EBSD_PREP_PAIR_CASE(16, 1024)
#undef EBSD_PREP_PAIR_CASE
  return EBSD_ERR_SHAPE;
}

template <typename OutT>
int launch_prep(const float* pats, int n, int ld_in, const int* pix_idx, int n_pix, void* out,
                int ld, float* out_scale, int flags, hipStream_t st) {
  OutT* o = static_cast<OutT*>(out);
  const bool v4 = prep_vec4(pats, ld_in, pix_idx, n_pix) && ld % 4 == 0 &&
                  (reinterpret_cast<uintptr_t>(out) & 15) == 0;
#define EBSD_PREP_CASE(I, B)                                                                                 \
  if (n_pix <= (I) * (B)) {                                                                                  \
    if (v4)                                                                                                  \
      hipLaunchKernelGGL((prep_kernel<I, B, OutT, 4>), dim3(n), dim3(B), 0, st, pats, ld_in, pix_idx, n_pix,   \
                         o, ld, out_scale, flags);                                                           \
    else                                                                                                     \
      hipLaunchKernelGGL((prep_kernel<I, B, OutT, 1>), dim3(n), dim3(B), 0, st, pats, ld_in, pix_idx, n_pix,   \
                         o, ld, out_scale, flags);                                                           \
    EBSD_LAUNCH_CHECK();                                                                                     \
    return EBSD_OK;                                                                                          \
  }
  EBSD_PREP_CASE(4, 256)
  EBSD_PREP_CASE(8, 256)
  EBSD_PREP_CASE(16, 256)
  EBSD_PREP_CASE(8, 1024)
  EBSD_PREP_CASE(16, 1024)
#undef EBSD_PREP_CASE
  return EBSD_ERR_SHAPE;
}

}  // namespace

extern "C" int ebsd_pad_multiple(int dtype) {
  switch (dtype) {
    case EBSD_DTYPE_F32: return 32;
    case EBSD_DTYPE_BF16:
    case EBSD_DTYPE_F16: return 64;
    case EBSD_DTYPE_FP8_E4M3: return 128;
    default: return EBSD_ERR_DTYPE;
  }
}

static int check_ld(int dtype, int n_pix, int ld) {
  const int pm = ebsd_pad_multiple(dtype);
  if (pm < 0) return EBSD_ERR_DTYPE;
  if (n_pix <= 0 || n_pix > 64 * 256) return EBSD_ERR_SHAPE;
  if (ld < n_pix || ld % pm != 0) return EBSD_ERR_SHAPE;
  return EBSD_OK;
}

extern "C" int ebsd_project(const float* mp, int mp_rows, int mp_cols, const float* rays,
                            int n_pix, const float* quats, int n_ori, void* out, int out_dtype,
                            int ld_out, float* out_scale, int flags, void* stream) {
  if (!mp || !rays || !quats || !out) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || n_ori < 0) return EBSD_ERR_SHAPE;
  if ((int64_t)mp_rows * mp_cols > (int64_t)INT_MAX) return EBSD_ERR_SHAPE;
  const int chk = check_ld(out_dtype, n_pix, ld_out);
  if (chk) return chk;
  if (out_dtype == EBSD_DTYPE_FP8_E4M3 && !out_scale) return EBSD_ERR_NULL;
  if ((reinterpret_cast<uintptr_t>(quats) & 15) != 0) return EBSD_ERR_ALIGN;
  if (n_ori == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (out_dtype) {
    case EBSD_DTYPE_F32:
      return launch_project<float>(mp, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale, flags, st);
    case EBSD_DTYPE_BF16:
      return launch_project<__hip_bfloat16>(mp, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale, flags, st);
    case EBSD_DTYPE_F16:
      return launch_project<__half>(mp, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale, flags, st);
    case EBSD_DTYPE_FP8_E4M3:
      return launch_project<uint8_t>(mp, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale, flags, st);
  }
  return EBSD_ERR_DTYPE;
}

extern "C" int ebsd_project_quads(const float* quads, int mp_rows, int mp_cols, const float* rays, int n_pix,
                                  const float* quats, int n_ori, void* out, int out_dtype, int ld_out,
                                  float* out_scale, int flags, const int* order, void* stream) {
  if (!quads || !rays || !quats || !out) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || n_ori < 0) return EBSD_ERR_SHAPE;
  if ((int64_t)mp_rows * mp_cols > (int64_t)INT_MAX) return EBSD_ERR_SHAPE;
  const int chk = check_ld(out_dtype, n_pix, ld_out);
  if (chk) return chk;
  if (out_dtype == EBSD_DTYPE_FP8_E4M3 && !out_scale) return EBSD_ERR_NULL;
  if (((reinterpret_cast<uintptr_t>(quats) | reinterpret_cast<uintptr_t>(quads)) & 15) != 0) return EBSD_ERR_ALIGN;
  if (n_ori == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float4* q4 = reinterpret_cast<const float4*>(quads);
  switch (out_dtype) {
    case EBSD_DTYPE_F32:
      return launch_project<float>(nullptr, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale,
                                   flags, st, nullptr, q4, order);
    case EBSD_DTYPE_BF16:
      return launch_project<__hip_bfloat16>(nullptr, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out,
                                            out_scale, flags, st, nullptr, q4, order);
    case EBSD_DTYPE_F16:
      return launch_project<__half>(nullptr, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale,
                                    flags, st, nullptr, q4, order);
    case EBSD_DTYPE_FP8_E4M3:
      return launch_project<uint8_t>(nullptr, mp_rows, mp_cols, rays, n_pix, quats, n_ori, out, ld_out, out_scale,
                                     flags, st, nullptr, q4, order);
  }
  return EBSD_ERR_DTYPE;
}

namespace {

// Interleave the low 7 bits of x with the low 6 bits of y (y bits above x's).
__device__ __forceinline__ int morton_7x6(int x, int y) {
  int k = 0;
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    k |= ((x >> b) & 1) << (2 * b);
    if (b < 6) k |= ((y >> b) & 1) << (2 * b + 1);
  }
  return k;
}

__global__ __launch_bounds__(256) void view_keys_kernel(const float* __restrict__ quats, int n, float dx, float dy,
                                                        float dz, int* __restrict__ keys) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 q = *reinterpret_cast<const float4*>(quats + 4 * (size_t)i);
  float x, y, z, gx, gy;
  rotate_ray(q.x, q.y, q.z, q.w, dx, dy, dz, x, y, z);
  lambert_sbs(x, y, z, gx, gy);
  const int cx = min(max((int)((gx + 1.0f) * 64.0f), 0), 127);  // 128 columns over the 2S-wide image
  const int cy = min(max((int)((gy + 1.0f) * 32.0f), 0), 63);   // 64 rows
  keys[i] = morton_7x6(cx, cy);
}

}  // namespace

extern "C" int ebsd_view_keys(const float* quats, int n_ori, float dir_x, float dir_y, float dir_z, int* keys,
                              void* stream) {
  if (!quats || !keys) return EBSD_ERR_NULL;
  if (n_ori < 0) return EBSD_ERR_SHAPE;
  if ((reinterpret_cast<uintptr_t>(quats) & 15) != 0) return EBSD_ERR_ALIGN;
  if (n_ori == 0) return EBSD_OK;
  hipLaunchKernelGGL(view_keys_kernel, dim3((n_ori + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                     quats, n_ori, dir_x, dir_y, dir_z, keys);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_project_shifted(const float* mp, int mp_rows, int mp_cols, const float* coords,
                                    int n_pix, const float* quats, const float* offsets, int n_ori, void* out,
                                    int out_dtype, int ld_out, float* out_scale, int flags, void* stream) {
  if (!mp || !coords || !quats || !offsets || !out) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || n_ori < 0) return EBSD_ERR_SHAPE;
  if ((int64_t)mp_rows * mp_cols > (int64_t)INT_MAX) return EBSD_ERR_SHAPE;
  const int chk = check_ld(out_dtype, n_pix, ld_out);
  if (chk) return chk;
  if (out_dtype == EBSD_DTYPE_FP8_E4M3 && !out_scale) return EBSD_ERR_NULL;
  if ((reinterpret_cast<uintptr_t>(quats) & 15) != 0) return EBSD_ERR_ALIGN;
  if (n_ori == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (out_dtype) {
    case EBSD_DTYPE_F32:
      return launch_project<float>(mp, mp_rows, mp_cols, coords, n_pix, quats, n_ori, out, ld_out, out_scale,
                                   flags, st, offsets);
    case EBSD_DTYPE_BF16:
      return launch_project<__hip_bfloat16>(mp, mp_rows, mp_cols, coords, n_pix, quats, n_ori, out, ld_out,
                                            out_scale, flags, st, offsets);
    case EBSD_DTYPE_F16:
      return launch_project<__half>(mp, mp_rows, mp_cols, coords, n_pix, quats, n_ori, out, ld_out, out_scale,
                                    flags, st, offsets);
    case EBSD_DTYPE_FP8_E4M3:
      return launch_project<uint8_t>(mp, mp_rows, mp_cols, coords, n_pix, quats, n_ori, out, ld_out, out_scale,
                                     flags, st, offsets);
  }
  return EBSD_ERR_DTYPE;
}

extern "C" int ebsd_sample_dirs(const float* mp, int mp_rows, int mp_cols, const float* dirs,
                                long long n, float* out, int renorm, void* stream) {
  if (!mp || !dirs || !out) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || n < 0) return EBSD_ERR_SHAPE;
  if ((int64_t)mp_rows * mp_cols > (int64_t)INT_MAX) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return EBSD_ERR_SHAPE;
  hipLaunchKernelGGL(sample_dirs_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), mp, mp_rows, mp_cols, dirs, (int64_t)n, out, renorm);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

// Launch of the value kernel (grad == nullptr: renorm applies) or the
// value + gradient kernel for one (mode, padding, alignment).
template <int MODE, int PAD, bool ALIGN>
static void launch_sample_one(dim3 grid, hipStream_t st, const float* mp, int H, int W, const float* dirs,
                              int64_t n, float* out, int renorm, float* grad) {
  if (grad)
    hipLaunchKernelGGL((sample_dirs_grad_kernel<MODE, PAD, ALIGN>), grid, dim3(256), 0, st, mp, H, W, dirs, n, out,
                       grad);
  else
    hipLaunchKernelGGL((sample_dirs_mode_kernel<MODE, PAD, ALIGN>), grid, dim3(256), 0, st, mp, H, W, dirs, n, out,
                       renorm);
}

template <int MODE, int PAD>
static void launch_sample_align(bool align, dim3 grid, hipStream_t st, const float* mp, int H, int W,
                                const float* dirs, int64_t n, float* out, int renorm, float* grad) {
  if (align)
    launch_sample_one<MODE, PAD, true>(grid, st, mp, H, W, dirs, n, out, renorm, grad);
  else
    launch_sample_one<MODE, PAD, false>(grid, st, mp, H, W, dirs, n, out, renorm, grad);
}

template <int MODE>
static void launch_sample_pad(int padding, bool align, dim3 grid, hipStream_t st, const float* mp, int H, int W,
                              const float* dirs, int64_t n, float* out, int renorm, float* grad) {
  switch (padding) {
    case EBSD_PAD_ZEROS:
      launch_sample_align<MODE, EBSD_PAD_ZEROS>(align, grid, st, mp, H, W, dirs, n, out, renorm, grad);
      break;
    case EBSD_PAD_BORDER:
      launch_sample_align<MODE, EBSD_PAD_BORDER>(align, grid, st, mp, H, W, dirs, n, out, renorm, grad);
      break;
    default:
      launch_sample_align<MODE, EBSD_PAD_REFLECTION>(align, grid, st, mp, H, W, dirs, n, out, renorm, grad);
      break;
  }
}

static int sample_dirs_any(const float* mp, int mp_rows, int mp_cols, const float* dirs, long long n, float* out,
                           int renorm, int mode, int padding, int align_corners, float* grad, void* stream) {
  if (!mp || !dirs || !out) return EBSD_ERR_NULL;
  if (mp_rows < 1 || mp_cols < 1 || n < 0) return EBSD_ERR_SHAPE;
  if ((int64_t)mp_rows * mp_cols > (int64_t)INT_MAX) return EBSD_ERR_SHAPE;
  if (mode < EBSD_SAMPLE_BILINEAR || mode > EBSD_SAMPLE_BICUBIC) return EBSD_ERR_SHAPE;
  if (padding < EBSD_PAD_ZEROS || padding > EBSD_PAD_REFLECTION) return EBSD_ERR_SHAPE;
  if (n == 0) return EBSD_OK;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return EBSD_ERR_SHAPE;
  const dim3 grid((unsigned)blocks);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool align = align_corners != 0;
  const int H = mp_rows, W = mp_cols;
  switch (mode) {
    case EBSD_SAMPLE_BILINEAR:
      launch_sample_pad<EBSD_SAMPLE_BILINEAR>(padding, align, grid, st, mp, H, W, dirs, n, out, renorm, grad);
      break;
    case EBSD_SAMPLE_NEAREST:
      launch_sample_pad<EBSD_SAMPLE_NEAREST>(padding, align, grid, st, mp, H, W, dirs, n, out, renorm, grad);
      break;
    default:
      launch_sample_pad<EBSD_SAMPLE_BICUBIC>(padding, align, grid, st, mp, H, W, dirs, n, out, renorm, grad);
      break;
  }
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_sample_dirs_mode(const float* mp, int mp_rows, int mp_cols, const float* dirs, long long n,
                                     float* out, int renorm, int mode, int padding, int align_corners,
                                     void* stream) {
  return sample_dirs_any(mp, mp_rows, mp_cols, dirs, n, out, renorm, mode, padding, align_corners, nullptr, stream);
}

extern "C" int ebsd_sample_dirs_grad(const float* mp, int mp_rows, int mp_cols, const float* dirs, long long n,
                                     float* out, float* grad, int mode, int padding, int align_corners,
                                     void* stream) {
  if (!grad) return EBSD_ERR_NULL;
  return sample_dirs_any(mp, mp_rows, mp_cols, dirs, n, out, 0, mode, padding, align_corners, grad, stream);
}

extern "C" int ebsd_prep_patterns(const float* pats, int n, int ld_in, const int* pix_idx,
                                  int n_pix, void* out, int out_dtype, int ld_out,
                                  float* out_scale, int flags, void* stream) {
  if (!pats || !out) return EBSD_ERR_NULL;
  if (n < 0 || ld_in < 1) return EBSD_ERR_SHAPE;
  if (!pix_idx && n_pix > ld_in) return EBSD_ERR_SHAPE;
  const int chk = check_ld(out_dtype, n_pix, ld_out);
  if (chk) return chk;
  if (out_dtype == EBSD_DTYPE_FP8_E4M3 && !out_scale) return EBSD_ERR_NULL;
  if (n == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (out_dtype) {
    case EBSD_DTYPE_F32:
      return launch_prep<float>(pats, n, ld_in, pix_idx, n_pix, out, ld_out, out_scale, flags, st);
    case EBSD_DTYPE_BF16:
      return launch_prep<__hip_bfloat16>(pats, n, ld_in, pix_idx, n_pix, out, ld_out, out_scale, flags, st);
    case EBSD_DTYPE_F16:
      return launch_prep<__half>(pats, n, ld_in, pix_idx, n_pix, out, ld_out, out_scale, flags, st);
    case EBSD_DTYPE_FP8_E4M3:
      return launch_prep<uint8_t>(pats, n, ld_in, pix_idx, n_pix, out, ld_out, out_scale, flags, st);
  }
  return EBSD_ERR_DTYPE;
}

extern "C" int ebsd_prep_patterns_pair(const float* pats, int n, int ld_in, const int* pix_idx, int n_pix,
                                       float* out_f32, int ld_f32, void* out2, int out2_dtype, int ld2, int flags,
                                       void* stream) {
  if (!pats || !out_f32 || !out2) return EBSD_ERR_NULL;
  if (n < 0 || ld_in < 1) return EBSD_ERR_SHAPE;
  if (!pix_idx && n_pix > ld_in) return EBSD_ERR_SHAPE;
  if (out2_dtype != EBSD_DTYPE_BF16 && out2_dtype != EBSD_DTYPE_F16) return EBSD_ERR_DTYPE;
  int chk = check_ld(EBSD_DTYPE_F32, n_pix, ld_f32);
  if (chk) return chk;
  chk = check_ld(out2_dtype, n_pix, ld2);
  if (chk) return chk;
  if (n == 0) return EBSD_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (out2_dtype == EBSD_DTYPE_BF16)
    return launch_prep_pair<__hip_bfloat16>(pats, n, ld_in, pix_idx, n_pix, out_f32, ld_f32, out2, ld2, flags, st);
  return launch_prep_pair<__half>(pats, n, ld_in, pix_idx, n_pix, out_f32, ld_f32, out2, ld2, flags, st);
}
