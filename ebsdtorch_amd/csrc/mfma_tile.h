// Shared MFMA tile machinery of the similarity and GEMM kernels (gfx950).
//
// A 256-thread workgroup computes a 128 (A rows) x 128 (B rows) block of
// A * B^T, with A and B row-major and K contiguous; wave w owns B rows
// [32 w, 32 w + 32) and all 128 A rows, i.e. four 32x32 MFMA accumulator
// tiles. K streams through double-buffered LDS in 128-byte stages.
#pragma once

#include "common.h"

namespace ebsd_tile {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

constexpr int TILE = 128;          // rows of A and of B per block
constexpr int STAGE_BYTES = 128;   // bytes of K per row per LDS stage
constexpr int LDS_ROW = 144;       // padded LDS row stride (bytes): conflict-free b128 reads
constexpr int LDS_TILE = TILE * LDS_ROW;
constexpr int LDS_BYTES = 4 * LDS_TILE;  // 2 stages x (dictionary tile + query tile)

// E8M0 exponent of 1.0: unit block scales for the MX-scaled FP8 MFMA.
constexpr int kUnitScale = 127;

// Per-lane position of accumulator element r of a 32x32 MFMA tile:
// column = lane % 32, row = 8 * (r / 4) + 4 * (lane / 32) + r % 4.
__device__ __forceinline__ int mfma32_row(int r, int lane) {
  return ((r >> 2) << 3) + ((lane >> 5) << 2) + (r & 3);
}

// Byte offset of lane half h = lane / 32 inside a row's K stage. Each MFMA
// of the low-precision types covers one contiguous block of K: FP8 (32x32x64)
// bytes [64 s, 64 s + 64), lane half h the 32 bytes at 64 s + 32 h; BF16/FP16
// (32x32x16) bytes [32 s, 32 s + 32), lane half h the 16 bytes at 32 s + 16 h.
// Every similarity kernel uses this K order, so their FP8/BF16/FP16 scores are
// the same sums in the same order (bit-identical across kernels). FP32
// (32x32x2) keeps lane half h on bytes [64 h, 64 h + 64) of the stage.
template <int DT>
__device__ __forceinline__ int lane_k_off(int lane) {
  const int h = lane >> 5;
  if constexpr (DT == EBSD_DTYPE_F32) return h * 64;
  else if constexpr (DT == EBSD_DTYPE_FP8_E4M3) return h * 32;
  else return h * 16;
}

// Power-of-two part of a positive scale (its exponent field): FP8 row scales
// act as E8M0 block scales, whose mantissa is implicitly 1.
__device__ __forceinline__ float pow2_part(float s) { return __uint_as_float(__float_as_uint(s) & 0x7f800000u); }

// One LDS stage of the K loop for a 128 (A) x 32 (B) wave tile; a_off and
// b_off include lane_k_off. A and B use the same K order, so every MFMA sums
// matching products.
template <int DT>
__device__ __forceinline__ void stage_mfma(const unsigned char* As, const unsigned char* Bs, int a_off,
                                           int b_off, floatx16 (&acc)[4]) {
  if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const intx4 b0 = *reinterpret_cast<const intx4*>(Bs + b_off + s * 64);
      const intx4 b1 = *reinterpret_cast<const intx4*>(Bs + b_off + s * 64 + 16);
      const intx8 b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) {
        const unsigned char* ap = As + a_off + bi * 32 * LDS_ROW + s * 64;
        const intx4 a0 = *reinterpret_cast<const intx4*>(ap);
        const intx4 a1 = *reinterpret_cast<const intx4*>(ap + 16);
        const intx8 a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
        // cbsz = blgp = 0: both operands OCP e4m3
        acc[bi] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc[bi], 0, 0, 0, kUnitScale, 0,
                                                                  kUnitScale);
      }
    }
  } else {
    constexpr int step = DT == EBSD_DTYPE_F32 ? 16 : 32;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const floatx4 braw = *reinterpret_cast<const floatx4*>(Bs + b_off + s * step);
      floatx4 araw[4];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
        araw[bi] = *reinterpret_cast<const floatx4*>(As + a_off + bi * 32 * LDS_ROW + s * step);
      if constexpr (DT == EBSD_DTYPE_F32) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int bi = 0; bi < 4; ++bi)
            acc[bi] = __builtin_amdgcn_mfma_f32_32x32x2f32(araw[bi][e], braw[e], acc[bi], 0, 0, 0);
      } else if constexpr (DT == EBSD_DTYPE_BF16) {
        const bf16x8 b = __builtin_bit_cast(bf16x8, braw);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
          acc[bi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, araw[bi]), b, acc[bi],
                                                            0, 0, 0);
      } else {
        const halfx8 b = __builtin_bit_cast(halfx8, braw);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
          acc[bi] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(halfx8, araw[bi]), b, acc[bi],
                                                           0, 0, 0);
      }
    }
  }
}


// Full K loop of one 128 x 128 block. arow / brow: this thread's four global
// row pointers (rows lr + 32 i of the A and B blocks, already offset by its
// 16-byte column lc); nk = ld_bytes / STAGE_BYTES stages. On return acc holds
// the block and every thread has passed a barrier after its last LDS read.
template <int DT>
__device__ __forceinline__ void tile_product(const unsigned char* const (&arow)[4],
                                             const unsigned char* const (&brow)[4], int nk,
                                             unsigned char* smem, int a_off, int b_off, int st_off,
                                             floatx16 (&acc)[4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[a][r] = 0.0f;

  floatx4 ra[4], rb[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ra[i] = *reinterpret_cast<const floatx4*>(arow[i]);
    rb[i] = *reinterpret_cast<const floatx4*>(brow[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    *reinterpret_cast<floatx4*>(smem + st_off + 32 * i * LDS_ROW) = ra[i];
    *reinterpret_cast<floatx4*>(smem + LDS_TILE + st_off + 32 * i * LDS_ROW) = rb[i];
  }
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = (kt + 1) < nk;
    if (more) {
      const int ko = (kt + 1) * STAGE_BYTES;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ra[i] = *reinterpret_cast<const floatx4*>(arow[i] + ko);
        rb[i] = *reinterpret_cast<const floatx4*>(brow[i] + ko);
      }
    }
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
    // Orders this stage's reads before the next overwrite of its buffer,
    // and (last stage) before the caller's next use of the LDS.
    __syncthreads();
  }
}

}  // namespace ebsd_tile
