// Similarity GEMM with a fused streaming top-k for gfx950.
//
// Scores S[q, d] = <Q[q, :], Dict[d, :]> (utils/knn.py:96-97, metric
// "angular") are produced tile by tile on the matrix cores and reduced at once
// to a running top-k per query held in VGPRs; S never reaches memory.
//
// Element types: FP32 (v_mfma_f32_32x32x2_f32), BF16 / FP16
// (v_mfma_f32_32x32x16_{bf16,f16}) and OCP FP8 e4m3 (v_mfma_scale_f32_
// 32x32x64_f8f6f4 with unit block scales; FP8 scores are multiplied by the
// per-row query and dictionary scales in the epilogue). All accumulate in
// FP32. Data movement is byte based and identical for every type: a stage is
// 128 bytes of K per row (32 FP32, 64 BF16/FP16 or 128 FP8 elements).
//
// Work decomposition
//   grid    : 512 persistent workgroups = 8 XCDs x 64 resident slots; each
//             XCD works through groups of A query tiles (128 rows each) x B
//             dictionary parts (A * B = 64), so that both operands of its
//             concurrent workgroups are shared in its L2 (plan_sim)
//   block   : 256 threads = 4 waves; wave w owns queries [32 w, 32 w + 32)
//             of the block and all 128 dictionary rows of the current tile,
//             i.e. a 128 x 32 score block = 4 MFMA 32x32 tiles (64 acc VGPRs)
// The K dimension streams through double-buffered LDS in 128-byte stages;
// the next stage is fetched into registers while the current one feeds the
// MFMAs. After each dictionary tile, every lane folds its 64 fresh scores
// (1 query column x 64 dictionary rows) into one register top-k list, so a
// query is covered by two lists (the two lane halves of its wave). At the
// end they are merged through LDS into one partial list per (query, split);
// a second kernel merges the splits (and, optionally, the caller's running
// state).

#include "mfma_tile.h"

namespace {

using namespace ebsd_tile;

// Work schedule of the persistent similarity kernel (see plan_sim).
struct SimSched {
  int A;               // query tiles per XCD block
  int B;               // dictionary parts per XCD block (A * B = 64 slots)
  int n_qtiles;
  int n_qgroups;       // ceil(n_qtiles / A)
  int R;               // dictionary super-parts when fewer than 8 query groups
  int tiles_per_part;  // dictionary tiles per part (n_parts = B * R parts)
};

template <int DT, int K>
__global__ __launch_bounds__(256, 2) void sim_topk_kernel(
    const unsigned char* __restrict__ Q, const float* __restrict__ q_scale, int M,
    const unsigned char* __restrict__ Dm, const float* __restrict__ d_scale, int N, int ld_bytes,
    SimSched sc, float* __restrict__ pval, int* __restrict__ pidx, int index_base) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_BYTES];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  // Workgroups are dispatched round-robin over the 8 XCDs, so blockIdx % 8 is
  // this workgroup's XCD and blockIdx / 8 its slot among that XCD's 64.
  // Slot (a, b) takes query tile a of the XCD's current group of A tiles and
  // dictionary part b: the B workgroups of a query tile share its stages in
  // the XCD's L2, and the A workgroups of a part stream the same dictionary
  // tiles in the same order.
  const int xcd = blockIdx.x & 7;
  const int slot = blockIdx.x >> 3;
  const int a = slot / sc.B;
  const int b = slot - a * sc.B;
  int g_first, g_step, super;
  if (sc.R == 1) {
    g_first = xcd;
    g_step = 8;
    super = 0;
  } else {
    g_first = xcd % sc.n_qgroups;
    g_step = sc.n_qgroups;
    super = xcd / sc.n_qgroups;
    if (super >= sc.R) return;
  }
  const int part = super * sc.B + b;
  const int n_tiles = (N + TILE - 1) / TILE;
  const int t_begin = part * sc.tiles_per_part;
  const int t_end = min(t_begin + sc.tiles_per_part, n_tiles);

  for (int g = g_first; g < sc.n_qgroups; g += g_step) {
  const int qt = g * sc.A + a;
  if (qt >= sc.n_qtiles) continue;  // uniform over the workgroup
  const int q0 = qt * TILE;
  const int split = part;

  // Global->LDS mapping: thread copies 16 bytes at byte lc of rows lr + 32 i
  // (i < 4) of both the dictionary tile and the query tile.
  const int lr = tid >> 3;
  const int lc = (tid & 7) * 16;
  const unsigned char* qrow[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) qrow[i] = Q + (size_t)min(q0 + lr + 32 * i, M - 1) * ld_bytes + lc;

  // FP8: this lane's query scale (its query column is fixed)
  float qs = 1.0f;
  if constexpr (DT == EBSD_DTYPE_FP8_E4M3) qs = q_scale[min(q0 + wave * 32 + (lane & 31), M - 1)];

  float tv[K];
  int ti[K];
  topk_init<K>(tv, ti);

  const int nk = ld_bytes / STAGE_BYTES;
  const int a_off = (lane & 31) * LDS_ROW + (lane >> 5) * 64;
  const int b_off = (wave * 32 + (lane & 31)) * LDS_ROW + (lane >> 5) * 64;
  const int st_off = lr * LDS_ROW + lc;

  for (int t = t_begin; t < t_end; ++t) {
    const int d0 = t * TILE;
    const unsigned char* drow[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) drow[i] = Dm + (size_t)min(d0 + lr + 32 * i, N - 1) * ld_bytes + lc;

    floatx16 acc[4];
    tile_product<DT>(drow, qrow, nk, smem, a_off, b_off, st_off, acc);

    // Fold the tile into the running list: this lane holds query column
    // wave*32 + lane%32 and 64 of the tile's 128 dictionary rows.
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int d = d0 + bi * 32 + mfma32_row(r, lane);
        if (d < N) {
          float v = acc[bi][r];
          if constexpr (DT == EBSD_DTYPE_FP8_E4M3) v = v * (qs * d_scale[d]);
          topk_offer<K>(tv, ti, v, index_base + d);
        }
      }
    }
  }

  // Merge the 2 lists per query (the two lane halves of its wave) in LDS.
  float* cv = reinterpret_cast<float*>(smem);  // [128 queries][2 lists][K]
  int* ci = reinterpret_cast<int*>(smem) + TILE * 2 * K;
  {
    const int ql = wave * 32 + (lane & 31);
    const int list = lane >> 5;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      cv[(ql * 2 + list) * K + j] = tv[j];
      ci[(ql * 2 + list) * K + j] = ti[j];
    }
  }
  __syncthreads();
  if (tid < TILE) {
    const int q = q0 + tid;
    const float* v = cv + tid * 2 * K;
    const int* ix = ci + tid * 2 * K;
    int h0 = 0, h1 = 0;
    for (int j = 0; j < K; ++j) {
      // A list can only run dry on the K-th pick, so both heads stay < K
      // whenever they are read.
      const float v0 = v[h0], v1 = v[K + h1];
      const int i0 = ix[h0], i1 = ix[K + h1];
      const bool take1 = topk_better(v1, i1, v0, i0);
      if (q < M) {
        const size_t o = (size_t)(split * K + j) * M + q;
        pval[o] = take1 ? v1 : v0;
        pidx[o] = take1 ? i1 : i0;
      }
      h0 += take1 ? 0 : 1;
      h1 += take1 ? 1 : 0;
    }
  }
  // the LDS lists are read before the next work unit reuses the stages
  __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Low-precision variant (BF16 / FP16 / FP8): larger tiles, direct-to-LDS
// loads, a three-stage LDS ring with one barrier per stage.
//
//   block   : 512 threads = 8 waves, one workgroup per CU (256 persistent
//             workgroups = 8 XCDs x 32); block tile 256 dictionary rows x 128
//             queries; wave (wd = w % 4, wq = w / 4) owns dictionary rows
//             [64 wd, 64 wd + 64) and queries [64 wq, 64 wq + 64): 2 x 2 MFMA
//             32x32 tiles, so every operand byte read from LDS feeds two MFMAs
//   stages  : 128 bytes of K for all 384 rows (48 KiB) land in LDS straight
//             from global memory (global_load_lds_dwordx4: 6 per wave); the
//             loads for stage g + 2 are issued as stage g starts, so two
//             stages are in flight while one is consumed
//   LDS     : rows are 128 bytes, unpadded; 16-byte chunk c of row r sits at
//             chunk c ^ ((r / 2) % 8), so the sixteen rows read by sixteen
//             consecutive lanes hit sixteen different bank groups (see
//             lp_swz) while each load instruction still writes one
//             contiguous KiB
//   top-k   : each lane keeps a list for each of its 2 query columns; after
//             a dictionary tile every lane offers its 64 fresh scores (FP8:
//             times the query and dictionary scales, the latter staged in
//             LDS per tile); at the end of a work unit the 8 lists of each
//             query are merged through LDS into one partial list
constexpr int LP_BM = 256;
constexpr int LP_BN = 128;
constexpr int LP_ROWS = LP_BM + LP_BN;
constexpr int LP_STAGE = LP_ROWS * STAGE_BYTES;   // 49,152 B
constexpr int LP_NBUF = 3;
constexpr int LP_SCALE_OFF = LP_NBUF * LP_STAGE;  // two tiles' dictionary scales
constexpr int LP_THR_OFF = LP_SCALE_OFF + 2 * LP_BM * 4;  // shared list thresholds
constexpr int LP_LDS = LP_THR_OFF + 2 * 4 * LP_BN * 4;
constexpr int LP_THREADS = 512;
constexpr int LP_LOADS = LP_ROWS / 8 / 8;  // 1-KiB direct loads per wave per stage
static_assert(LP_LOADS == 6, "the stage wait in sim_topk_lp_kernel counts 6 loads per wave");
constexpr int LP_SLOTS_PER_XCD = 32;

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// LDS has 64 four-byte banks: a 128-byte row covers half of them (even rows
// the lower half, odd rows the upper), and a ds_read_b128 serves 16 lanes at
// a time. XOR-ing the chunk with (row / 2) % 8 gives the 16 consecutive rows
// read by 16 lanes 16 different (row parity, chunk) slots: no bank conflict.
__device__ __forceinline__ int lp_swz_key(int row) { return (row >> 1) & 7; }

__device__ __forceinline__ int lp_swz(int row, int chunk) {
  return row * STAGE_BYTES + ((chunk ^ lp_swz_key(row)) << 4);
}

template <int DT>
__device__ __forceinline__ void lp_stage_mfma(const unsigned char* __restrict__ buf, int wd, int wq, int lane,
                                              floatx16 (&acc)[2][2]) {
  const int r32 = lane & 31, h = lane >> 5;
  if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int c0 = 4 * s + 2 * h;
      intx8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int ra = 64 * wd + 32 * i + r32;
        const int rb = LP_BM + 64 * wq + 32 * i + r32;
        const intx4 a0 = *reinterpret_cast<const intx4*>(buf + lp_swz(ra, c0));
        const intx4 a1 = *reinterpret_cast<const intx4*>(buf + lp_swz(ra, c0 + 1));
        const intx4 b0 = *reinterpret_cast<const intx4*>(buf + lp_swz(rb, c0));
        const intx4 b1 = *reinterpret_cast<const intx4*>(buf + lp_swz(rb, c0 + 1));
        a[i] = intx8{a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
        b[i] = intx8{b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, kUnitScale,
                                                                      0, kUnitScale);
    }
  } else {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int c = 4 * h + s;
      floatx4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const floatx4*>(buf + lp_swz(64 * wd + 32 * i + r32, c));
        b[i] = *reinterpret_cast<const floatx4*>(buf + lp_swz(LP_BM + 64 * wq + 32 * i + r32, c));
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if constexpr (DT == EBSD_DTYPE_BF16)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                                __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(halfx8, a[i]),
                                                               __builtin_bit_cast(halfx8, b[j]), acc[i][j], 0, 0, 0);
        }
    }
  }
}

// Requires N >= 256 and at least 3 stages (ld_bytes >= 384); see plan_sim.
// Wave 0 issues 4 extra loads (the scales) with the first stage of each tile,
// so its wait for "at most one batch outstanding" may cover a few loads of
// the next batch too: conservative, never short.
template <int DT, int K>
__global__ __launch_bounds__(LP_THREADS, 1) void sim_topk_lp_kernel(
    const unsigned char* __restrict__ Q, const float* __restrict__ q_scale, int M,
    const unsigned char* __restrict__ Dm, const float* __restrict__ d_scale, int N, int ld_bytes,
    SimSched sc, float* __restrict__ pval, int* __restrict__ pidx, int index_base) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[LP_LDS];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wd = wave & 3, wq = wave >> 2;
  // Workgroups are dispatched round-robin over the XCDs: blockIdx % 8 is the
  // XCD, blockIdx / 8 the slot (a, b): query tile a of the XCD's group of A,
  // dictionary part b (see the FP32 kernel for the grouping).
  const int xcd = blockIdx.x & 7;
  const int slot = blockIdx.x >> 3;
  const int a = slot / sc.B;
  const int b = slot - a * sc.B;
  int g_first, g_step, super;
  if (sc.R == 1) {
    g_first = xcd;
    g_step = 8;
    super = 0;
  } else {
    g_first = xcd % sc.n_qgroups;
    g_step = sc.n_qgroups;
    super = xcd / sc.n_qgroups;
    if (super >= sc.R) return;
  }
  const int part = super * sc.B + b;
  const int n_tiles = (N + LP_BM - 1) / LP_BM;
  const int t_begin = part * sc.tiles_per_part;
  const int t_end = min(t_begin + sc.tiles_per_part, n_tiles);
  const int nk = ld_bytes / STAGE_BYTES;
  const int G = max(t_end - t_begin, 0) * nk;  // stages per work unit
  // direct-load lane mapping: load i writes rows 8 i .. 8 i + 7, lane -> row
  // 8 i + lane / 8, physical chunk lane % 8, which holds logical chunk
  // (lane % 8) ^ lp_swz_key(8 i + lane / 8); i = wave + 8 j, so the key
  // depends on i only through its parity, the wave's parity
  const int lrow = lane >> 3;
  const int lchunk = (lane & 7) ^ lp_swz_key(8 * (wave & 1) + lrow);
  float* sscale = reinterpret_cast<float*>(smem + LP_SCALE_OFF);
  float* lthr = reinterpret_cast<float*>(smem + LP_THR_OFF);

  for (int grp = g_first; grp < sc.n_qgroups; grp += g_step) {
    const int qt = grp * sc.A + a;
    if (qt >= sc.n_qtiles) continue;  // uniform over the workgroup
    // (a part with no tiles, G == 0, still writes its empty lists)
    const int q0 = qt * LP_BN;

    float qs[2] = {1.0f, 1.0f};
    if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
#pragma unroll
      for (int j = 0; j < 2; ++j) qs[j] = q_scale[min(q0 + 64 * wq + 32 * j + (lane & 31), M - 1)];
    }
    float tv[2][K];
    int ti[2][K];
    topk_init<K>(tv[0], ti[0]);
    topk_init<K>(tv[1], ti[1]);
    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;


    auto issue = [&](int g) {
      const int t = t_begin + g / nk;
      const int kt = g - (g / nk) * nk;
      unsigned char* dst = smem + (g % LP_NBUF) * LP_STAGE;
      const int d0 = t * LP_BM;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = wave + 8 * j;
        const unsigned char* src = Dm + (size_t)min(d0 + 8 * i + lrow, N - 1) * ld_bytes + lchunk * 16 +
                                   kt * STAGE_BYTES;
        __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(dst + i * 1024), 16, 0, 0);
      }
#pragma unroll
      for (int j = 4; j < 6; ++j) {
        const int i = wave + 8 * j;  // query rows (loads 0-3 cover dictionary rows)
        const unsigned char* src = Q + (size_t)min(q0 + 8 * i + lrow - LP_BM, M - 1) * ld_bytes + lchunk * 16 +
                                   kt * STAGE_BYTES;
        __builtin_amdgcn_global_load_lds(src, (lds_ptr_t)(dst + i * 1024), 16, 0, 0);
      }
      if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
        // the tile's 256 dictionary scales, 4-byte direct loads (clamped rows)
        if (kt == 0 && wave == 0) {
          unsigned char* sdst = smem + LP_SCALE_OFF + (t & 1) * LP_BM * 4;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            __builtin_amdgcn_global_load_lds(d_scale + min(d0 + 64 * e + lane, N - 1), (lds_ptr_t)(sdst + e * 256),
                                             4, 0, 0);
        }
      }
    };

    if (G > 0) issue(0);
    if (G > 1) issue(1);
    for (int g = 0; g < G; ++g) {
      const int tl = g / nk;
      const int kt = g - tl * nk;
      // this wave's loads of stage g are complete once at most the next
      // stage's batch is outstanding
      if (g + 1 < G)
        asm volatile("s_waitcnt vmcnt(6)" ::: "memory");
      else
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      // stage g (and, at a tile's first stage, the tile's dictionary scales,
      // issued right after it) visible to all waves; every wave is done with
      // stage g - 1, whose buffer the loads of stage g + 2 overwrite, and
      // with the epilogue of the tile two back, whose scale slot the next
      // tile's scales overwrite
      __syncthreads();
      if (g + 2 < G) issue(g + 2);
      lp_stage_mfma<DT>(smem + (g % LP_NBUF) * LP_STAGE, wd, wq, lane, acc);

      if (kt == nk - 1) {
        const int d0 = (t_begin + tl) * LP_BM;
        const float* tsc = sscale + ((t_begin + tl) & 1) * LP_BM;
        // A score below the 16th best of any of the query's 8 lists cannot
        // reach its final top-K (the union's K-th best is at least as high),
        // so offers are filtered with the largest list threshold known: the
        // two lane halves of this wave now, the other three dictionary waves
        // as of the previous tile (thresholds only rise).
        float thr[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const float own = fmaxf(tv[j][K - 1], __shfl_xor(tv[j][K - 1], 32, 64));
          const int ql = 64 * wq + 32 * j + (lane & 31);
          float t = own;
          if (tl > 0) {
            const float* prev = lthr + ((tl - 1) & 1) * 4 * LP_BN;
#pragma unroll
            for (int w = 0; w < 4; ++w) t = fmaxf(t, prev[w * LP_BN + ql]);
          }
          thr[j] = t;
          if (lane < 32) lthr[(tl & 1) * 4 * LP_BN + wd * LP_BN + ql] = own;
        }
        // One offer site per (list, tile row block), looping over the 16
        // accumulator rows with a uniform register index: the epilogue stays
        // a few hundred instructions (fully unrolled, its 64 inlined list
        // insertions would not fit the instruction cache).
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i) {
#pragma unroll 1
            for (int r = 0; r < 16; ++r) {
              const int dl = 64 * wd + 32 * i + mfma32_row(r, lane);
              float v = acc[i][j][r];
              if (d0 + dl < N) {
                if constexpr (DT == EBSD_DTYPE_FP8_E4M3) v = v * (qs[j] * tsc[dl]);
                if (!(v < thr[j])) topk_offer<K>(tv[j], ti[j], v, index_base + d0 + dl);
              }
            }
          }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
      }
    }

    // Merge the 8 lists of each query: [128 queries][8 lists][K] in LDS.
    __syncthreads();
    float* cv = reinterpret_cast<float*>(smem);
    int* ci = reinterpret_cast<int*>(smem) + LP_BN * 8 * K;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int ql = 64 * wq + 32 * j + (lane & 31);
      const int list = wd * 2 + (lane >> 5);
#pragma unroll
      for (int e = 0; e < K; ++e) {
        cv[(ql * 8 + list) * K + e] = tv[j][e];
        ci[(ql * 8 + list) * K + e] = ti[j][e];
      }
    }
    __syncthreads();
    if (tid < LP_BN) {
      const int q = q0 + tid;
      const float* v = cv + tid * 8 * K;
      const int* ix = ci + tid * 8 * K;
      int head[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int e = 0; e < K; ++e) {
        int best = -1;
        float bv = -INFINITY;
        int bi = INT_MAX;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
          if (head[l] < K) {
            const float hv = v[l * K + head[l]];
            const int hi = ix[l * K + head[l]];
            if (best < 0 || topk_better(hv, hi, bv, bi)) {
              best = l;
              bv = hv;
              bi = hi;
            }
          }
        }
#pragma unroll
        for (int l = 0; l < 8; ++l) head[l] += (l == best) ? 1 : 0;
        if (q < M) {
          const size_t o = (size_t)(part * K + e) * M + q;
          pval[o] = bv;
          pidx[o] = bi;
        }
      }
    }
    // the lists are read before the next unit's loads overwrite the LDS
    __syncthreads();
  }
}

// Merge n_splits sorted partial lists (layout [split][K][M]) and, if `merge`,
// the caller's running state (M, topk), into the sorted state (M, topk).
template <int K>
__global__ __launch_bounds__(256) void topk_merge_kernel(
    const float* __restrict__ pval, const int* __restrict__ pidx, int n_splits, int M, int topk,
    float* __restrict__ out_val, int* __restrict__ out_idx, int merge) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= M) return;
  float tv[K];
  int ti[K];
  topk_init<K>(tv, ti);
  if (merge) {
    for (int j = 0; j < topk; ++j) {
      const float v = out_val[(size_t)q * topk + j];
      const int id = out_idx[(size_t)q * topk + j];
      topk_offer<K>(tv, ti, v, id);
    }
  }
  for (int s = 0; s < n_splits; ++s) {
    for (int j = 0; j < K; ++j) {
      const size_t o = (size_t)(s * K + j) * M + q;
      const float v = pval[o];
      const int id = pidx[o];
      if (!topk_better(v, id, tv[K - 1], ti[K - 1])) break;  // lists are sorted
      topk_insert<K>(tv, ti, v, id);
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < topk) {
      out_val[(size_t)q * topk + j] = tv[j];
      out_idx[(size_t)q * topk + j] = ti[j] == INT_MAX ? -1 : ti[j];
    }
  }
}

// FP32 re-scoring of candidate lists: one wave per query computes the exact
// FP32 dot product with each candidate row, then keeps the best `topk`.
template <int C>
__global__ __launch_bounds__(256) void rescore_kernel(
    const float* __restrict__ Q, int M, const float* __restrict__ Dm, int N, int ld,
    const int* __restrict__ cand, int n_cand, int topk, float* __restrict__ out_val,
    int* __restrict__ out_idx) {
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= M) return;
  const floatx4* qr = reinterpret_cast<const floatx4*>(Q + (size_t)q * ld);
  const int n4 = ld / 4;
  float tv[C];
  int ti[C];
  topk_init<C>(tv, ti);
  for (int j = 0; j < n_cand; ++j) {
    const int id = cand[(size_t)q * n_cand + j];  // wave-uniform
    if (id < 0 || id >= N) continue;  // empty slot or not a dictionary row
    const floatx4* dr = reinterpret_cast<const floatx4*>(Dm + (size_t)id * ld);
    float acc = 0.0f;
    for (int c = lane; c < n4; c += 64) {
      const floatx4 a = qr[c];
      const floatx4 b = dr[c];
      acc += a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    }
    acc = wave_sum(acc);
    topk_offer<C>(tv, ti, acc, id);
  }
  if (lane == 0) {
    for (int j = 0; j < topk; ++j) {
      out_val[(size_t)q * topk + j] = tv[j];
      out_idx[(size_t)q * topk + j] = ti[j] == INT_MAX ? -1 : ti[j];
    }
  }
}

struct SimPlan {
  int k_list;   // register list length (1 or 16)
  bool lp;      // low-precision kernel (sim_topk_lp_kernel) instead of the 128 x 128 one
  SimSched sc;
  int n_parts;  // partial lists per query (= B * R)
  int grid;
};

// 8 XCDs x 32 CUs on MI355X. The 128 x 128 kernel runs two resident
// workgroups per CU (64 slots per XCD), the low-precision one a single
// 8-wave workgroup per CU (32 slots).
constexpr int kXcds = 8;
constexpr int kSlotsPerXcd = 64;

void plan_groups(SimSched& s, int slots, int a_max, int n_tiles, int& n_parts) {
  // Largest A <= a_max that still gives every XCD a query group.
  s.A = a_max;
  while (s.A > 1 && (s.n_qtiles + s.A - 1) / s.A < kXcds) s.A >>= 1;
  s.B = slots / s.A;
  s.n_qgroups = (s.n_qtiles + s.A - 1) / s.A;
  s.R = s.n_qgroups >= kXcds ? 1 : kXcds / s.n_qgroups;
  n_parts = s.B * s.R;
  s.tiles_per_part = (n_tiles + n_parts - 1) / n_parts;
}

SimPlan plan_sim(int m, int n, int topk, int dtype, int ld_bytes) {
  SimPlan p;
  p.k_list = topk <= 1 ? 1 : EBSD_MAX_TOPK;
  SimSched& s = p.sc;
  p.lp = dtype != EBSD_DTYPE_F32 && n >= LP_BM && ld_bytes / STAGE_BYTES >= 3;
  if (p.lp) {
    s.n_qtiles = (m + LP_BN - 1) / LP_BN;
    plan_groups(s, LP_SLOTS_PER_XCD, 16, (n + LP_BM - 1) / LP_BM, p.n_parts);
    p.grid = kXcds * LP_SLOTS_PER_XCD;
  } else {
    s.n_qtiles = (m + TILE - 1) / TILE;
    plan_groups(s, kSlotsPerXcd, 8, (n + TILE - 1) / TILE, p.n_parts);
    p.grid = kXcds * kSlotsPerXcd;
  }
  return p;
}

size_t workspace_bytes(const SimPlan& p, int m) {
  return (size_t)p.n_parts * p.k_list * (size_t)m * (sizeof(float) + sizeof(int));
}

}  // namespace

extern "C" int ebsd_rescore(const float* q, int m, const float* dict, int n, int ld, const int* cand,
                            int n_cand, int topk, float* out_val, int* out_idx, void* stream) {
  if (!q || !dict || !cand || !out_val || !out_idx) return EBSD_ERR_NULL;
  if (m < 0 || n < 1 || ld < 1 || ld % 32 != 0) return EBSD_ERR_SHAPE;
  if (n_cand < 1 || n_cand > EBSD_MAX_TOPK || topk < 1 || topk > n_cand) return EBSD_ERR_TOPK;
  if (((reinterpret_cast<uintptr_t>(q) | reinterpret_cast<uintptr_t>(dict)) & 15) != 0) return EBSD_ERR_ALIGN;
  if (m == 0) return EBSD_OK;
  hipLaunchKernelGGL(rescore_kernel<EBSD_MAX_TOPK>, dim3((m + 3) / 4), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const float*>(q), m,
                     static_cast<const float*>(dict), n, ld, cand, n_cand, topk, out_val, out_idx);
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

static int elem_bytes(int dtype) {
  switch (dtype) {
    case EBSD_DTYPE_F32: return 4;
    case EBSD_DTYPE_BF16:
    case EBSD_DTYPE_F16: return 2;
    case EBSD_DTYPE_FP8_E4M3: return 1;
  }
  return 0;
}

extern "C" int ebsd_sim_topk_workspace(int m, int n, int topk, int dtype, size_t* bytes) {
  if (!bytes) return EBSD_ERR_NULL;
  if (m < 0 || n < 1) return EBSD_ERR_SHAPE;
  if (topk < 1 || topk > EBSD_MAX_TOPK) return EBSD_ERR_TOPK;
  if (elem_bytes(dtype) == 0) return EBSD_ERR_DTYPE;
  // the plan also depends on the row length; cover both kernels
  const size_t b1 = workspace_bytes(plan_sim(m, n, topk, EBSD_DTYPE_F32, STAGE_BYTES), m);
  const size_t b2 = workspace_bytes(plan_sim(m, n, topk, dtype, 3 * STAGE_BYTES), m);
  *bytes = b1 > b2 ? b1 : b2;
  return EBSD_OK;
}

template <int DT, int K>
static void launch_sim_kernel(const SimPlan& p, const unsigned char* Qb, const float* q_scale, int m,
                              const unsigned char* Db, const float* d_scale, int n, int ld_bytes, float* pval,
                              int* pidx, int index_base, hipStream_t st) {
  if constexpr (DT != EBSD_DTYPE_F32) {
    if (p.lp) {
      hipLaunchKernelGGL((sim_topk_lp_kernel<DT, K>), dim3(p.grid), dim3(LP_THREADS), 0, st, Qb, q_scale, m, Db,
                         d_scale, n, ld_bytes, p.sc, pval, pidx, index_base);
      return;
    }
  }
  hipLaunchKernelGGL((sim_topk_kernel<DT, K>), dim3(p.grid), dim3(256), 0, st, Qb, q_scale, m, Db, d_scale, n,
                     ld_bytes, p.sc, pval, pidx, index_base);
}

template <int DT>
static int launch_sim(const SimPlan& p, const void* q, const float* q_scale, int m, const void* dict,
                      const float* d_scale, int n, int ld_bytes, int topk, float* top_val, int* top_idx,
                      int index_base, int merge, float* pval, int* pidx, hipStream_t st) {
  const dim3 mgrid((m + 255) / 256);
  const unsigned char* Qb = static_cast<const unsigned char*>(q);
  const unsigned char* Db = static_cast<const unsigned char*>(dict);
  if (p.k_list == 1) {
    launch_sim_kernel<DT, 1>(p, Qb, q_scale, m, Db, d_scale, n, ld_bytes, pval, pidx, index_base, st);
    EBSD_LAUNCH_CHECK();
    hipLaunchKernelGGL(topk_merge_kernel<1>, mgrid, dim3(256), 0, st, pval, pidx, p.n_parts, m, topk, top_val,
                       top_idx, merge);
  } else {
    launch_sim_kernel<DT, EBSD_MAX_TOPK>(p, Qb, q_scale, m, Db, d_scale, n, ld_bytes, pval, pidx, index_base, st);
    EBSD_LAUNCH_CHECK();
    hipLaunchKernelGGL(topk_merge_kernel<EBSD_MAX_TOPK>, mgrid, dim3(256), 0, st, pval, pidx, p.n_parts, m,
                       topk, top_val, top_idx, merge);
  }
  EBSD_LAUNCH_CHECK();
  return EBSD_OK;
}

extern "C" int ebsd_sim_topk(const void* q, const float* q_scale, int m, const void* dict,
                             const float* d_scale, int n, int ld, int dtype, int topk,
                             float* top_val, int* top_idx, int index_base, int merge,
                             void* workspace, size_t ws_bytes, void* stream) {
  if (!q || !dict || !top_val || !top_idx) return EBSD_ERR_NULL;
  if (m < 0 || n < 1 || ld < 1) return EBSD_ERR_SHAPE;
  if (topk < 1 || topk > EBSD_MAX_TOPK) return EBSD_ERR_TOPK;
  const int eb = elem_bytes(dtype);
  if (eb == 0) return EBSD_ERR_DTYPE;
  if (dtype == EBSD_DTYPE_FP8_E4M3 && (!q_scale || !d_scale)) return EBSD_ERR_NULL;
  if (ld % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  if ((int64_t)index_base + n > (int64_t)INT_MAX || index_base < 0) return EBSD_ERR_SHAPE;
  if (((reinterpret_cast<uintptr_t>(q) | reinterpret_cast<uintptr_t>(dict)) & 15) != 0)
    return EBSD_ERR_ALIGN;
  if (m == 0) return EBSD_OK;
  const int ld_bytes = ld * eb;
  const SimPlan p = plan_sim(m, n, topk, dtype, ld_bytes);
  if (!workspace || ws_bytes < workspace_bytes(p, m)) return EBSD_ERR_WORKSPACE;
  float* pval = static_cast<float*>(workspace);
  int* pidx = reinterpret_cast<int*>(pval + (size_t)p.n_parts * p.k_list * m);
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case EBSD_DTYPE_F32:
      return launch_sim<EBSD_DTYPE_F32>(p, q, q_scale, m, dict, d_scale, n, ld_bytes, topk, top_val, top_idx,
                                        index_base, merge, pval, pidx, st);
    case EBSD_DTYPE_BF16:
      return launch_sim<EBSD_DTYPE_BF16>(p, q, q_scale, m, dict, d_scale, n, ld_bytes, topk, top_val, top_idx,
                                         index_base, merge, pval, pidx, st);
    case EBSD_DTYPE_F16:
      return launch_sim<EBSD_DTYPE_F16>(p, q, q_scale, m, dict, d_scale, n, ld_bytes, topk, top_val, top_idx,
                                        index_base, merge, pval, pidx, st);
    case EBSD_DTYPE_FP8_E4M3:
      return launch_sim<EBSD_DTYPE_FP8_E4M3>(p, q, q_scale, m, dict, d_scale, n, ld_bytes, topk, top_val,
                                             top_idx, index_base, merge, pval, pidx, st);
  }
  return EBSD_ERR_DTYPE;
}
