// Similarity GEMM with a fused streaming top-k for gfx950.
//
// Scores S[q, d] = <Q[q, :], Dict[d, :]> (utils/knn.py:96-97, metric
// "angular") are produced tile by tile on the matrix cores and reduced at once
// to a running top-k per query held in VGPRs; S never reaches memory.
//
// Element types: FP32 (v_mfma_f32_32x32x2_f32), BF16 / FP16
// (v_mfma_f32_32x32x16_{bf16,f16}) and OCP FP8 e4m3 (v_mfma_scale_f32_
// 32x32x64_f8f6f4). FP8 rows carry power-of-two scales: the 192 x 256 kernel
// feeds their exponents to the MFMA as E8M0 block scales, the 128 x 128 one
// multiplies its unit-scale scores by the same powers of two, exactly. All
// accumulate in FP32, in the K order of lane_k_off, so both kernels give the
// same scores bit for bit.
//
// Two kernels: the 128 x 128 one below (every type, any shape), and for
// BF16 / FP16 / FP8 with at least 192 dictionary rows and 4 stages of K the
// 192 x 256 one (sim_topk_bt_kernel, further down). The 128 x 128 kernel:
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

#include <algorithm>
#include <mutex>
#include <type_traits>

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

// What one persistent workgroup of either similarity kernel does: for query
// groups g = g_first, g_first + g_step, ... < n_qgroups it takes query tile
// g * A + a (when that tile exists) against dictionary tiles
// [t_begin, t_end) and writes partial list `part` of the tile's queries.
// Workgroups are dispatched round-robin over the 8 XCDs, so block % 8 is the
// XCD and block / 8 the slot (a, b): the B workgroups of a query tile share
// its stages in the XCD's L2, and the A workgroups of a part stream the same
// dictionary tiles in the same order. With fewer than 8 query groups the
// XCDs split the dictionary into R super-parts instead.
struct SlotWork {
  bool idle;
  int a, part, g_first, g_step, t_begin, t_end;
};

__host__ __device__ inline SlotWork slot_work(const SimSched& sc, int block, int n_tiles) {
  SlotWork w;
  const int xcd = block & 7;
  const int slot = block >> 3;
  w.a = slot / sc.B;
  const int b = slot - w.a * sc.B;
  int super;
  if (sc.R == 1) {
    w.g_first = xcd;
    w.g_step = 8;
    super = 0;
  } else {
    w.g_first = xcd % sc.n_qgroups;
    w.g_step = sc.n_qgroups;
    super = xcd / sc.n_qgroups;
  }
  w.idle = super >= sc.R;
  w.part = super * sc.B + b;
  w.t_begin = w.part * sc.tiles_per_part;
  w.t_end = w.t_begin + sc.tiles_per_part < n_tiles ? w.t_begin + sc.tiles_per_part : n_tiles;
  return w;
}

template <int DT, int K>
__global__ __launch_bounds__(256, 2) void sim_topk_kernel(
    const unsigned char* __restrict__ Q, const float* __restrict__ q_scale, int M,
    const unsigned char* __restrict__ Dm, const float* __restrict__ d_scale, int N, int ld_bytes,
    SimSched sc, float* __restrict__ pval, int* __restrict__ pidx, int index_base) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_BYTES];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  // 64 slots per XCD (two workgroups per CU); the mapping is slot_work's
  const SlotWork wk = slot_work(sc, blockIdx.x, (N + TILE - 1) / TILE);
  if (wk.idle) return;
  const int part = wk.part;
  const int t_begin = wk.t_begin;
  const int t_end = wk.t_end;

  for (int g = wk.g_first; g < sc.n_qgroups; g += wk.g_step) {
  const int qt = g * sc.A + wk.a;
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
  if constexpr (DT == EBSD_DTYPE_FP8_E4M3) qs = pow2_part(q_scale[min(q0 + wave * 32 + (lane & 31), M - 1)]);

  float tv[K];
  int ti[K];
  topk_init<K>(tv, ti);

  const int nk = ld_bytes / STAGE_BYTES;
  const int a_off = (lane & 31) * LDS_ROW + lane_k_off<DT>(lane);
  const int b_off = (wave * 32 + (lane & 31)) * LDS_ROW + lane_k_off<DT>(lane);
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
          if constexpr (DT == EBSD_DTYPE_FP8_E4M3) v = v * (qs * pow2_part(d_scale[d]));
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
// Low-precision variant (BF16 / FP16 / FP8): 192 x 256 block tiles, one
// 4-wave workgroup per CU, direct-to-LDS loads through a four-stage ring.
//
//   block   : 256 threads = 4 waves, one workgroup per CU (256 persistent
//             workgroups = 8 XCDs x 32); block tile 192 dictionary rows x 256
//             queries; wave w owns queries [64 w, 64 w + 64) against all 192
//             dictionary rows of the tile: 6 x 2 MFMA 32x32 tiles (192
//             accumulator registers, leaving the register allocator a quarter
//             of the AGPR file). Every score of a query is made in one wave,
//             so each query keeps a single list whose own K-th best is the
//             exact bound of what may still enter it
//   stages  : 64 bytes of K for all 448 rows (28 KiB) land in LDS straight
//             from global memory (global_load_lds_dwordx4, 7 per wave per
//             stage, plus one 4-byte load of the next tile's dictionary
//             scales), spread over the stage's MFMAs. Each wave reads stage
//             g + 1's operands into registers while stage g's MFMAs run
//             (each register as soon as its last MFMA of stage g issues), so
//             a stage starts with its operands in hand; stage g + 3 goes into
//             the buffer of stage g - 1, which the MFMAs of stage g - 1 have
//             consumed. Every tile is read from a uniform base (its first
//             row) at fixed per-lane offsets; the last, partial tile is read
//             as the dictionary's last 192 rows, and the rows it shares with
//             the tile before are skipped in the fold. Issuing is
//             branch-free: past the last stage the loads fetch the last tile
//             again into the free buffer, so every stage waits for the same
//             count and the MFMAs can be scheduled across the issue.
//             Rows of 256 bytes (4 stages) keep the work unit's 256 query
//             rows resident in LDS instead (loaded once, 64 KiB), and the
//             ring carries the 192 dictionary rows only (12 KiB, 3 loads per
//             wave and stage plus the scales): the query reload was a
//             quarter of the kernel's time at that width
//   barrier : one per stage, a bare s_barrier: each wave's own vmcnt wait
//             covers the direct loads it issued, so no full memory fence
//   LDS     : rows are 64 bytes, unpadded; 16-byte chunk c of row r sits at
//             chunk c ^ ((r / 4) % 4), so the sixteen rows read by sixteen
//             consecutive lanes hit sixteen different bank groups (and so do
//             eight rows of each lane half) while each load instruction still
//             writes one contiguous KiB
//   FP8     : the per-row scales are powers of two and enter the MFMA as its
//             E8M0 block scales (one byte per 32-row block, selected by
//             op_sel), so scores come out scaled with no epilogue multiply
//   top-k   : the two lane halves of a 32x32 tile hold the same 32 query
//             columns but different rows; lane half h keeps the lists of
//             column block h, one per lane. A tile's scores are screened in
//             place (per lane and 16 rows, the maximum against the K-th best
//             of the score's column); a 32-row block that some lane may take
//             is written to the wave's LDS scratch in the rows of the lanes
//             that own its columns, and each owner offers its marked scores
//             in ascending row order. At the end of a work unit each lane
//             writes its list as the partial list of its query.
constexpr int BT_D = 192;                                 // dictionary rows per tile
constexpr int BT_Q = 256;                                 // queries per tile
constexpr int BT_WQ = 64;                                 // queries per wave
constexpr int BT_KB = 64;                                 // bytes of K per row per stage
constexpr int BT_STAGE = (BT_D + BT_Q) * BT_KB;           // 28 KiB
constexpr int BT_NBUF = 4;
constexpr int BT_SCALE_OFF = BT_NBUF * BT_STAGE;          // two tiles' dictionary scales
// Scratch row of one lane in the fold: the 32 scores of a block for the query
// column whose list the lane owns, in row order (36 floats, so the 16-byte
// pieces written by 16 consecutive lanes start in 16 different bank groups).
constexpr int BT_SCR_ROW = 36;
constexpr int BT_SCR_OFF = BT_SCALE_OFF + 2 * 256 * 4;    // fold scratch, 9 KiB per wave
constexpr int BT_LDS = BT_SCR_OFF + 4 * 64 * BT_SCR_ROW * 4;
constexpr int BT_THREADS = 256;
// direct loads per wave per stage: 1-KiB data loads plus one scale load
constexpr int BT_LOADS = BT_STAGE / 1024 / (BT_THREADS / 64) + 1;
static_assert(BT_LOADS == 8, "the stage waits of sim_topk_bt_kernel count 8 loads per wave");
// Resident-query mode, for rows of exactly 4 stages (256 bytes): the 256
// query rows of a work unit (64 KiB) are loaded into LDS once, and the stage
// ring carries only the 192 dictionary rows (12 KiB per stage). Ring and query
// block together take the 112 KiB of the four full stages.
constexpr int BT_RQ_ROW_BYTES = 4 * BT_KB;
constexpr int BT_RQ_STAGE = BT_D * BT_KB;                    // 12 KiB
constexpr int BT_RQ_QOFF = BT_NBUF * BT_RQ_STAGE;            // query block
constexpr int BT_RQ_QSTAGE = BT_Q * BT_KB;                   // 16 KiB of it per stage
constexpr int BT_RQ_LOADS = BT_RQ_STAGE / 1024 / (BT_THREADS / 64) + 1;
static_assert(BT_RQ_LOADS == 4, "the stage waits of the resident-query mode count 4 loads per wave");
static_assert(BT_RQ_QOFF + 4 * BT_RQ_QSTAGE == BT_SCALE_OFF, "resident queries fill the full stage ring");
static_assert(BT_LDS <= 160 * 1024, "LDS budget");
static_assert(BT_Q == 4 * BT_WQ && BT_D == 6 * 32, "four waves of 64 queries, six 32-row blocks");
constexpr int BT_SLOTS_PER_XCD = 32;
constexpr int BT_MIN_STAGES = 4;  // the pipeline and the scale slots assume at least 4 stages per tile

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Direct global -> LDS loads of the 192 x 256 kernel: lane l's 16 (4) bytes
// land at LDS byte m0 + 16 l (4 l). Written as asm so that the compiler's wait
// tracking does not see them: it would otherwise wait for every outstanding
// LDS read before the first MFMA of a stage. Their completion is awaited
// explicitly (vmcnt in bt_wait_barrier); m0 must be wave-uniform.
__device__ __forceinline__ void bt_dma16(const void* src, unsigned m0) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(m0)
               : "memory", "m0");
}
__device__ __forceinline__ void bt_dma4(const void* src, unsigned m0) {
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off" ::"v"(src), "s"(m0)
               : "memory", "m0");
}
// The same loads from a wave-uniform base (SGPRs) plus a 32-bit per-lane
// byte offset: no 64-bit address arithmetic on the vector unit.
__device__ __forceinline__ void bt_dma16_s(const void* base, unsigned off, unsigned m0) {
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(off), "s"(base), "s"(m0)
               : "memory", "m0");
}
__device__ __forceinline__ void bt_dma4_s(const void* base, unsigned off, unsigned m0) {
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dword %0, %1" ::"v"(off), "s"(base), "s"(m0)
               : "memory", "m0");
}

__device__ __forceinline__ int bt_swz(int row, int chunk) { return row * BT_KB + ((chunk ^ ((row >> 2) & 3)) << 4); }

// Wait until at most N of this wave's vector memory operations are in flight
// (they complete in issue order) and this wave's LDS operations are done,
// then a bare workgroup barrier. The waits are the builtin (the compiler's
// wait tracking sees them); the empty asm statements keep memory accesses
// from moving across.
template <int N>
__device__ __forceinline__ void bt_wait_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  // s_waitcnt fields (gfx9): vmcnt[3:0] and [15:14], expcnt[6:4] (7: none),
  // lgkmcnt[11:8] (0)
  constexpr int imm = (N & 15) | ((N >> 4) << 14) | (7 << 4);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(imm);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// As bt_wait_barrier, but up to L of this wave's LDS reads may stay in flight
// across the barrier (reads of a buffer that no load of the next stage
// overwrites).
template <int N, int L>
__device__ __forceinline__ void bt_wait_barrier_l() {
  static_assert(N >= 0 && N < 64 && L >= 0 && L < 16, "vmcnt / lgkmcnt range");
  constexpr int imm = (N & 15) | ((N >> 4) << 14) | (7 << 4) | (L << 8);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(imm);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// MFMA operands of one stage for one wave: A = all 192 dictionary rows (six
// 32-row blocks), B = its 64 queries (two blocks), 32 bytes of K per lane
// and block.
struct BtFrag {
  intx8 a[6];
  intx8 b[2];
};

// Lane half h reads, per 32-row block, the 32 bytes of the 64-byte stage that
// its MFMAs consume in the K order of lane_k_off: FP8 one 32x32x64 step (bytes
// 32 h .. 32 h + 31 = chunks 2 h, 2 h + 1), BF16/FP16 two 32x32x16 steps
// (chunk h, then chunk 2 + h). Reads are issued in first-use order. abuf:
// the stage's dictionary rows, bbuf: its query rows (the 16 bytes of row r,
// chunk c at bt_swz(r, c) from either base).
template <int DT>
__device__ __forceinline__ intx8 bt_rd(const unsigned char* __restrict__ buf, int row, int lane) {
  const int h = lane >> 5;
  const int c0 = DT == EBSD_DTYPE_FP8_E4M3 ? 2 * h : h;
  const int c1 = DT == EBSD_DTYPE_FP8_E4M3 ? 2 * h + 1 : 2 + h;
  const intx4 x0 = *reinterpret_cast<const intx4*>(buf + bt_swz(row, c0));
  const intx4 x1 = *reinterpret_cast<const intx4*>(buf + bt_swz(row, c1));
  return intx8{x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
}

template <int DT>
__device__ __forceinline__ void bt_read(const unsigned char* __restrict__ abuf,
                                        const unsigned char* __restrict__ bbuf, int wq, int lane, BtFrag& f) {
  const int r32 = lane & 31;
  f.a[0] = bt_rd<DT>(abuf, r32, lane);
#pragma unroll
  for (int j = 0; j < 2; ++j) f.b[j] = bt_rd<DT>(bbuf, BT_WQ * wq + 32 * j + r32, lane);
#pragma unroll
  for (int i = 1; i < 6; ++i) f.a[i] = bt_rd<DT>(abuf, 32 * i + r32, lane);
}

// MFMAs IJ .. END - 1 (I = IJ / 2 dictionary block, J = IJ % 2 query block)
// of K step S of one stage (FP8: one step of 64, BF16/FP16: two of 16).
// FP8: byte i % 4 of sa[i / 4] / byte j of sb hold the E8M0 exponents of the
// lane's dictionary / query row in block i / j. Z: the first step of a tile,
// which starts the accumulators from zero (the MFMA's C operand is the
// constant 0).
template <int DT, int S, int IJ, int END, bool Z>
__device__ __forceinline__ void bt_mfma_step(const intx8 (&a)[6], const intx8 (&b)[2], const int (&sa)[2], int sb,
                                             floatx16 (&acc)[6][2]) {
  if constexpr (IJ < END) {
    constexpr int I = IJ / 2, J = IJ % 2;
    const floatx16 c = Z ? floatx16{} : acc[I][J];
    if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
      acc[I][J] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[I], b[J], c, 0, 0, I & 3, sa[I >> 2], J, sb);
    } else {
      const intx4 a4 = {a[I][4 * S], a[I][4 * S + 1], a[I][4 * S + 2], a[I][4 * S + 3]};
      const intx4 b4 = {b[J][4 * S], b[J][4 * S + 1], b[J][4 * S + 2], b[J][4 * S + 3]};
      if constexpr (DT == EBSD_DTYPE_BF16)
        acc[I][J] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a4),
                                                            __builtin_bit_cast(bf16x8, b4), c, 0, 0, 0);
      else
        acc[I][J] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(halfx8, a4),
                                                           __builtin_bit_cast(halfx8, b4), c, 0, 0, 0);
    }
    bt_mfma_step<DT, S, IJ + 1, END, Z>(a, b, sa, sb, acc);
  }
}

// All MFMAs of one stage (K steps S < NS, MFMAs IJ < 12 each: FP8 one step,
// BF16 / FP16 two), with hook(integral_constant<12 S + IJ>) called after each:
// the loads of a later stage go out in the shadow of the MFMAs, and the
// resident-query mode reads the next stage's dictionary operands into the
// registers an MFMA pair has finished with. Z: the tile's first stage
// (accumulators start from zero).
template <int DT, bool Z, int K = 0, typename Hook>
__device__ __forceinline__ void bt_mfma_hooked(const intx8 (&a)[6], const intx8 (&b)[2], const int (&sa)[2], int sb,
                                               floatx16 (&acc)[6][2], Hook&& hook) {
  constexpr int NS = DT == EBSD_DTYPE_FP8_E4M3 ? 1 : 2;
  if constexpr (K < 12 * NS) {
    constexpr int S = K / 12, IJ = K % 12;
    bt_mfma_step<DT, S, IJ, IJ + 1, Z && S == 0>(a, b, sa, sb, acc);
    hook(std::integral_constant<int, K>{});
    bt_mfma_hooked<DT, Z, K + 1>(a, b, sa, sb, acc, hook);
  }
}

// As bt_mfma_hooked, with the MFMAs of a K step in query-block-major order
// (K = 12 S + 6 J + I): query block 0's operands are done with after MFMA
// 12 (NS - 1) + 5, dictionary block i's after 12 (NS - 1) + 6 + i.
template <int DT, bool Z, int K = 0, typename Hook>
__device__ __forceinline__ void bt_mfma_hooked_jm(const intx8 (&a)[6], const intx8 (&b)[2], const int (&sa)[2], int sb,
                                                  floatx16 (&acc)[6][2], Hook&& hook) {
  constexpr int NS = DT == EBSD_DTYPE_FP8_E4M3 ? 1 : 2;
  if constexpr (K < 12 * NS) {
    constexpr int S = K / 12, J = (K % 12) / 6, I = K % 6;
    bt_mfma_step<DT, S, 2 * I + J, 2 * I + J + 1, Z && S == 0>(a, b, sa, sb, acc);
    hook(std::integral_constant<int, K>{});
    bt_mfma_hooked_jm<DT, Z, K + 1>(a, b, sa, sb, acc, hook);
  }
}

__device__ __forceinline__ int e8m0_of(float s) { return (int)((__float_as_uint(s) >> 23) & 0xffu); }

// Reads an accumulator register where it is used (the asm operand pins the
// value to its AGPR), so the compiler does not copy all accumulators to VGPRs
// ahead of the fold.
__device__ __forceinline__ float bt_acc(float a) {
  float v;
  asm volatile("v_accvgpr_read_b32 %0, %1" : "=v"(v) : "a"(a));
  return v;
}

// Maximum of three scores (v_max3_f32 as written: the scores are finite, so
// the NaN quieting that fmaxf would add in front of every operand is moot).
__device__ __forceinline__ float bt_max3(float a, float b, float c) {
  float d;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

// Largest of the 16 accumulator registers of one 32x32 tile in this lane.
__device__ __forceinline__ float bt_max16(const floatx16& a) {
  float v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = bt_acc(a[r]);
  const float m0 = bt_max3(v[0], v[1], v[2]), m1 = bt_max3(v[3], v[4], v[5]);
  const float m2 = bt_max3(v[6], v[7], v[8]), m3 = bt_max3(v[9], v[10], v[11]);
  const float m4 = bt_max3(v[12], v[13], v[14]);
  return bt_max3(bt_max3(m0, m1, m2), m3, bt_max3(m4, v[15], v[15]));
}

// Insert (v, id) into a descending list if it beats the last entry, for
// offers that arrive in ascending id order: an equal score then never
// outranks a listed one (whose id is smaller), so one comparison per step
// gives the order of topk_better.
//
// The new entry goes right after the last entry it does not beat: with
// c_j = v > tv[j] (monotone in j), entry j becomes tv[j - 1] where c_{j-1},
// else v where c_j, else stays, i.e. the median of (tv[j - 1], v, tv[j]) for
// the value (one v_med3_f32; the scores are finite) and two selects for the
// index. An offer that does not beat tv[K - 1] (all c_j false) leaves the
// list as it is, so the update runs without a branch.
template <int K>
__device__ __forceinline__ void topk_offer_ascending(float (&tv)[K], int (&ti)[K], float v, int id) {
  bool cj = v > tv[K - 1];
#pragma unroll
  for (int j = K - 1; j > 0; --j) {
    const bool cp = v > tv[j - 1];
    const int t = cj ? id : ti[j];
    ti[j] = cp ? ti[j - 1] : t;
    tv[j] = __builtin_amdgcn_fmed3f(tv[j - 1], v, tv[j]);
    cj = cp;
  }
  ti[0] = cj ? id : ti[0];
  tv[0] = cj ? v : tv[0];
}

// Bit j = 4 q + rr of a 16-bit mask to bit 8 q + rr (q, rr < 4).
__device__ __forceinline__ unsigned bt_spread_rows(unsigned p) {
  const unsigned y = (p | (p << 8)) & 0x00ff00ffu;
  return (y | (y << 4)) & 0x0f0f0f0fu;
}

// Offer queues of the resident-query mode: the query block of the LDS is free
// once the query operands sit in registers, and holds per lane a FIFO of
// BT_QCAP offers (score, index bits), slot s of lane l at qbuf[64 s + l] of
// its wave's part. Marked scores are queued at a few instructions each and
// inserted into the list in batches (bt_drain): every BT_DRAIN_TILES tiles,
// at the end of a work unit, and whenever a queue could overflow. Batching
// lets one insertion pass serve the marks of many lanes, and gives every wave
// the same insertion work at the same tiles, so the waves wait less for each
// other at the stage barriers.
constexpr int BT_QCAP = 32;
constexpr int BT_DRAIN_TILES = 16;
static_assert(4 * 64 * BT_QCAP * 8 <= 4 * BT_RQ_QSTAGE, "the offer queues fit in the query block");

// Offer the lane's queued entries to its list in queue order (ascending
// index, as offers arrive), the next one read ahead; empties the queue.
template <int K>
__device__ __forceinline__ void bt_drain(float (&tv)[K], int (&ti)[K], const float2* qbuf, int& qlen, int lane) {
  float2 e = qbuf[lane];
  for (int p = 0; p < qlen; ++p) {
    const float2 en = qbuf[64 * min(p + 1, BT_QCAP - 1) + lane];
    topk_offer_ascending<K>(tv, ti, e.x, __float_as_int(e.y));
    e = en;
  }
  qlen = 0;
}

// First dictionary row that tile t's loads read: the tile's own, except for
// the last, partial tile, which reads the dictionary's last BT_D rows (the
// leading ones were scored by the tile before and are skipped in the fold);
// loads past the end read that last tile again. Every tile is thus read from
// a uniform base at fixed per-lane offsets. Requires N >= BT_D.
__device__ __forceinline__ int bt_row0(int t, int N) { return min(t * BT_D, N - BT_D); }

// Fold a finished 192 x 64 wave tile into the lane's list. Lane (h, c)
// (h = lane / 32, c = lane % 32) holds, of every 32 x 32 tile, column c and
// rows 8 q + 4 h + rr (accumulator register 4 q + rr); its list belongs to
// column c of column block h. dbase: index of the tile's first dictionary
// row; lo: its first row not scored before (rows below lo were scored by
// the tile before, see bt_row0); scratch: the wave's 64
// scratch rows. Offers reach the list in ascending dictionary index. QUEUE:
// marked scores go to the lane's offer queue (qbuf, qlen) instead of straight
// into the list, and the bound of a column is its list's K-th best without the
// queued offers (a lower bound, so nothing that belongs is screened out).
template <int K, bool QUEUE>
__device__ __forceinline__ void bt_fold(const floatx16 (&acc)[6][2], float (&tv)[K], int (&ti)[K], int dbase,
                                        int lo, float* scratch, int lane, float2* qbuf, int& qlen) {
  const int h = lane >> 5, c = lane & 31;
  // the K-th best scores of the two columns whose scores the lane holds:
  // column c of block 0 (the list of lane c) and of block 1 (lane 32 + c)
  const auto t = __builtin_amdgcn_permlane32_swap(__float_as_uint(tv[K - 1]), __float_as_uint(tv[K - 1]), false,
                                                  false);
  const float ta = __uint_as_float(t[0]);
  const float tb = __uint_as_float(t[1]);
  // Screen every 32-row block first, with no branch in between: bit 2 i
  // (resp. 2 i + 1) is set when some lane may take a score of column block 0
  // (resp. 1) in row block i.
  // (kept as the ballots themselves and tested where they are used, so the
  // test stays a scalar compare)
  unsigned long long hit_a[6], hit_b[6];
  unsigned long long any = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const float ma = bt_max16(acc[i][0]);
    const float mb = bt_max16(acc[i][1]);
    hit_a[i] = __builtin_amdgcn_ballot_w64(!(ma < ta));
    hit_b[i] = __builtin_amdgcn_ballot_w64(!(mb < tb));
    any |= hit_a[i] | hit_b[i];
  }
  if (any == 0) return;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if ((hit_a[i] | hit_b[i]) == 0) continue;
    // Some lane may take a score: write the scores of the column block (or
    // blocks) concerned into the scratch rows of the lanes that own their
    // columns (row 8 q + 4 h + rr at float 8 q + 4 h + rr), mark the ones
    // that pass the owner's bound, and let every owner walk its marked rows
    // in ascending order (a divergent loop of max-popcount trips).
    // fail bit j = 4 q + rr: the score of row 8 q + 4 h + rr is below the
    // owner's bound, i.e. the sign of score - bound (+0 for equal finite
    // values), shifted in from j = 15 down to j = 0; a block that is not
    // staged fails everywhere
    unsigned fail_a = 0xffffu, fail_b = 0xffffu;
    if (hit_a[i] != 0) {
#pragma unroll
      for (int q = 3; q >= 0; --q) {
        floatx4 va;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) va[rr] = bt_acc(acc[i][0][4 * q + rr]);
        *reinterpret_cast<floatx4*>(scratch + BT_SCR_ROW * c + 8 * q + 4 * h) = va;
#pragma unroll
        for (int rr = 3; rr >= 0; --rr) fail_a = __builtin_amdgcn_alignbit(fail_a, __float_as_uint(va[rr] - ta), 31);
      }
    }
    if (hit_b[i] != 0) {
#pragma unroll
      for (int q = 3; q >= 0; --q) {
        floatx4 vb;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) vb[rr] = bt_acc(acc[i][1][4 * q + rr]);
        *reinterpret_cast<floatx4*>(scratch + BT_SCR_ROW * (32 + c) + 8 * q + 4 * h) = vb;
#pragma unroll
        for (int rr = 3; rr >= 0; --rr) fail_b = __builtin_amdgcn_alignbit(fail_b, __float_as_uint(vb[rr] - tb), 31);
      }
    }
    const unsigned bits_a = bt_spread_rows(~fail_a & 0xffffu) << (4 * h);
    const unsigned bits_b = bt_spread_rows(~fail_b & 0xffffu) << (4 * h);
    // the marks of the owned column from the partner lane (the other half)
    const unsigned other = h ? bits_a : bits_b;
    const auto o = __builtin_amdgcn_permlane32_swap(other, other, false, false);
    unsigned m = (h ? bits_b : bits_a) | (h ? o[0] : o[1]);
    // The partner's stores precede these loads in the wave's LDS order; the
    // clobbers keep the compiler from moving memory accesses across.
    asm volatile("" ::: "memory");
    if constexpr (QUEUE) {
      // make room first if some lane's marks could overflow its queue (after
      // a drain, 32 marks fit)
      if (__builtin_amdgcn_ballot_w64(qlen + __builtin_popcount(m) > BT_QCAP) != 0)
        bt_drain<K>(tv, ti, qbuf, qlen, lane);
      while (m != 0) {
        const int rho = __builtin_ctz(m);
        m &= m - 1;
        const int row = 32 * i + rho;
        if (row >= lo) {
          qbuf[64 * qlen + lane] = make_float2(scratch[BT_SCR_ROW * lane + rho], __int_as_float(dbase + row));
          ++qlen;
        }
      }
      asm volatile("" ::: "memory");
      continue;
    }
    // the score of the next marked row is read ahead of the current row's
    // insertion (for an empty rest, row 31's: a harmless in-bounds read)
    int rho = __builtin_ctz(m | 0x80000000u);
    float v = scratch[BT_SCR_ROW * lane + rho];
    while (m != 0) {
      m &= m - 1;
      const int rho_next = __builtin_ctz(m | 0x80000000u);
      const float v_next = scratch[BT_SCR_ROW * lane + rho_next];
      const int row = 32 * i + rho;
      // a row scored by the tile before is offered as -inf (no change)
      topk_offer_ascending<K>(tv, ti, row >= lo ? v : -INFINITY, dbase + row);
      rho = rho_next;
      v = v_next;
    }
    asm volatile("" ::: "memory");
  }
}

// Requires N >= BT_D and at least BT_MIN_STAGES stages (ld_bytes >= 256); see
// plan_sim. FP8 scales are used through their exponents (powers of two). RQ:
// resident-query mode (ld_bytes == BT_RQ_ROW_BYTES).
template <int DT, int K, bool RQ>
__global__ __launch_bounds__(BT_THREADS, 1) void sim_topk_bt_kernel(
    const unsigned char* __restrict__ Q, const float* __restrict__ q_scale, int M,
    const unsigned char* __restrict__ Dm, const float* __restrict__ d_scale, int N, int ld_bytes,
    SimSched sc, float* __restrict__ pval, int* __restrict__ pidx, int index_base) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[BT_LDS];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wq = wave;
  const int r32 = lane & 31, h = lane >> 5;
  // 32 slots per XCD (one workgroup per CU); the mapping is slot_work's
  const SlotWork wk = slot_work(sc, blockIdx.x, (N + BT_D - 1) / BT_D);
  if (wk.idle) return;
  const int part = wk.part;
  const int t_begin = wk.t_begin;
  const int t_end = wk.t_end;
  const int nk = ld_bytes / BT_KB;
  const int G = max(t_end - t_begin, 0) * nk;  // stages per work unit
  // direct-load lane mapping: load L writes rows 16 L .. 16 L + 15 of the
  // stage (rows < 192 dictionary, then queries), lane -> row 16 L + lane / 4
  // at physical chunk lane % 4, which holds logical chunk
  // (lane % 4) ^ ((row / 4) % 4) = (lane % 4) ^ ((lane / 16) % 4); wave w
  // issues loads L = w + 4 j, j < 7: three dictionary blocks, four query blocks
  const int lrow = lane >> 2;
  const int lchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const float* sscale = reinterpret_cast<const float*>(smem + BT_SCALE_OFF);
  float* scratch = reinterpret_cast<float*>(smem + BT_SCR_OFF) + wave * 64 * BT_SCR_ROW;

  for (int grp = wk.g_first; grp < sc.n_qgroups; grp += wk.g_step) {
    const int qt = grp * sc.A + wk.a;
    if (qt >= sc.n_qtiles) continue;  // uniform over the workgroup
    // (a part with no tiles, G == 0, still writes its empty lists)
    const int q0 = qt * BT_Q;
    // this wave's four query loads of every stage: query rows
    // 16 (wave + 4 j) + lrow, j < 4 (loads L = 12 + wave + 4 j), clamped, as
    // byte offsets from the unit's first row qbase (stage kt: qbase + 64 kt)
    const unsigned char* qbase = Q + (size_t)q0 * ld_bytes;
    unsigned qoff[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      qoff[j] = (unsigned)(min(16 * (wave + 4 * j) + lrow, M - 1 - q0) * ld_bytes + lchunk * 16);
    // FP8: E8M0 exponents of the lane's query rows, byte j for block j
    int sb = 0;
    if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
#pragma unroll
      for (int j = 0; j < 2; ++j) sb |= e8m0_of(q_scale[min(q0 + BT_WQ * wq + 32 * j + r32, M - 1)]) << (8 * j);
    }
    float tv[K];
    int ti[K];
    topk_init<K>(tv, ti);
    floatx16 acc[6][2];  // written from zero by the first stage of every tile

    // Loads of stage g, in 8 parts: k < 3 dictionary rows 16 (wave + 4 k) ..,
    // k = 3 .. 6 query rows 16 (wave + 4 (k - 3)) .., k = 7 the dictionary
    // scales of the tile after g's (tile t's into LDS slot t % 2: 256
    // entries, 64 per wave, entry e the scale of the tile's row min(e, 191):
    // entries 192 .. 255 are unused and stay inside the tile; BF16 / FP16
    // load bytes of the tile's first row instead, so every batch has 8
    // loads). Every address lies in the dictionary or the query block (tiles
    // from bt_row0, query rows clamped), also for g >= G; the buffer is
    // g % BT_NBUF.
    const unsigned lds0 = (unsigned)(uintptr_t)(lds_ptr_t)smem;
    const unsigned wave_s = __builtin_amdgcn_readfirstlane(wave);
    auto scale_m0 = [&](int t) { return lds0 + BT_SCALE_OFF + ((t & 1) * 256 + 64 * wave_s) * 4; };
    auto issue_scales = [&](int t) {
      if constexpr (DT == EBSD_DTYPE_FP8_E4M3)
        bt_dma4_s(d_scale + bt_row0(t, N), (unsigned)min(tid, BT_D - 1) * 4u, scale_m0(t));
      else
        bt_dma4_s(Dm + (size_t)bt_row0(t, N) * ld_bytes, (unsigned)lane * 4u, scale_m0(t));
    };
    // RQ: the loads of stage g are its three dictionary parts and the scales
    // (k = 0 .. 3); the query rows are loaded once per work unit (qload)
    constexpr int LOADS = RQ ? BT_RQ_LOADS : BT_LOADS;
    constexpr int STAGE = RQ ? BT_RQ_STAGE : BT_STAGE;
    auto issue_part = [&](int g, int k) {
      const int tl = g / nk;
      const int kt = g - tl * nk;
      const int t = t_begin + tl;
      const unsigned mb = lds0 + (g % BT_NBUF) * STAGE;
      if (k < 3) {
        const unsigned off = (unsigned)((16 * (wave + 4 * k) + lrow) * ld_bytes + lchunk * 16);
        bt_dma16_s(Dm + (size_t)bt_row0(t, N) * ld_bytes + kt * BT_KB, off, mb + (wave_s + 4 * k) * 1024);
      } else if (!RQ && k < 7) {
        bt_dma16_s(qbase + kt * BT_KB, qoff[k - 3], mb + (12 + wave_s + 4 * (k - 3)) * 1024);
      } else {
        issue_scales(t + 1);
      }
    };
    auto issue = [&](int g) {
#pragma unroll
      for (int k = 0; k < LOADS; ++k) issue_part(g, k);
    };
    // RQ: the work unit's query rows, stage kt's 64 bytes of all 256 rows at
    // BT_RQ_QOFF + kt * BT_RQ_QSTAGE in the stages' row layout (16 loads)
    auto qload = [&]() {
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          bt_dma16_s(qbase + kt * BT_KB, qoff[j], lds0 + BT_RQ_QOFF + kt * BT_RQ_QSTAGE + (wave_s + 4 * j) * 1024);
    };
    // E8M0 exponents of the lane's dictionary rows of tile t (byte i % 4 of
    // s[i / 4]: block i)
    auto dict_scales = [&](int t, int (&s)[2]) {
      s[0] = s[1] = 0;
      if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
        const float* sl = sscale + (t & 1) * 256 + r32;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i >> 2] |= e8m0_of(sl[32 * i]) << (8 * (i & 3));
      }
    };

    int sa[2] = {0, 0};
    const int n_t = G / max(nk, 1);
    // RQ: the wave's offer queues in the query block (see BT_QCAP)
    float2* qbuf = reinterpret_cast<float2*>(smem + BT_RQ_QOFF) + wave * 64 * BT_QCAP;
    int qlen = 0;
    auto fold = [&](int tl) {
      const int d0 = (t_begin + tl) * BT_D;
      const int r0 = bt_row0(t_begin + tl, N);
      bt_fold<K, RQ>(acc, tv, ti, index_base + r0, d0 - r0, scratch, lane, qbuf, qlen);
      if constexpr (RQ) {
        if ((tl % BT_DRAIN_TILES == BT_DRAIN_TILES - 1 || tl + 1 == n_t) && __builtin_amdgcn_ballot_w64(qlen > 0) != 0)
          bt_drain<K>(tv, ti, qbuf, qlen, lane);
      }
      // the next tile's scales came with this tile's stages
      if (tl + 1 < n_t) dict_scales(t_begin + tl + 1, sa);
      // no LDS operation of the fold stays in flight into the next tile's
      // stages (lgkmcnt(0); vmcnt and expcnt unconstrained), so the waits
      // there only count the stage's own reads
      __builtin_amdgcn_s_waitcnt(0xC07F);
    };
    if constexpr (!RQ) {
      // Operands one stage ahead in registers: during stage g the MFMAs use
      // stage g's operands (read during stage g - 1) and the hooks read stage
      // g + 1's into the registers as their last MFMA of stage g issues; the
      // loads of stage g go to stage g + 3, into the buffer of stage g - 1,
      // whose reads the MFMAs of stage g - 1 consumed. Stage g's barrier
      // therefore means: stage g + 1 has landed (one batch stays in flight)
      // and every wave is done with stage g - 1's buffer; the reads of stage
      // g's own buffer may still be in flight.
      // prologue: the first tile's scales and stages 0 .. 2 in flight, then
      // stage 0's operands
      BtFrag f;
      if (G > 0) {
        issue_scales(t_begin);
        for (int p = 0; p < BT_NBUF - 1; ++p) issue(p);
        bt_wait_barrier<2 * LOADS>();
        dict_scales(t_begin, sa);
        bt_read<DT>(smem, smem + BT_D * BT_KB, wq, lane, f);
      }
      constexpr int NS = DT == EBSD_DTYPE_FP8_E4M3 ? 1 : 2;
      constexpr int NMF = 12 * NS;
      constexpr int LAST = 12 * (NS - 1);  // first MFMA of the stage's last K step
      for (int tl = 0; tl < n_t; ++tl) {
        auto stage = [&](int g, auto first) {
          bt_wait_barrier_l<LOADS, 15>();
          const unsigned char* nbuf = smem + ((g + 1) % BT_NBUF) * STAGE;
          bt_mfma_hooked_jm<DT, decltype(first)::value>(f.a, f.b, sa, sb, acc, [&](auto k) {
            constexpr int kk = decltype(k)::value;
            // next-stage operands into the registers this MFMA was the last to use
            if constexpr (kk == LAST + 5) f.b[0] = bt_rd<DT>(nbuf + BT_D * BT_KB, BT_WQ * wq + r32, lane);
            if constexpr (kk >= LAST + 6) {
              f.a[kk - LAST - 6] = bt_rd<DT>(nbuf, 32 * (kk - LAST - 6) + r32, lane);
              if constexpr (kk == LAST + 11) f.b[1] = bt_rd<DT>(nbuf + BT_D * BT_KB, BT_WQ * wq + 32 + r32, lane);
            }
            // part j of stage g + 3's loads goes out after MFMA j * NMF / LOADS
            constexpr int j = (kk * LOADS + NMF - 1) / NMF;
            if constexpr (j < LOADS && j * NMF / LOADS == kk) {
              issue_part(g + BT_NBUF - 1, j);
              __builtin_amdgcn_sched_barrier(0);
            }
          });
        };
        stage(tl * nk, std::true_type{});
        for (int kt = 1; kt < nk; ++kt) stage(tl * nk + kt, std::false_type{});
        fold(tl);
      }
    } else {
      // Resident queries (nk == 4): the wave's query operands of all four
      // stages stay in registers for the whole unit (bres, 64 VGPRs), and the
      // dictionary operands of stage g + 1 are read during stage g's MFMAs,
      // block i into a[i] as soon as its two MFMAs have been issued. Stage
      // g's barrier therefore means: stage g + 1 has landed, and every wave
      // has read stage g, whose buffer then takes stage g + 4 (the lead stays
      // three stages: two batches in flight at every wait).
      // prologue: the queries, the first tile's scales and stages 0 .. 3
      intx8 bres[4][2];
      intx8 a[6];
      if (G > 0) {
        issue_scales(t_begin);
        qload();
        for (int p = 0; p < BT_NBUF; ++p) issue(p);
        bt_wait_barrier<(BT_NBUF - 1) * LOADS>();
        dict_scales(t_begin, sa);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            bres[kt][j] = bt_rd<DT>(smem + BT_RQ_QOFF + kt * BT_RQ_QSTAGE, BT_WQ * wq + 32 * j + r32, lane);
#pragma unroll
        for (int i = 0; i < 6; ++i) a[i] = bt_rd<DT>(smem, 32 * i + r32, lane);
      }
      auto stage = [&](int g, auto kt, auto first) {
        bt_wait_barrier<2 * LOADS>();
        const unsigned char* nbuf = smem + ((g + 1) % BT_NBUF) * STAGE;
        constexpr int NS = DT == EBSD_DTYPE_FP8_E4M3 ? 1 : 2;
        bt_mfma_hooked<DT, decltype(first)::value>(a, bres[decltype(kt)::value], sa, sb, acc, [&](auto kk) {
          constexpr int k = decltype(kk)::value;
          if constexpr (k < LOADS) issue_part(g + BT_NBUF, k);
          if constexpr (k / 12 == NS - 1 && (k % 12) % 2 == 1)
            a[(k % 12) / 2] = bt_rd<DT>(nbuf, 32 * ((k % 12) / 2) + r32, lane);
        });
      };
      for (int tl = 0; tl < n_t; ++tl) {
        stage(4 * tl, std::integral_constant<int, 0>{}, std::true_type{});
        stage(4 * tl + 1, std::integral_constant<int, 1>{}, std::false_type{});
        stage(4 * tl + 2, std::integral_constant<int, 2>{}, std::false_type{});
        stage(4 * tl + 3, std::integral_constant<int, 3>{}, std::false_type{});
        fold(tl);
      }
    }
    // The loads past the last stage land, and every wave is done with
    // the stage ring, the scales and its scratch, before the next unit's
    // loads reuse the LDS (this wait includes lgkmcnt(0) and a barrier).
    bt_wait_barrier<0>();
    // the lane's list is its query's partial list
    const int q = q0 + BT_WQ * wq + 32 * h + r32;
    if (q < M) {
#pragma unroll
      for (int e = 0; e < K; ++e) {
        const size_t o = (size_t)(part * K + e) * M + q;
        pval[o] = tv[e];
        pidx[o] = ti[e];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Two-workgroups-per-CU variant for rows of 256 bytes (sim_topk_b2_kernel):
// a wave's vector instructions do not overlap its own matrix instructions,
// but they do overlap those of another wave on the same SIMD. With one
// 4-wave workgroup per CU the fold of every tile (screen, staging, list
// insertion) adds to the MFMA time; here each CU runs two independent 4-wave
// workgroups (one wave of each per SIMD), each with its own stage ring and
// barriers, so that one workgroup's fold runs while the other's MFMAs do.
//
//   block   : 256 threads = 4 waves, two workgroups per CU (512 persistent
//             workgroups = 8 XCDs x 64 slots); block tile 192 dictionary rows
//             x 128 queries; wave w owns queries [32 w, 32 w + 32) against all
//             192 rows: 6 MFMA 32x32 tiles (96 accumulator registers), so that
//             the whole wave fits in the 256 registers of two waves per SIMD.
//             Every score of a query is made in one wave: one list per query,
//             held by lane c < 32 of its column (the other lane half holds
//             the column's other 16 rows of each tile and no list)
//   operands: the unit's 128 query rows are loaded once (through the stage
//             ring's memory, before the ring starts) into registers: each
//             wave's query operands of all four stages (32 VGPRs). The ring
//             carries the 192 dictionary rows (12 KiB per stage, three 1-KiB
//             direct loads per wave plus the scales); the dictionary operands
//             of stage g + 1 are read during stage g's MFMAs, block i as soon
//             as its MFMA is issued. Stage g's barrier means: stage g + 1 has
//             landed and every wave has read stage g, whose buffer then takes
//             stage g + 4
//   top-k   : the screen, staging and offer queues of sim_topk_bt_kernel's
//             resident mode, for one column block per wave (queues only for
//             the 32 list-holding lanes)
//   LDS     : 48 KiB ring (first the query block), 8 KiB queues, 2 KiB
//             scales, 18 KiB fold scratch: 76 KiB per workgroup
constexpr int B2_Q = 128;                                  // queries per tile
constexpr int B2_WQ = 32;                                  // queries per wave
constexpr int B2_QCAP = 8;                                 // queue slots per list
constexpr int B2_QUEUE_OFF = BT_NBUF * BT_RQ_STAGE;        // 48 KiB
constexpr int B2_SCALE_OFF = B2_QUEUE_OFF + 4 * 32 * B2_QCAP * 8;
constexpr int B2_SCR_OFF = B2_SCALE_OFF + 2 * 256 * 4;
constexpr int B2_LDS = B2_SCR_OFF + 4 * 32 * BT_SCR_ROW * 4;
constexpr int B2_SLOTS_PER_XCD = 64;
constexpr int B2_ASLOTS = 2;                               // dictionary operand registers (x 8)
static_assert(B2_LDS <= 80 * 1024, "two workgroups per CU");
static_assert(B2_Q * BT_RQ_ROW_BYTES <= B2_QUEUE_OFF, "the query block is loaded through the ring");

// Largest of the 16 accumulator registers of one 32x32 tile in this lane, for
// accumulators the compiler keeps in VGPRs. Written with fmaxf, not asm: the
// values come straight from MFMAs, and only compiler-visible reads get the
// wait states an MFMA result needs before a vector instruction may read it
// (the MFMA results are canonical, so fmaxf adds no quieting).
__device__ __forceinline__ float b2_max3(float a, float b, float c) {
  return __builtin_elementwise_maximum(__builtin_elementwise_maximum(a, b), c);
}
__device__ __forceinline__ float b2_max16(const floatx16& a) {
  // eight three-way maxima (v_maximum3_f32): five over triples, two over
  // their results and the last register, one pair
  const float r0 = b2_max3(a[0], a[1], a[2]), r1 = b2_max3(a[3], a[4], a[5]), r2 = b2_max3(a[6], a[7], a[8]);
  const float r3 = b2_max3(a[9], a[10], a[11]), r4 = b2_max3(a[12], a[13], a[14]);
  return __builtin_elementwise_maximum(b2_max3(r0, r1, r2), b2_max3(r3, r4, a[15]));
}

// Offer the lane's queued entries (slot s at qbuf[32 s + c]) to its list in
// queue order, the next one read ahead; empties the queue.
template <int K>
__device__ __forceinline__ void b2_drain(float (&tv)[K], int (&ti)[K], const float2* qbuf, int& qlen, int c) {
  float2 e = qbuf[c];
  for (int p = 0; p < qlen; ++p) {
    const float2 en = qbuf[32 * min(p + 1, B2_QCAP - 1) + c];
    topk_offer_ascending<K>(tv, ti, e.x, __float_as_int(e.y));
    e = en;
  }
  qlen = 0;
}

// Fold a finished 192 x 32 wave tile. Lane (h, c) holds column c and rows
// 8 q + 4 h + rr of every 32 x 32 tile; lane c (h = 0) owns the column's
// list. Marked scores are staged in the owner's scratch row and queued. Rows
// below `lo` were scored by the tile before (see bt_row0) and are not offered.
template <int K>
__device__ __forceinline__ void b2_fold(const floatx16 (&acc)[6], float (&tv)[K], int (&ti)[K], int dbase,
                                        int lo, float* scratch, int lane, float2* qbuf, int& qlen) {
  const int h = lane >> 5, c = lane & 31;
  // the column's K-th best (lane c's) in both lane halves
  const auto t = __builtin_amdgcn_permlane32_swap(__float_as_uint(tv[K - 1]), __float_as_uint(tv[K - 1]), false,
                                                  false);
  const float th = __uint_as_float(t[0]);
  // screen: one ballot per 32-row block, kept as a scalar mask
  unsigned long long hit[6];
  unsigned long long any = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    hit[i] = __builtin_amdgcn_ballot_w64(!(b2_max16(acc[i]) < th));
    any |= hit[i];
  }
  if (any == 0) return;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (hit[i] == 0) continue;
    // stage the block's scores in the owner's row (row 8 q + 4 h + rr at
    // float 8 q + 4 h + rr) and mark the ones that pass the bound (fail bit
    // 4 q + rr: sign of score - bound, shifted in from bit 15 down)
    unsigned fail = 0xffffu;
#pragma unroll
    for (int q = 3; q >= 0; --q) {
      floatx4 v;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) v[rr] = acc[i][4 * q + rr];
      *reinterpret_cast<floatx4*>(scratch + BT_SCR_ROW * c + 8 * q + 4 * h) = v;
#pragma unroll
      for (int rr = 3; rr >= 0; --rr) fail = __builtin_amdgcn_alignbit(fail, __float_as_uint(v[rr] - th), 31);
    }
    const unsigned bits = bt_spread_rows(~fail & 0xffffu) << (4 * h);
    // the owner's marks: its own rows and those of lane 32 + c
    const auto o = __builtin_amdgcn_permlane32_swap(bits, bits, false, false);
    unsigned m = h == 0 ? (o[0] | o[1]) : 0u;
    // the partner's stores precede these loads in the wave's LDS order
    asm volatile("" ::: "memory");
    if (__builtin_amdgcn_ballot_w64(qlen + __builtin_popcount(m) > B2_QCAP) != 0) {
      // a queue would overflow: insert the queued offers, then this block's
      // marks straight away (up to 32 of them, more than a queue holds)
      b2_drain<K>(tv, ti, qbuf, qlen, c);
      while (m != 0) {
        const int rho = __builtin_ctz(m);
        m &= m - 1;
        const int row = 32 * i + rho;
        topk_offer_ascending<K>(tv, ti, row >= lo ? scratch[BT_SCR_ROW * c + rho] : -INFINITY, dbase + row);
      }
    } else {
      while (m != 0) {
        const int rho = __builtin_ctz(m);
        m &= m - 1;
        const int row = 32 * i + rho;
        if (row >= lo) {
          qbuf[32 * qlen + c] = make_float2(scratch[BT_SCR_ROW * c + rho], __int_as_float(dbase + row));
          ++qlen;
        }
      }
    }
    asm volatile("" ::: "memory");
  }
}

// The six MFMAs (dictionary blocks I .. 5, block i's operand in ar[i % 3]) of
// K step S of one stage for the wave's query block b; hook(I) after the MFMA
// of block I. (FP8 only takes the two-workgroup kernel: one step.)
template <int DT, int S, bool Z, int I = 0, typename Hook>
__device__ __forceinline__ void b2_mfma_step(const intx8 (&ar)[B2_ASLOTS], const intx8& b, const int (&sa)[2], int sb,
                                             floatx16 (&acc)[6], Hook&& hook) {
  if constexpr (I < 6) {
    const floatx16 c = Z ? floatx16{} : acc[I];
    const intx8& a = ar[I % B2_ASLOTS];
    if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
      acc[I] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, I & 3, sa[I >> 2], 0, sb);
    } else {
      const intx4 a4 = {a[4 * S], a[4 * S + 1], a[4 * S + 2], a[4 * S + 3]};
      const intx4 b4 = {b[4 * S], b[4 * S + 1], b[4 * S + 2], b[4 * S + 3]};
      if constexpr (DT == EBSD_DTYPE_BF16)
        acc[I] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a4), __builtin_bit_cast(bf16x8, b4),
                                                         c, 0, 0, 0);
      else
        acc[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(halfx8, a4), __builtin_bit_cast(halfx8, b4),
                                                        c, 0, 0, 0);
    }
    hook(std::integral_constant<int, I>{});
    b2_mfma_step<DT, S, Z, I + 1>(ar, b, sa, sb, acc, hook);
  }
}

// Requires N >= BT_D and ld_bytes == BT_RQ_ROW_BYTES (4 stages); see plan_sim.
template <int DT, int K>
__global__ __launch_bounds__(BT_THREADS, 2) void sim_topk_b2_kernel(
    const unsigned char* __restrict__ Q, const float* __restrict__ q_scale, int M,
    const unsigned char* __restrict__ Dm, const float* __restrict__ d_scale, int N, int ld_bytes,
    SimSched sc, float* __restrict__ pval, int* __restrict__ pidx, int index_base) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[B2_LDS];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wq = wave;
  const int r32 = lane & 31, h = lane >> 5;
  constexpr int NK = BT_RQ_ROW_BYTES / BT_KB;  // 4 stages per tile
  constexpr int LOADS = BT_RQ_LOADS;
  constexpr int STAGE = BT_RQ_STAGE;
  static_assert(DT == EBSD_DTYPE_FP8_E4M3, "the two-workgroup kernel takes FP8 rows with per-row scales");
  // 64 slots per XCD (two workgroups per CU); the mapping is slot_work's
  const SlotWork wk = slot_work(sc, blockIdx.x, (N + BT_D - 1) / BT_D);
  if (wk.idle) return;
  const int part = wk.part;
  const int t_begin = wk.t_begin;
  const int t_end = wk.t_end;
  const int n_t = max(t_end - t_begin, 0);
  // direct-load lane mapping as in sim_topk_bt_kernel: load L writes rows
  // 16 L .. 16 L + 15, lane -> row 16 L + lane / 4 at physical chunk lane % 4
  const int lrow = lane >> 2;
  const int lchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const float* sscale = reinterpret_cast<const float*>(smem + B2_SCALE_OFF);
  float* scratch = reinterpret_cast<float*>(smem + B2_SCR_OFF) + wave * 32 * BT_SCR_ROW;
  float2* qbuf = reinterpret_cast<float2*>(smem + B2_QUEUE_OFF) + wave * 32 * B2_QCAP;
  const unsigned lds0 = (unsigned)(uintptr_t)(lds_ptr_t)smem;
  const unsigned wave_s = __builtin_amdgcn_readfirstlane(wave);

  for (int grp = wk.g_first; grp < sc.n_qgroups; grp += wk.g_step) {
    const int qt = grp * sc.A + wk.a;
    if (qt >= sc.n_qtiles) continue;  // uniform over the workgroup
    const int q0 = qt * B2_Q;
    // FP8: E8M0 exponent of the lane's query row
    int sb = 0;
    if constexpr (DT == EBSD_DTYPE_FP8_E4M3) sb = e8m0_of(q_scale[min(q0 + B2_WQ * wq + r32, M - 1)]);
    float tv[K];
    int ti[K];
    topk_init<K>(tv, ti);
    floatx16 acc[6];
    int qlen = 0;

    auto scale_m0 = [&](int t) { return lds0 + B2_SCALE_OFF + ((t & 1) * 256 + 64 * wave_s) * 4; };
    // Loads of stage g: k < 3 dictionary rows 16 (wave + 4 k) .. of its tile,
    // k = 3 the scales of the tile after g's (entry e of the 256: the scale
    // of the tile's row min(e, 191), so every read stays inside the tile).
    // Every tile is read from a uniform base (its first row, bt_row0) at the
    // lane's fixed byte offset.
    auto issue_scales = [&](int t) {
      bt_dma4_s(d_scale + bt_row0(t, N), (unsigned)min(tid, BT_D - 1) * 4u, scale_m0(t));
    };
    auto issue_part = [&](int g, int k) {
      const int tl = g / NK;
      const int kt = g - tl * NK;
      const int t = t_begin + tl;
      if (k < 3) {
        const unsigned off = (unsigned)((16 * (wave + 4 * k) + lrow) * BT_RQ_ROW_BYTES + lchunk * 16);
        bt_dma16_s(Dm + (size_t)bt_row0(t, N) * BT_RQ_ROW_BYTES + kt * BT_KB, off,
                   lds0 + (g % BT_NBUF) * STAGE + (wave_s + 4 * k) * 1024);
      } else {
        issue_scales(t + 1);
      }
    };
    auto dict_scales = [&](int t, int (&s)[2]) {
      s[0] = s[1] = 0;
      if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
        const float* sl = sscale + (t & 1) * 256 + r32;
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i >> 2] |= e8m0_of(sl[32 * i]) << (8 * (i & 3));
      }
    };

    int sa[2] = {0, 0};
    intx8 bres[NK];
    // dictionary operands rotate through three registers: block i of a stage
    // in ar[i % 3]; after the MFMA of block i, block i + 3 of the same stage
    // (i < 3) or block i - 3 of the next (i >= 3) is read into its place.
    // Stage g's buffer is read during stages g - 1 and g, so stage g's loads
    // go into the buffer of stage g - 1 (issued during stage g - 3: a lead of
    // two stages) and every wait leaves one batch in flight.
    intx8 ar[B2_ASLOTS];
    if (n_t > 0) {
      // the unit's 128 query rows, stage kt's 64 bytes of them at kt * 8 KiB
      // (the ring's memory, before the ring starts): 8 loads per wave
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const unsigned char* qs = Q + (size_t)min(q0 + 16 * (wave + 4 * j) + lrow, M - 1) * ld_bytes + lchunk * 16;
#pragma unroll
        for (int kt = 0; kt < NK; ++kt) bt_dma16(qs + kt * BT_KB, lds0 + kt * B2_Q * BT_KB + (wave_s + 4 * j) * 1024);
      }
      bt_wait_barrier<0>();
#pragma unroll
      for (int kt = 0; kt < NK; ++kt) bres[kt] = bt_rd<DT>(smem + kt * B2_Q * BT_KB, B2_WQ * wq + r32, lane);
      // every wave has its query operands before the ring overwrites them
      bt_wait_barrier<0>();
      // the first tile's scales and stages 0 .. 2
      issue_scales(t_begin);
#pragma unroll
      for (int p = 0; p < BT_NBUF - 1; ++p)
#pragma unroll
        for (int k = 0; k < LOADS; ++k) issue_part(p, k);
      bt_wait_barrier<2 * LOADS>();
      dict_scales(t_begin, sa);
#pragma unroll
      for (int i = 0; i < B2_ASLOTS; ++i) ar[i] = bt_rd<DT>(smem, 32 * i + r32, lane);
    }
    auto stage = [&](int tl, auto kt, auto first) {
      const int g = NK * tl + decltype(kt)::value;
      // the next stage's first two blocks (four reads of buffer g) may still
      // be in flight: no load of stage g writes that buffer
      bt_wait_barrier_l<LOADS, 2 * B2_ASLOTS>();
      const unsigned char* cbuf = smem + (g % BT_NBUF) * STAGE;
      const unsigned char* nbuf = smem + ((g + 1) % BT_NBUF) * STAGE;
      b2_mfma_step<DT, 0, decltype(first)::value>(ar, bres[decltype(kt)::value], sa, sb, acc, [&](auto i) {
        constexpr int I = decltype(i)::value;
        if constexpr (I < LOADS) issue_part(g + BT_NBUF - 1, I);
        if constexpr (I + B2_ASLOTS < 6)
          ar[I % B2_ASLOTS] = bt_rd<DT>(cbuf, 32 * (I + B2_ASLOTS) + r32, lane);
        else
          ar[I % B2_ASLOTS] = bt_rd<DT>(nbuf, 32 * (I + B2_ASLOTS - 6) + r32, lane);
      });
    };
    for (int tl = 0; tl < n_t; ++tl) {
      stage(tl, std::integral_constant<int, 0>{}, std::true_type{});
      stage(tl, std::integral_constant<int, 1>{}, std::false_type{});
      stage(tl, std::integral_constant<int, 2>{}, std::false_type{});
      stage(tl, std::integral_constant<int, 3>{}, std::false_type{});
      const int d0 = (t_begin + tl) * BT_D;
      const int r0 = bt_row0(t_begin + tl, N);
      b2_fold<K>(acc, tv, ti, index_base + r0, d0 - r0, scratch, lane, qbuf, qlen);
      if ((tl % BT_DRAIN_TILES == BT_DRAIN_TILES - 1 || tl + 1 == n_t) && __builtin_amdgcn_ballot_w64(qlen > 0) != 0)
        b2_drain<K>(tv, ti, qbuf, qlen, r32);
      if (tl + 1 < n_t) dict_scales(t_begin + tl + 1, sa);
      __builtin_amdgcn_s_waitcnt(0xC07F);
    }
    // the loads past the last stage land and every wave is done with
    // the LDS before the next unit's query rows overwrite the ring
    bt_wait_barrier<0>();
    const int q = q0 + B2_WQ * wq + r32;
    if (h == 0 && q < M) {
#pragma unroll
      for (int e = 0; e < K; ++e) {
        const size_t o = (size_t)(part * K + e) * M + q;
        pval[o] = tv[e];
        pidx[o] = ti[e];
      }
    }
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
// Candidates go in groups of kRescoreGroup, so each load of the query feeds
// that many products and as many dictionary rows are fetched at once (16
// candidates of 3,600 pixels for 65,536 queries: 4.07 ms one at a time,
// 3.04 ms in groups of 4, 2.87 ms of 8, 3.05 ms of 16; non-temporal row
// loads 2.72 ms, and at 14,400 pixels 11.0 -> 10.0 ms); a slot that holds no
// dictionary row (or pads the last group) reads row 0 and is dropped.
constexpr int kRescoreGroup = 8;

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
  for (int j0 = 0; j0 < n_cand; j0 += kRescoreGroup) {
    int id[kRescoreGroup];
    bool ok[kRescoreGroup];
    const floatx4* dr[kRescoreGroup];
#pragma unroll
    for (int u = 0; u < kRescoreGroup; ++u) {
      id[u] = j0 + u < n_cand ? cand[(size_t)q * n_cand + j0 + u] : -1;  // wave-uniform
      ok[u] = id[u] >= 0 && id[u] < N;  // else an empty slot or not a dictionary row
      dr[u] = reinterpret_cast<const floatx4*>(Dm + (size_t)(ok[u] ? id[u] : 0) * ld);
    }
    float acc[kRescoreGroup] = {};
    for (int c = lane; c < n4; c += 64) {
      const floatx4 a = qr[c];
#pragma unroll
      for (int u = 0; u < kRescoreGroup; ++u) {
        const floatx4 b = __builtin_nontemporal_load(&dr[u][c]);  // read once: streamed past the caches
        acc[u] += a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
      }
    }
#pragma unroll
    for (int u = 0; u < kRescoreGroup; ++u) {
      const float v = wave_sum(acc[u]);
      if (ok[u]) topk_offer<C>(tv, ti, v, id[u]);
    }
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
  bool bt;      // a low-precision large-tile kernel instead of the 128 x 128 one
  bool b2;      // ... the two-workgroups-per-CU one (sim_topk_b2_kernel)
  SimSched sc;
  int n_parts;  // partial lists per query (= B * R)
  int grid;
};

// 8 XCDs x 32 CUs on MI355X. The 128 x 128 kernel runs two resident
// workgroups per CU (64 slots per XCD), the 192 x 256 one a single 4-wave
// workgroup per CU (32 slots).
constexpr int kXcds = 8;
constexpr int kSlotsPerXcd = 64;
// Query tiles per XCD group of the 192 x 256 kernel (A of plan_groups): 16
// tiles x 2 dictionary parts; each dictionary tile is then streamed by 16 CUs
// of an XCD at once, which measured fastest (A = 2, 4, 8 were 5-20% slower;
// A = 32 took 16% longer with 3,600-byte rows, as long with 448).
// Rows of at most 256 bytes take 32 tiles x 1 part: the group's queries
// (2 MiB) still fit in the XCD's 4 MiB L2, and one part per query tile
// halves the fold's early, low-threshold work (28.1 -> 24.1 ms for
// 65,536 x 995,277 at 256 bytes).
constexpr int kBtGroupTiles = 16;
constexpr int kBtGroupTilesShort = 32;
constexpr int kBtShortRowBytes = 256;

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
  p.bt = dtype != EBSD_DTYPE_F32 && n >= BT_D && ld_bytes / BT_KB >= BT_MIN_STAGES;
  p.b2 = p.bt && ld_bytes == BT_RQ_ROW_BYTES && dtype == EBSD_DTYPE_FP8_E4M3;
  if (p.b2) {
    s.n_qtiles = (m + B2_Q - 1) / B2_Q;
    plan_groups(s, B2_SLOTS_PER_XCD, B2_SLOTS_PER_XCD, (n + BT_D - 1) / BT_D, p.n_parts);
    p.grid = kXcds * B2_SLOTS_PER_XCD;
  } else if (p.bt) {
    s.n_qtiles = (m + BT_Q - 1) / BT_Q;
    plan_groups(s, BT_SLOTS_PER_XCD, ld_bytes <= kBtShortRowBytes ? kBtGroupTilesShort : kBtGroupTiles,
                (n + BT_D - 1) / BT_D, p.n_parts);
    p.grid = kXcds * BT_SLOTS_PER_XCD;
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
  // the plan also depends on the row length; cover both kernels and both
  // groupings of the 192 x 256 one
  const size_t b1 = workspace_bytes(plan_sim(m, n, topk, EBSD_DTYPE_F32, STAGE_BYTES), m);
  const size_t b2 = workspace_bytes(plan_sim(m, n, topk, dtype, kBtShortRowBytes), m);
  const size_t b3 = workspace_bytes(plan_sim(m, n, topk, dtype, kBtShortRowBytes + BT_KB), m);
  *bytes = std::max(b1, std::max(b2, b3));
  return EBSD_OK;
}

extern "C" int ebsd_sim_schedule(int m, int n, int topk, int dtype, int ld, int* info, int* tile_visits,
                                 long long visits_len, int* list_writes, long long writes_len) {
  if (!info) return EBSD_ERR_NULL;
  if (m < 1 || n < 1 || ld < 1) return EBSD_ERR_SHAPE;
  if (topk < 1 || topk > EBSD_MAX_TOPK) return EBSD_ERR_TOPK;
  const int eb = elem_bytes(dtype);
  if (eb == 0) return EBSD_ERR_DTYPE;
  if (ld % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  const SimPlan p = plan_sim(m, n, topk, dtype, ld * eb);
  const int q_rows = p.b2 ? B2_Q : p.bt ? BT_Q : TILE;
  const int d_rows = p.bt ? BT_D : TILE;
  const int n_dtiles = (n + d_rows - 1) / d_rows;
  const int vals[EBSD_SCHEDULE_INFO] = {p.b2 ? 2 : p.bt ? 1 : 0, p.grid, p.n_parts, p.k_list, q_rows,
                                        d_rows, p.sc.n_qtiles, n_dtiles, p.sc.A, p.sc.B, p.sc.R};
  for (int i = 0; i < EBSD_SCHEDULE_INFO; ++i) info[i] = vals[i];
  const long long nv = (long long)p.sc.n_qtiles * n_dtiles;
  const long long nw = (long long)p.sc.n_qtiles * p.n_parts;
  if (tile_visits && visits_len < nv) return EBSD_ERR_WORKSPACE;
  if (list_writes && writes_len < nw) return EBSD_ERR_WORKSPACE;
  if (tile_visits) std::fill(tile_visits, tile_visits + nv, 0);
  if (list_writes) std::fill(list_writes, list_writes + nw, 0);
  // the kernels' own loop over query groups and dictionary tiles, per block
  for (int block = 0; block < p.grid; ++block) {
    const SlotWork w = slot_work(p.sc, block, n_dtiles);
    if (w.idle) continue;
    for (int g = w.g_first; g < p.sc.n_qgroups; g += w.g_step) {
      const int qt = g * p.sc.A + w.a;
      if (qt >= p.sc.n_qtiles) continue;
      if (w.part < 0 || w.part >= p.n_parts) return EBSD_ERR_INTERNAL;
      if (list_writes) ++list_writes[(long long)qt * p.n_parts + w.part];
      for (int t = w.t_begin; t < w.t_end; ++t) {
        if (t < 0 || t >= n_dtiles) return EBSD_ERR_INTERNAL;
        if (tile_visits) ++tile_visits[(long long)qt * n_dtiles + t];
      }
    }
  }
  return EBSD_OK;
}

template <int DT, int K>
static void launch_sim_kernel(const SimPlan& p, const unsigned char* Qb, const float* q_scale, int m,
                              const unsigned char* Db, const float* d_scale, int n, int ld_bytes, float* pval,
                              int* pidx, int index_base, hipStream_t st) {
  if constexpr (DT != EBSD_DTYPE_F32) {
    if constexpr (DT == EBSD_DTYPE_FP8_E4M3) {
      if (p.b2) {
        hipLaunchKernelGGL((sim_topk_b2_kernel<DT, K>), dim3(p.grid), dim3(BT_THREADS), 0, st, Qb, q_scale, m, Db,
                           d_scale, n, ld_bytes, p.sc, pval, pidx, index_base);
        return;
      }
    }
    if (p.bt) {
      if (ld_bytes == BT_RQ_ROW_BYTES)
        hipLaunchKernelGGL((sim_topk_bt_kernel<DT, K, true>), dim3(p.grid), dim3(BT_THREADS), 0, st, Qb, q_scale,
                           m, Db, d_scale, n, ld_bytes, p.sc, pval, pidx, index_base);
      else
        hipLaunchKernelGGL((sim_topk_bt_kernel<DT, K, false>), dim3(p.grid), dim3(BT_THREADS), 0, st, Qb, q_scale,
                           m, Db, d_scale, n, ld_bytes, p.sc, pval, pidx, index_base);
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

// ---------------------------------------------------------------------------
// Allocation-free variant: the partial top-k lists live in a static device
// buffer instead of a caller workspace. Every GPU that loads this library
// holds its own copy of `scratch` (EBSD_NOALLOC_SCRATCH_BYTES) until the
// process ends. Host code never uses the symbol's name as an address: the
// per-device address comes from hipGetSymbolAddress, taken once per device
// (with that device current) on the first call there. Nothing in the buffer
// needs clearing: the similarity kernels write every partial list before the
// merge kernel reads it. Calls that overlap on one GPU share the buffer and
// race; callers serialise them (one stream, or events).
// ---------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[EBSD_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (set up on first use).
int scratch_address(void** out) {
  int dev = 0;
  EBSD_HIP_RET(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return EBSD_ERR_SHAPE;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    EBSD_HIP_RET(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return EBSD_ERR_NULL;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return EBSD_OK;
}

}  // namespace

extern "C" int ebsd_sim_topk_noalloc(const void* q, const float* q_scale, int m, const void* dict,
                                     const float* d_scale, int n, int ld, int dtype, int topk,
                                     float* top_val, int* top_idx, int index_base, int merge, void* stream) {
  // same argument checks as ebsd_sim_topk, then the size check against the
  // static buffer before any device work
  if (!q || !dict || !top_val || !top_idx) return EBSD_ERR_NULL;
  if (m < 0 || n < 1 || ld < 1) return EBSD_ERR_SHAPE;
  if (topk < 1 || topk > EBSD_MAX_TOPK) return EBSD_ERR_TOPK;
  const int eb = elem_bytes(dtype);
  if (eb == 0) return EBSD_ERR_DTYPE;
  if (ld % ebsd_pad_multiple(dtype) != 0) return EBSD_ERR_SHAPE;
  if (m == 0) return EBSD_OK;
  const size_t need = workspace_bytes(plan_sim(m, n, topk, dtype, ld * eb), m);
  if (need > (size_t)EBSD_NOALLOC_SCRATCH_BYTES) return EBSD_ERR_SCRATCH;
  void* ptr = nullptr;
  const int st = scratch_address(&ptr);
  if (st != EBSD_OK) return st;
  return ebsd_sim_topk(q, q_scale, m, dict, d_scale, n, ld, dtype, topk, top_val, top_idx, index_base, merge,
                       ptr, (size_t)EBSD_NOALLOC_SCRATCH_BYTES, stream);
}
