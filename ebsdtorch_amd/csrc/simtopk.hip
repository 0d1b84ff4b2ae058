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
  int k_list;  // register list length (1 or 16)
  SimSched sc;
  int n_parts;  // partial lists per query (= B * R)
  int grid;
};

// 8 XCDs x 32 CUs x 2 resident workgroups on MI355X: one persistent
// workgroup per resident slot.
constexpr int kXcds = 8;
constexpr int kSlotsPerXcd = 64;

SimPlan plan_sim(int m, int n, int topk) {
  SimPlan p;
  p.k_list = topk <= 1 ? 1 : EBSD_MAX_TOPK;
  SimSched& s = p.sc;
  s.n_qtiles = (m + TILE - 1) / TILE;
  const int n_tiles = (n + TILE - 1) / TILE;
  // Largest A <= 8 that still gives every XCD a query group.
  s.A = 8;
  while (s.A > 1 && (s.n_qtiles + s.A - 1) / s.A < kXcds) s.A >>= 1;
  s.B = kSlotsPerXcd / s.A;
  s.n_qgroups = (s.n_qtiles + s.A - 1) / s.A;
  s.R = s.n_qgroups >= kXcds ? 1 : kXcds / s.n_qgroups;
  p.n_parts = s.B * s.R;
  s.tiles_per_part = (n_tiles + p.n_parts - 1) / p.n_parts;
  p.grid = kXcds * kSlotsPerXcd;
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
  const SimPlan p = plan_sim(m, n, topk);
  *bytes = workspace_bytes(p, m);
  return EBSD_OK;
}

template <int DT>
static int launch_sim(const SimPlan& p, const void* q, const float* q_scale, int m, const void* dict,
                      const float* d_scale, int n, int ld_bytes, int topk, float* top_val, int* top_idx,
                      int index_base, int merge, float* pval, int* pidx, hipStream_t st) {
  const dim3 grid(p.grid), block(256);
  const dim3 mgrid((m + 255) / 256);
  const unsigned char* Qb = static_cast<const unsigned char*>(q);
  const unsigned char* Db = static_cast<const unsigned char*>(dict);
  if (p.k_list == 1) {
    hipLaunchKernelGGL((sim_topk_kernel<DT, 1>), grid, block, 0, st, Qb, q_scale, m, Db, d_scale, n, ld_bytes,
                       p.sc, pval, pidx, index_base);
    EBSD_LAUNCH_CHECK();
    hipLaunchKernelGGL(topk_merge_kernel<1>, mgrid, dim3(256), 0, st, pval, pidx, p.n_parts, m, topk, top_val,
                       top_idx, merge);
  } else {
    hipLaunchKernelGGL((sim_topk_kernel<DT, EBSD_MAX_TOPK>), grid, block, 0, st, Qb, q_scale, m, Db, d_scale, n,
                       ld_bytes, p.sc, pval, pidx, index_base);
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
  const SimPlan p = plan_sim(m, n, topk);
  if (!workspace || ws_bytes < workspace_bytes(p, m)) return EBSD_ERR_WORKSPACE;
  float* pval = static_cast<float*>(workspace);
  int* pidx = reinterpret_cast<int*>(pval + (size_t)p.n_parts * p.k_list * m);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ld_bytes = ld * eb;
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
