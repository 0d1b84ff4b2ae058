"""Streaming k-nearest-neighbour search on the matrix cores
(utils/knn.py:106-287, metric "angular").

Same call sequence as ebsdtorch's ChunkedKNN: set_data_chunk(), then
query_all() or query_chunk(), then retrieve_topk(). Each query call runs
ebsd_sim_topk, which scores the chunk and merges it into the running
per-query top-k in one pass; scores are never materialised. Indices are
global dictionary rows (int64 on return), values are the angular metric.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import kernels
from ._lib import DTYPE_F32, EBSD_MAX_TOPK


class ChunkedKNN:
    def __init__(self, data_size: int, query_size: int, topk: int, match_device: torch.device,
                 distance_metric: str = "angular", match_dtype: torch.dtype = torch.float32,
                 quantized_via_ao: bool = False, precision: Optional[str] = None):
        if distance_metric != "angular":
            raise NotImplementedError(f"'{distance_metric}' not supported.")
        if not 1 <= topk <= EBSD_MAX_TOPK:
            raise ValueError(f"topk must be in [1, {EBSD_MAX_TOPK}], got {topk}")
        match_device = torch.device(match_device)
        if match_device.type != "cuda":
            raise ValueError("ChunkedKNN runs on the GPU; pass a HIP device")
        if precision is not None:
            self.dtype = kernels.PRECISIONS[precision]
        elif match_dtype in kernels.MATCH_DTYPES:
            self.dtype = kernels.MATCH_DTYPES[match_dtype]
        else:
            raise NotImplementedError(f"match_dtype {match_dtype} is not supported")
        self.data_size, self.query_size, self.topk = data_size, query_size, topk
        self.match_device = match_device
        self.distance_metric = distance_metric
        self.match_dtype = match_dtype
        self.big_better = True
        self.top_val = torch.full((query_size, topk), float("-inf"), dtype=torch.float32, device=match_device)
        self.top_idx = torch.full((query_size, topk), -1, dtype=torch.int32, device=match_device)
        self.prepared_data: Optional[kernels.PatternMatrix] = None
        self.curr_end = 0
        self.curr_start = 0

    def _as_matrix(self, x) -> kernels.PatternMatrix:
        if isinstance(x, kernels.PatternMatrix):
            if x.dtype != self.dtype:
                raise ValueError("pattern matrix dtype does not match the search dtype")
            return x
        x = x.to(self.match_device)
        return kernels.prep_patterns(x.reshape(x.shape[0], -1), self.dtype, zero_mean=False)

    def set_data_chunk(self, data_chunk) -> None:
        """Make `data_chunk` (N, D) (or a prepared PatternMatrix) the
        current dictionary chunk; its rows get the next global indices."""
        m = self._as_matrix(data_chunk)
        self.curr_start = self.curr_end
        self.curr_end += m.rows
        self.prepared_data = m

    def _run(self, q: kernels.PatternMatrix, lo: int, hi: int) -> None:
        if self.prepared_data is None:
            raise ValueError("Data chunk is not set.")
        if self.dtype != DTYPE_F32:
            kernels.sim_topk(q, self.prepared_data, self.topk, self.top_val[lo:hi], self.top_idx[lo:hi],
                             index_base=self.curr_start, merge=True)
            return
        # FP32: the chunk's best EBSD_MAX_TOPK candidates from the matrix
        # cores are re-scored exactly (ebsd_rescore) and merged into the
        # running lists, ordered by score then index (as the kernels order
        # ties)
        k_s = min(EBSD_MAX_TOPK, max(self.topk, self.prepared_data.rows))
        v = torch.empty((q.rows, k_s), dtype=torch.float32, device=self.match_device)
        i = torch.empty((q.rows, k_s), dtype=torch.int32, device=self.match_device)
        kernels.sim_topk(q, self.prepared_data, k_s, v, i, index_base=0, merge=False)
        k_r = min(self.topk, k_s)
        v, i = kernels.rescore(q, self.prepared_data, i, k_r)
        i = torch.where(i >= 0, i + self.curr_start, i)
        cv = torch.cat([self.top_val[lo:hi], v], 1)
        ci = torch.cat([self.top_idx[lo:hi], i], 1)
        # empty slots (index -1, score -inf) sort last
        key_i = torch.where(ci < 0, torch.iinfo(torch.int32).max, ci)
        o1 = torch.argsort(key_i, dim=1, stable=True)
        cv, ci = torch.gather(cv, 1, o1), torch.gather(ci, 1, o1)
        o2 = torch.argsort(cv, dim=1, descending=True, stable=True)[:, :self.topk]
        self.top_val[lo:hi] = torch.gather(cv, 1, o2)
        self.top_idx[lo:hi] = torch.gather(ci, 1, o2)

    def query_all(self, query) -> None:
        if self.prepared_data is None:
            raise ValueError("Data chunk is not set.")
        q = self._as_matrix(query)
        if q.rows != self.query_size:
            raise ValueError("query_all needs every query row")
        self._run(q, 0, q.rows)

    def query_chunk(self, query_chunk, query_start: int) -> None:
        if self.prepared_data is None:
            raise ValueError("Data chunk is not set.")
        q = self._as_matrix(query_chunk)
        self._run(q, int(query_start), int(query_start) + q.rows)

    def retrieve_topk(self) -> Tuple[torch.Tensor, torch.Tensor]:
        """(indices (N, k) int64, metrics (N, k)) sorted best first."""
        return self.top_idx.to(torch.int64), self.top_val
