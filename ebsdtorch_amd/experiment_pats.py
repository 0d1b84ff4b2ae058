"""Experimental pattern container (ebsd/experiment_pats.py:19-451): stores
the scan, hands out (binned) pattern chunks and keeps indexing results."""

from __future__ import annotations

from typing import List, Optional

import torch


class ExperimentPatterns(torch.nn.Module):
    def __init__(self, patterns: torch.Tensor):
        super().__init__()
        if patterns.dim() < 2:
            raise ValueError("'patterns' requires >= 2 dimensions (B, H, W) or (H, W)")
        if patterns.dim() < 3:
            patterns = patterns.unsqueeze(0)
        if patterns.dim() > 4:
            raise ValueError("3D Volume not yet supported... 'patterns' requires dimensions: "
                             "(SCAN_H, SCAN_W, H, W), (N_PATS, H, W), or (H, W)")
        self.pattern_shape = tuple(patterns.shape[-2:])
        self.spatial_shape = tuple(patterns.shape[:-2]) if patterns.dim() > 2 else (1, 1)
        self.patterns = patterns.reshape(-1, *self.pattern_shape)
        self.n_patterns = self.patterns.shape[0]
        self.n_pixels = self.pattern_shape[0] * self.pattern_shape[1]
        self.phases = None
        self.orientations = None
        self.best_metrics = None
        self.spatial_coords = None
        self.indexing_indices = None

    def set_spatial_coords(self, spatial_coords: torch.Tensor, indices: Optional[torch.Tensor] = None):
        if indices is None:
            if spatial_coords.shape[0] != self.n_patterns:
                raise ValueError("Spatial coordinates must have the same number of patterns as the "
                                 f"ExperimentPatterns object. Got {spatial_coords.shape[0]} and {self.n_patterns}.")
            self.spatial_coords = spatial_coords
        else:
            if spatial_coords.shape[0] != indices.shape[0]:
                raise ValueError("Spatial coordinates must have the same number of patterns as the indices.")
            self.spatial_coords[indices] = spatial_coords

    def get_spatial_coords(self, indices: Optional[torch.Tensor] = None) -> torch.Tensor:
        if self.spatial_coords is None:
            raise ValueError("Spatial coordinates must be set before retrieving them.")
        return self.spatial_coords if indices is None else self.spatial_coords[indices]

    def set_orientations(self, orientations: torch.Tensor, indices: Optional[torch.Tensor] = None):
        if orientations.dim() != 2:
            raise ValueError(f"Orientations must be quaternions (N_ORI, 4). Got {tuple(orientations.shape)}.")
        if orientations.shape[1] != 4:
            raise ValueError(f"Orientations must be quaternions (w, x, y, z). Got {orientations.shape[1]}.")
        if indices is None:
            if orientations.shape[0] != self.n_patterns:
                raise ValueError("Orientations must have the same number of patterns as the ExperimentPatterns "
                                 f"object. Got {orientations.shape[0]} orientations and {self.n_patterns} patterns.")
            self.orientations = orientations
        else:
            if indices.dim() != 1:
                raise ValueError(f"Indices must be (N_ORI,). Got {tuple(indices.shape)}.")
            if orientations.shape[0] != indices.shape[0]:
                raise ValueError("Orientations must have the same number of patterns as the indices.")
            self.orientations[indices] = orientations

    def get_orientations(self, indices: Optional[torch.Tensor] = None) -> torch.Tensor:
        return self.orientations if indices is None else self.orientations[indices]

    def get_patterns(self, indices: torch.Tensor, binning: int) -> torch.Tensor:
        """Patterns at `indices`, average-pooled by `binning`
        (experiment_pats.py:318-346)."""
        pats = self.patterns[indices]
        if binning > 1:
            pats = torch.nn.functional.avg_pool2d(pats, binning)
        return pats

    def subtract_overall_background(self) -> None:
        """Subtract the mean pattern of the scan (experiment_pats.py:223-232)."""
        self.patterns -= torch.mean(self.patterns, dim=0, keepdim=True)

    def contrast_enhance_clahe(self, clip_limit: float = 40.0, tile_grid_size: int = 4, n_bins: int = 64) -> None:
        """CLAHE of every pattern (experiment_pats.py:234-254, clahe.py)."""
        from .preprocessing import clahe_grayscale

        self.patterns = clahe_grayscale(self.patterns, clip_limit=clip_limit, n_bins=n_bins,
                                        grid_shape=(tile_grid_size, tile_grid_size))

    def normalize_per_pattern(self, norm_type: str = "minmax") -> None:
        """Per-pattern normalisation (experiment_pats.py:256-281)."""
        from .preprocessing import normalize_patterns

        self.patterns = self.patterns.to(torch.float32).contiguous()
        normalize_patterns(self.patterns.view(self.n_patterns, -1), norm_type)

    def standard_clean(self) -> None:
        """Background subtraction, minmax, CLAHE, minmax
        (experiment_pats.py:283-293)."""
        self.subtract_overall_background()
        self.normalize_per_pattern(norm_type="minmax")
        self.contrast_enhance_clahe()
        self.normalize_per_pattern(norm_type="minmax")

    def do_nlpar(self, k_rad: int = 3, coeff: float = 0.375) -> None:
        """NLPAR-denoise the scan in place (experiment_pats.py:296-316,
        preprocessing/nlpar.py); needs a 2D scan on the GPU."""
        from .preprocessing import nlpar

        if len(self.spatial_shape) != 2:
            raise ValueError(f"NLPAR only supported for 2D EBSD scans, got {self.spatial_shape}")
        self.patterns = nlpar(self.patterns.view(tuple(self.spatial_shape) + tuple(self.pattern_shape)),
                              k_rad=k_rad, coeff=coeff).view(-1, *self.pattern_shape)

    def get_indices_per_phase(self) -> List[torch.Tensor]:
        if self.phases is None:
            raise ValueError("Phases must be indexed before retrieving indices per phase.")
        return [torch.nonzero(self.phases == i, as_tuple=False).squeeze()
                for i in range(int(self.phases.max().item()) + 1)]

    def set_raw_indexing_results(self, quaternions: torch.Tensor, metrics: torch.Tensor, phase_id: int,
                                 indices: Optional[torch.Tensor] = None):
        """Store one phase's (N, k, 4) quaternions and (N, k) metrics
        (experiment_pats.py:367-385); `indices` additionally keeps the (N, k)
        dictionary indices."""
        if not hasattr(self, "raw_indexing_results"):
            self.raw_indexing_results = {}
        self.raw_indexing_results[phase_id] = (quaternions, metrics)
        if indices is not None:
            if not hasattr(self, "raw_indexing_indices"):
                self.raw_indexing_indices = {}
            self.raw_indexing_indices[phase_id] = indices

    def combine_indexing_results(self, higher_is_better: bool) -> None:
        """Keep the best top-1 match over phases (experiment_pats.py:387-451).
        With several phases the comparison is made on each phase's top-1
        metric, the evident intent of the reference (whose element-wise
        comparison of (N, k) against (N,) metrics does not broadcast)."""
        if not hasattr(self, "raw_indexing_results"):
            raise ValueError("Raw indexing results must be set before combining.")
        res = self.raw_indexing_results
        # results live where the indexing ran (the GPU), which need not be
        # where the patterns are kept: combine on the first phase's device
        dev = next(iter(res.values()))[1].device
        if len(res) == 1:
            pid = next(iter(res))
            self.phases = torch.full((self.n_patterns,), pid, dtype=torch.uint8, device=dev)
            self.orientations = res[pid][0][:, 0]
            self.best_metrics = res[pid][1][:, 0]
            return
        best = torch.full((self.n_patterns,), float("-inf") if higher_is_better else float("inf"), device=dev)
        phases = torch.full((self.n_patterns,), 255, dtype=torch.uint8, device=dev)
        ori = torch.zeros(self.n_patterns, 4, device=dev)
        for pid, (q, m) in res.items():
            q, m = q.to(dev), m.to(dev)
            m1 = m[:, 0].to(best.dtype)
            mask = m1 > best if higher_is_better else m1 < best
            best[mask] = m1[mask]
            phases[mask] = pid
            ori[mask] = q[mask, 0].to(ori.dtype)
        self.phases, self.orientations, self.best_metrics = phases, ori, best
