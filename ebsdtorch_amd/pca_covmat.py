"""Streaming covariance matrix on the GPU (utils/pca_covmat.py:94-158).

Same interface as ebsdtorch's OnlineCovMatrix: call the object on batches
(B, n_features), then get_covmat() / get_eigenvectors(). Batches are merged
by the HIP kernel ebsd_cov_update (Chan's batched Welford update, FP32
accumulation on the matrix cores); eigenvectors come from an FP64 symmetric
eigensolver on the host (fp64_eigh), ascending eigenvalues as in the
reference.
"""

from __future__ import annotations

import torch

from . import kernels
from ._lib import DTYPE_F32


def fp64_eigh(matrix: torch.Tensor, dtype: torch.dtype = torch.float64):
    """(eigenvalues ascending, eigenvectors as columns) of a symmetric
    matrix, solved in float64 (or `dtype`) by the host's LAPACK
    (torch.linalg.eigh on the CPU) and returned on the matrix's device.

    Every eigendecomposition of this package goes through here. The device
    solver (hipSOLVER syevd through torch) is not used: one profiled run
    (rocprofv3 with hardware counters) ended with a segmentation fault inside
    hipsolverDsyevd on a 3,600 x 3,600 matrix, with no cause established, so
    the solve stays on the host at every size. It runs once per dictionary
    (or covariance) and costs about 9 D^3 flops: ~1 s at D = 3,600, ~8 s at
    D = 6,400 and about a minute at D = 14,400 on 8-16 host cores."""
    dev = matrix.device
    evals, evecs = torch.linalg.eigh(matrix.detach().to("cpu", dtype))
    return evals.to(dev), evecs.to(dev)


class OnlineCovMatrix(torch.nn.Module):
    def __init__(self, n_features: int, covmat_dtype: torch.dtype = torch.float32,
                 delta_dtype: torch.dtype = torch.float32, correlation: bool = False,
                 device: torch.device = None):
        super().__init__()
        if covmat_dtype != torch.float32 or delta_dtype != torch.float32:
            raise NotImplementedError("the covariance is accumulated in float32")
        self.n_features = n_features
        self.covmat_dtype = covmat_dtype
        self.delta_dtype = delta_dtype
        self.correlation = correlation
        dev = device if device is not None else torch.device("cuda")
        self.register_buffer("mean", torch.zeros(n_features, dtype=torch.float32, device=dev))
        self.register_buffer("covmat_aggregate", torch.zeros((n_features, n_features), dtype=torch.float32,
                                                             device=dev))
        self.register_buffer("obs", torch.tensor([0], dtype=torch.int64))

    def forward(self, x) -> None:
        """Merge a batch: a (B, n_features) float tensor on the GPU, or a
        float32 PatternMatrix with n_pix == n_features."""
        if not isinstance(x, kernels.PatternMatrix):
            x = x.to(self.mean.device, torch.float32)
            if x.dim() != 2 or x.shape[1] != self.n_features:
                raise ValueError(f"batch must be (B, {self.n_features})")
            x = kernels.prep_patterns(x, DTYPE_F32, zero_mean=False)
        if x.n_pix != self.n_features:
            raise ValueError("batch width does not match n_features")
        if x.rows == 0:
            return
        kernels.cov_update(x, self.mean, self.covmat_aggregate, int(self.obs.item()))
        self.obs += x.rows

    def _full_aggregate(self) -> torch.Tensor:
        c = self.covmat_aggregate
        return torch.triu(c) + torch.triu(c, diagonal=1).T

    def get_covmat(self) -> torch.Tensor:
        """Unbiased covariance (or correlation) matrix (pca_covmat.py:133-147)."""
        cov = self._full_aggregate() / float(int(self.obs.item()) - 1)
        if self.correlation:
            d = 1.0 / torch.sqrt(torch.diag(cov))
            return cov * d[:, None] * d[None, :]
        return cov

    def get_eigenvectors(self, dtype: torch.dtype = torch.float64) -> torch.Tensor:
        """Eigenvectors as columns, ascending eigenvalues (pca_covmat.py:149-158);
        solved in FP64 on the host (fp64_eigh)."""
        return fp64_eigh(self.get_covmat(), dtype)[1]

    def get_eigensystem(self, dtype: torch.dtype = torch.float64):
        """(eigenvalues ascending, eigenvectors as columns), FP64 host solve."""
        return fp64_eigh(self.get_covmat(), dtype)
