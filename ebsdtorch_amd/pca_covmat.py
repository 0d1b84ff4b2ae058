"""Streaming covariance matrix on the GPU (utils/pca_covmat.py:94-158).

Same interface as ebsdtorch's OnlineCovMatrix: call the object on batches
(B, n_features), then get_covmat() / get_eigenvectors(). Batches are merged
by the HIP kernel ebsd_cov_update (Chan's batched Welford update, FP32
accumulation on the matrix cores); eigenvectors come from an FP64 symmetric
eigensolver on the host (fp64_eigh), ascending eigenvalues as in the
reference. leading_eigh finds only the leading eigenpairs, by subspace
iteration with the products on the device (the basis of the
principal-coordinate search).
"""

from __future__ import annotations

import torch

from . import kernels
from ._lib import DTYPE_F32


def fp64_eigh(matrix: torch.Tensor, dtype: torch.dtype = torch.float64):
    """(eigenvalues ascending, eigenvectors as columns) of a symmetric
    matrix, solved in float64 (or `dtype`) by the host's LAPACK
    (torch.linalg.eigh on the CPU) and returned on the matrix's device.

    Every symmetric eigensolve of this package goes through here (the full
    solves of OnlineCovMatrix, and the small Rayleigh-Ritz problems of
    leading_eigh). The device solver (hipSOLVER syevd through torch) is not
    used: one profiled run (rocprofv3 with hardware counters) ended with a
    segmentation fault inside hipsolverDsyevd on a 3,600 x 3,600 matrix, with
    no cause established, so the solve stays on the host at every size. A
    full solve costs about 9 D^3 flops: ~1 s at D = 3,600, ~3-8 s at D =
    6,400 and about half a minute at D = 14,400 on 8-16 host cores."""
    dev = matrix.device
    evals, evecs = torch.linalg.eigh(matrix.detach().to("cpu", dtype))
    return evals.to(dev), evecs.to(dev)


def _orthonormal_columns(y: torch.Tensor) -> torch.Tensor:
    """An orthonormal basis of the column span of y (n, p) float64, n >= p:
    Cholesky QR applied twice (the products on y's device, the p x p
    Cholesky factor and its inverse on the host); when the Gram matrix is
    not numerically positive definite, the host's Householder QR."""
    q = y
    for _ in range(2):
        gram = (q.T @ q).cpu()
        chol, info = torch.linalg.cholesky_ex(gram)
        if int(info) != 0:
            return torch.linalg.qr(y.cpu())[0].to(y.device)
        inv = torch.linalg.solve_triangular(chol, torch.eye(gram.shape[0], dtype=gram.dtype), upper=False)
        q = q @ inv.T.to(y.device)
    return q


def leading_eigh(matrix: torch.Tensor, k: int, oversample: int = 64, iters: int = 6, seed: int = 0x5eed):
    """The k largest eigenpairs of a symmetric positive semi-definite matrix
    (eigenvalues ascending, eigenvectors as columns, as fp64_eigh) by
    randomized subspace iteration in float64: Y = M Omega (a seeded Gaussian
    D x (k + oversample) block), then `iters` rounds of orthonormalisation
    and Y = M Q, and a Rayleigh-Ritz step on Q^T M Q solved by fp64_eigh.
    The products run on the matrix's device (plain FP64 GEMMs), only
    (k + oversample)-sized problems on the host, so the cost is
    ~2 (iters + 2) D^2 (k + oversample) flops instead of a full D^3 solve.
    Each eigenvector's error shrinks like (lambda_{p+1} / lambda_i)^iters,
    p = k + oversample. Small matrices (p above D / 2) are solved exactly."""
    d = matrix.shape[0]
    p = min(d, k + oversample)
    if 2 * p > d:
        evals, evecs = fp64_eigh(matrix)
        return evals[-k:], evecs[:, -k:]
    m = matrix.detach().to(torch.float64)
    g = torch.Generator().manual_seed(seed)
    y = m @ torch.randn((d, p), generator=g, dtype=torch.float64).to(m.device)
    for _ in range(iters):
        y = m @ _orthonormal_columns(y)
    q = _orthonormal_columns(y)
    b = q.T @ (m @ q)
    evals, v = fp64_eigh(0.5 * (b + b.T))
    return evals[-k:], (q @ v)[:, -k:]


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
