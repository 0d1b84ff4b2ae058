"""Leading eigenpairs by subspace iteration (pca_covmat.leading_eigh) against
the full FP64 solve, on the host: the subspace captures the optimal share
of the trace, the eigenvalues match, the basis is orthonormal; small and
rank-deficient matrices take the exact and the Householder paths."""

import torch

from ebsdtorch_amd.pca_covmat import fp64_eigh, leading_eigh


def _spd(d, decay, seed, rank=None):
    g = torch.Generator().manual_seed(seed)
    lam = 1.0 / (1.0 + torch.arange(d, dtype=torch.float64)) ** decay
    if rank is not None:
        lam[rank:] = 0.0
    u, _ = torch.linalg.qr(torch.randn(d, d, generator=g, dtype=torch.float64))
    return (u * lam) @ u.T


def _captured(m, b):
    return float(torch.trace(b.T @ m @ b) / torch.trace(m))


def test_leading_subspace_matches_the_full_solve():
    for d, k, decay in ((1200, 128, 1.3), (1500, 256, 2.0)):
        m = _spd(d, decay, seed=d)
        w, v = leading_eigh(m, k)
        we, ve = fp64_eigh(m)
        assert w.shape == (k,) and v.shape == (d, k)
        assert torch.all(w[1:] >= w[:-1])  # ascending, as fp64_eigh
        assert abs(_captured(m, v) - _captured(m, ve[:, -k:])) < 1e-4 * _captured(m, ve[:, -k:])
        assert float((w - we[-k:]).abs().max()) < 1e-5 * float(we[-1])
        assert float((v.T @ v - torch.eye(k, dtype=torch.float64)).abs().max()) < 1e-10
        # the leading eigenvectors themselves (well separated at the top)
        assert float((v[:, -8:].T @ ve[:, -8:]).abs().diagonal().min()) > 1 - 1e-8


def test_small_matrix_is_solved_exactly():
    m = _spd(200, 1.0, seed=3)
    w, v = leading_eigh(m, 150)
    we, ve = fp64_eigh(m)
    assert torch.equal(w, we[-150:]) and torch.equal(v, ve[:, -150:])


def test_rank_deficient_matrix():
    """Rank 40 < the subspace size: the Gram matrix is singular, the
    orthonormalisation falls back to Householder QR; the 40 nonzero
    eigenpairs are found."""
    m = _spd(1000, 1.0, seed=5, rank=40)
    w, v = leading_eigh(m, 64)
    we, _ = fp64_eigh(m)
    assert float((w[-40:] - we[-40:]).abs().max()) < 1e-10
    assert abs(_captured(m, v) - 1.0) < 1e-10
