"""Pin the CLAHE / normalisation oracle on the CPU."""

import math

import pytest
import torch

from oracle import preprocess_oracle as PP


def test_tile_cdfs_conserve_counts_and_end_at_one():
    g = torch.Generator().manual_seed(0)
    x = torch.rand(3, 30, 26, generator=g) ** 3  # skewed histograms, clipping active
    c = PP.tile_cdfs(x, clip_limit=4.0, n_bins=16, grid_shape=(3, 2))
    assert c.shape == (3, 3, 2, 16)
    assert bool((c[..., 1:] >= c[..., :-1]).all())
    assert torch.allclose(c[..., -1], torch.ones(3, 3, 2))
    # the clipped histogram never exceeds limit + redistribution + 1
    hist = torch.diff(torch.cat((torch.zeros(3, 3, 2, 1), c), -1), dim=-1) * 130
    limit = max(4.0 * 130 // 16, 1)
    assert float(hist.max()) <= limit + 130 / 16 + 1 + 1e-3


def test_global_equalisation_is_a_monotone_map():
    g = torch.Generator().manual_seed(1)
    x = torch.rand(2, 17, 23, generator=g)
    y = PP.clahe_grayscale(x, clip_limit=0.0, n_bins=32, grid_shape=(1, 1))
    for i in range(2):
        order = torch.argsort(x[i].reshape(-1))
        assert bool((torch.diff(y[i].reshape(-1)[order]) >= -1e-6).all())
    # roughly uniform output for uniform input
    assert abs(float(y.mean()) - 0.5) < 0.05


def test_lookup_matches_scalar_trilinear():
    g = torch.Generator().manual_seed(2)
    x = torch.rand(1, 12, 10, generator=g)
    gh, gw, nb = 3, 2, 8
    cdf = PP.tile_cdfs(x, 40.0, nb, (gh, gw))[0]
    y = PP.clahe_grayscale(x, 40.0, nb, (gh, gw))[0]

    def src(c, n):
        return min(max(((c + 1) * n - 1) / 2, 0.0), n - 1)

    for r in range(12):
        for col in range(10):
            fz = src(2 * float(x[0, r, col]) - 1, nb)
            fx = src(-1 + 2 * col / 9, gw)
            fy = src(-1 + 2 * r / 11, gh)
            acc = 0.0
            for dy, wy in ((0, 1 - (fy - math.floor(fy))), (1, fy - math.floor(fy))):
                for dx, wx in ((0, 1 - (fx - math.floor(fx))), (1, fx - math.floor(fx))):
                    for dz, wz in ((0, 1 - (fz - math.floor(fz))), (1, fz - math.floor(fz))):
                        yy = min(math.floor(fy) + dy, gh - 1)
                        xx = min(math.floor(fx) + dx, gw - 1)
                        zz = min(math.floor(fz) + dz, nb - 1)
                        acc += wy * wx * wz * float(cdf[yy, xx, zz])
            assert abs(float(y[r, col]) - acc) < 1e-5


@pytest.mark.parametrize("mode", ["minmax", "zeromean", "standard"])
def test_normalize_per_pattern(mode):
    g = torch.Generator().manual_seed(3)
    x = 5 * torch.rand(4, 50, generator=g) + 2
    y = PP.normalize_per_pattern(x, mode)
    if mode == "minmax":
        assert float(y.min()) == 0.0 and float(y.max()) < 1.0 and float(y.max()) > 0.9999
    else:
        assert float(y.mean(-1).abs().max()) < 1e-5
    if mode == "standard":
        assert torch.allclose(y.std(-1), torch.ones(4), atol=1e-5)
