"""Pin the NLPAR oracle (preprocessing/nlpar.py restated) on the CPU:
distances and averages against scalar loops, the window order, the
temperature fit against its defining equation, the softmax invariants, and
denoising of a noisy piecewise-constant scan."""

import math

import pytest
import torch

from oracle import nlpar_oracle as N


def test_window_order():
    offs = N.window_offsets(2)
    assert offs.shape == (25, 2)
    assert offs[0].tolist() == [0, 0]
    l1 = offs.abs().sum(-1)
    assert bool((l1[1:] >= l1[:-1]).all())
    assert sorted(map(tuple, offs[1:5].tolist())) == [(-1, 0), (0, -1), (0, 1), (1, 0)]


@pytest.mark.parametrize("k", [1, 2])
def test_distances_and_average_match_scalar_loops(k):
    g = torch.Generator().manual_seed(k)
    hs, ws, h, w = 3, 4, 3, 2
    x = torch.rand(hs, ws, h, w, generator=g)
    d, outside = N.distances(x, k)
    offs = N.window_offsets(k).tolist()
    wts = torch.softmax(torch.randn(hs, ws, len(offs), generator=g), -1)
    avg = N.average(x, wts, k)
    for i in range(hs):
        for j in range(ws):
            acc = torch.zeros(h, w, dtype=torch.float64)
            for o, (dy, dx) in enumerate(offs):
                yi, xj = i + dy, j + dx
                inside = 0 <= yi < hs and 0 <= xj < ws
                nb = x[yi, xj].double() if inside else torch.zeros(h, w, dtype=torch.float64)
                assert bool(outside[i, j, o]) == (not inside)
                assert abs(float(d[i, j, o]) - float(((nb - x[i, j].double()) ** 2).sum())) < 1e-5
                acc += nb * float(wts[i, j, o])
            assert torch.allclose(avg[i, j].double(), acc, atol=1e-6)


def _scene(hs=12, ws=10, h=8, w=8, sigma=0.1, seed=0):
    """A noisy scan of three grains (piecewise-constant clean patterns)."""
    g = torch.Generator().manual_seed(seed)
    grains = torch.rand(3, h, w, generator=g)
    yy, xx = torch.meshgrid(torch.arange(hs), torch.arange(ws), indexing="ij")
    label = (xx >= ws // 2).long() + (yy >= hs // 2).long() * (xx < ws // 2).long() * 2
    clean = grains[label]
    return clean, clean + sigma * torch.randn(clean.shape, generator=g)


def test_weights_fit_and_invariants():
    clean, noisy = _scene()
    k = 2
    d, outside = N.distances(noisy, k)
    w = N.weights(d, outside, 64, k, coeff=0.375)
    assert torch.allclose(w.sum(-1), torch.ones(w.shape[:2]), atol=1e-5)
    assert float(w[outside].abs().max()) == 0.0
    # the fitted temperature satisfies its equation on the 3x3 window
    dd = d.clone()
    sig = torch.sqrt(dd[:, :, 1:9].min(-1).values / 128)
    sig_p = torch.nn.functional.pad(sig, (k, k, k, k), value=1.0)
    for i, (dy, dx) in enumerate(N.window_offsets(k).tolist()):
        s2 = sig ** 2 + sig_p[k + dy:k + dy + 12, k + dx:k + dx + 10] ** 2
        dd[:, :, i] = (dd[:, :, i] - 64 * s2) / (math.sqrt(2) * 8 * s2)
    nn_w = -dd.clamp_min(0.0)[:, :, :9]
    t = N.fit_temperature(nn_w, 0.375, 1e-7)
    assert abs(float(torch.softmax(nn_w / t ** 2, -1)[..., 0].mean()) - 0.375) < 1e-4


def test_nlpar_denoises_a_grain_scan():
    clean, noisy = _scene()
    out = N.nlpar(noisy, k_rad=2)
    assert float(((out - clean) ** 2).mean()) < 0.5 * float(((noisy - clean) ** 2).mean())
