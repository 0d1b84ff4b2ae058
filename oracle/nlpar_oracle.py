"""CPU oracle of ebsdtorch's NLPAR pattern denoising (preprocessing/nlpar.py).

TEST INFRASTRUCTURE ONLY. Only tests/ and bench.py's CPU-baseline legs may
import this module, and only as the checker (or, in the bench, as the timed
CPU baseline); the product package never imports it.

The restatement follows nlpar.py:20-192 step by step with torch CPU float32
operations: zero-padded scan, squared distances to each window offset in the
reference's offset order, the noise estimate from the 8 nearest offsets, the
normalised distances, the Newton fit of the softmax temperature on the 3x3
window, the masked softmax and the weighted average. Pinned in
tests/test_oracle_nlpar.py against scalar loops (distances, averages), the
fit's defining equation and the softmax's invariants.
"""

from __future__ import annotations

import math

import torch
import torch.nn.functional as F


def window_offsets(k_rad: int) -> torch.Tensor:
    """nlpar.py:83-98: meshgrid 'ij' of the window, sorted by L1 length with
    torch.argsort (index 0 = (0, 0), 1..8 = the eight nearest)."""
    r = torch.arange(-k_rad, k_rad + 1)
    offs = torch.stack(torch.meshgrid(r, r, indexing="ij"), dim=-1).reshape(-1, 2)
    return offs[torch.argsort(offs.abs().sum(-1))]


def _shifted(t: torch.Tensor, k: int, dy: int, dx: int, hs: int, ws: int) -> torch.Tensor:
    return t[k + dy:k + dy + hs, k + dx:k + dx + ws]


def distances(data: torch.Tensor, k_rad: int):
    """nlpar.py:47-117: (squared distances (Hs, Ws, n), outside-scan mask)."""
    hs, ws = data.shape[:2]
    k = k_rad
    padded = F.pad(data, (0, 0, 0, 0, k, k, k, k))
    inside = F.pad(torch.ones((hs, ws), dtype=torch.bool), (k, k, k, k), value=False)
    offs = window_offsets(k)
    centre = padded[k:k + hs, k:k + ws]
    d = torch.empty((hs, ws, len(offs)), dtype=torch.float32)
    outside = torch.empty((hs, ws, len(offs)), dtype=torch.bool)
    for i, (dy, dx) in enumerate(offs.tolist()):
        d[:, :, i] = ((_shifted(padded, k, dy, dx, hs, ws) - centre) ** 2).sum(dim=(-1, -2))
        outside[:, :, i] = ~_shifted(inside, k, dy, dx, hs, ws)
    return d, outside


def fit_temperature(nn_w: torch.Tensor, coeff: float, tol: float) -> torch.Tensor:
    """nlpar.py:140-171: Newton's method on t so that the mean over points of
    softmax(nn_w / t^2)[..., 0] equals coeff (10 steps at most)."""
    t = 1.0
    for _ in range(10):
        t2 = t ** 2
        ex = torch.exp(nn_w / t2)
        f = torch.softmax(nn_w / t2, dim=-1)[..., 0].mean() - coeff
        df = (-2 * torch.exp(nn_w[..., 0] / t2)
              * ((nn_w[..., :1] * ex).sum(-1) - (nn_w * ex).sum(-1))
              / (t ** 3 * ex.sum(-1) ** 2)).mean()
        t = t - f / df
        if torch.abs(f) < tol:
            break
    return t


def weights(d: torch.Tensor, outside: torch.Tensor, n_pix: int, k_rad: int, coeff: float = 0.375,
            tol: float = 1e-6, center_logit_bias: float = 0.0) -> torch.Tensor:
    """nlpar.py:119-184: softmax weights (Hs, Ws, n) from the distances."""
    hs, ws = d.shape[:2]
    k = k_rad
    d = d.clone()
    sig = torch.sqrt(d[:, :, 1:9].min(-1).values / (2 * n_pix))
    sig_p = F.pad(sig, (k, k, k, k), value=1.0)
    for i, (dy, dx) in enumerate(window_offsets(k).tolist()):
        s2 = sig ** 2 + _shifted(sig_p, k, dy, dx, hs, ws) ** 2
        d[:, :, i] = (d[:, :, i] - n_pix * s2) / (math.sqrt(2) * math.sqrt(n_pix) * s2)
    w = -d.clamp_min(0.0)
    t = fit_temperature(w[:, :, :9], coeff, tol)
    w = w / t ** 2
    if center_logit_bias > 0:
        w[:, :, 0] += center_logit_bias
    w = w.masked_fill(outside, -torch.inf)
    return torch.softmax(w, dim=-1)


def average(data: torch.Tensor, w: torch.Tensor, k_rad: int) -> torch.Tensor:
    """nlpar.py:186-191: sum over the window, in offset order, of the
    (zero-padded) neighbour times its weight."""
    hs, ws = data.shape[:2]
    k = k_rad
    padded = F.pad(data, (0, 0, 0, 0, k, k, k, k))
    out = torch.zeros_like(data)
    for i, (dy, dx) in enumerate(window_offsets(k).tolist()):
        out += _shifted(padded, k, dy, dx, hs, ws) * w[:, :, i, None, None]
    return out


def nlpar(data: torch.Tensor, k_rad: int = 5, coeff: float = 0.375, coeff_fit_tol: float = 1e-6,
          center_logit_bias: float = 0.0) -> torch.Tensor:
    """nlpar.py:20-192 for a scan (Hs, Ws, h, w) float32 on the CPU."""
    d, outside = distances(data, k_rad)
    w = weights(d, outside, data.shape[2] * data.shape[3], k_rad, coeff, coeff_fit_tol, center_logit_bias)
    return average(data, w, k_rad)
