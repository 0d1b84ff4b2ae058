"""CPU oracle of the learned-proposal form.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker; the
product package never imports it.

The reference has no network, so these are restatements of the form's own
semantics on top of the reference's supports:
  * zh2om (orientations.py:940-963): 6D -> rotation by Gram-Schmidt, rows
    b1, b2, b3 = b1 x b2;
  * laue_elements (laue_generators.py:54-249): the symmetry rotations of the
    loss, taken from the DI oracle;
  * get_local_orientation_grid + qu_prod (ebsd/refining.py:13-68, :114):
    candidates g_j * c of the local dictionary;
  * the DI oracle's projection (qu_apply, Lambert, grid_sample) for the
    candidate patterns.

How each part is pinned (tests/test_learned_oracle.py):
  * philox4x32_10 ........ Random123 known-answer vectors (Salmon et al.)
  * noise ................ Poisson mean/variance, chi-square of small-rate
                           counts against the exact pmf, Gaussian moments
  * rot6d_matrix ......... orthonormality + det +1; a rotation's first two
                           rows map back to it; closed form for identity
  * geodesic_sym_loss .... closed form (rotation by a known angle about an
                           axis), invariance under G on the left of the
                           truth, gradient against float64 finite
                           differences (torch.autograd.gradcheck)
  * net_forward .......... torch's conv2d / linear (library routines); the
                           convolution's stride/padding against a scalar
                           brute-force loop on a tiny input
  * local_refine ......... the DI oracle's projection (pinned by the
                           reference's golden vectors) + float64 cosine;
                           brute force of the top-k selection
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Tuple

import numpy as np
import torch
import torch.nn.functional as F

from . import di_oracle as dio

# --------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11)
# --------------------------------------------------------------------------

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key: Tuple[int, int]) -> np.ndarray:
    """ctr (n, 4) uint32 counters, key (k0, k1) -> (n, 4) uint32."""
    c = np.asarray(ctr, dtype=np.uint32).astype(np.uint64)
    k0, k1 = np.uint64(key[0] & 0xFFFFFFFF), np.uint64(key[1] & 0xFFFFFFFF)
    x0, x1, x2, x3 = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    for _ in range(10):
        p0 = _M0 * x0
        p1 = _M1 * x2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        x0, x1, x2, x3 = (hi1 ^ x1 ^ k0) & _MASK, lo1, (hi0 ^ x3 ^ k1) & _MASK, lo0
        k0 = (k0 + np.uint64(_W0)) & _MASK
        k1 = (k1 + np.uint64(_W1)) & _MASK
    return np.stack([x0, x1, x2, x3], 1).astype(np.uint32)


def _u01_open0(v):
    return (((v >> np.uint32(8)).astype(np.float32) + np.float32(1.0)) * np.float32(1.0 / 16777216.0)).astype(
        np.float32)


def _u01(v):
    return ((v >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)).astype(np.float32)


def _poisson_scalar(lam: np.float32, draw) -> int:
    """Per-element restatement of the kernel's sampler: multiplication of
    uniforms below rate 10, Hormann's PTRS above; `draw()` returns the next
    4 uint32 of the element's counter stream."""
    f32 = np.float32
    if not lam > 0:
        return 0
    if lam < 10.0:
        lim = f32(np.exp(-lam, dtype=np.float32))
        prod = f32(1.0)
        k = -1
        while True:
            r = draw()
            for v in r:
                prod = f32(prod * _u01_open0(np.uint32(v)))
                k += 1
                if prod <= lim:
                    return k
    slam = f32(np.sqrt(lam, dtype=np.float32))
    loglam = f32(np.log(lam, dtype=np.float32))
    b = f32(f32(0.931) + f32(2.53) * slam)
    a = f32(f32(-0.059) + f32(0.02483) * b)
    inv_alpha = f32(f32(1.1239) + f32(1.1328) / f32(b - f32(3.4)))
    vr = f32(f32(0.9277) - f32(3.6224) / f32(b - f32(2.0)))
    for _ in range(64):
        r = draw()
        U = f32(_u01(np.uint32(r[0])) - f32(0.5))
        V = _u01_open0(np.uint32(r[1]))
        us = f32(f32(0.5) - abs(U))
        kf = f32(np.floor(f32(f32(f32(f32(2.0) * a) / us + b) * U + lam + f32(0.43))))
        if us >= 0.07 and V <= vr:
            return int(kf)
        if kf < 0 or (us < 0.013 and V > us):
            continue
        lhs = f32(np.log(V, dtype=np.float32) + np.log(inv_alpha, dtype=np.float32)
                  - np.log(f32(a / f32(us * us) + b), dtype=np.float32))
        rhs = f32(-lam + kf * loglam - f32(math.lgamma(float(kf) + 1.0)))
        if lhs <= rhs:
            return int(kf)
    return int(np.floor(lam + 0.5))


def poisson_gaussian_noise(clean: np.ndarray, mean_counts: np.ndarray, sigma: float, seed: int,
                           offset: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """(counts (n, d) int, out (n, d) float32) of the detector-noise model
    for clean (n, d) float32 patterns, counters (pixel, row, draw, offset),
    key (seed lo, seed hi); the Gaussian uses draw 0, Poisson draws 1, 2, ...
    A mean count <= 0 leaves out the Poisson stage (out = clean + Gaussian)."""
    clean = np.asarray(clean, dtype=np.float32)
    n, d = clean.shape
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    off = offset & 0xFFFFFFFF
    counts = np.zeros((n, d), dtype=np.int64)
    out = np.zeros((n, d), dtype=np.float32)
    for i in range(n):
        row = clean[i]
        mean = np.float32(row.sum(dtype=np.float64) / d)
        c = np.float32(mean_counts[i])
        poisson_on = c > 0
        scale = np.float32(c / mean) if (poisson_on and mean > 0) else np.float32(0)
        for j in range(d):
            state = {"draw": 1}

            def draw():
                r = philox4x32_10(np.array([[j, i, state["draw"], off]], dtype=np.uint32), (k0, k1))[0]
                state["draw"] += 1
                return r

            g = philox4x32_10(np.array([[j, i, 0, off]], dtype=np.uint32), (k0, k1))[0]
            rr = math.sqrt(-2.0 * math.log(float(_u01_open0(np.uint32(g[0])))))
            z = rr * math.cos(2.0 * math.pi * float(_u01(np.uint32(g[1]))))
            if poisson_on:
                k = _poisson_scalar(np.float32(row[j] * scale), draw)
                counts[i, j] = k
                out[i, j] = np.float32(k / c + sigma * z)
            else:
                out[i, j] = np.float32(row[j] + sigma * z)
    return counts, out


# --------------------------------------------------------------------------
# 6D rotation and the symmetric geodesic loss
# --------------------------------------------------------------------------


def rot6d_matrix(z6: torch.Tensor) -> torch.Tensor:
    """(B, 6) -> (B, 3, 3) with rows b1, b2, b3 (zh2om, orientations.py:940-963)."""
    a1, a2 = z6[:, :3], z6[:, 3:6]
    b1 = a1 / (a1.norm(dim=1, keepdim=True) + 1e-12)
    u2 = a2 - (b1 * a2).sum(1, keepdim=True) * b1
    b2 = u2 / (u2.norm(dim=1, keepdim=True) + 1e-12)
    b3 = torch.cross(b1, b2, dim=1)
    return torch.stack([b1, b2, b3], 1)


def geodesic_sym_loss(z6: torch.Tensor, q_true: torch.Tensor, ops: torch.Tensor,
                      clamp_eps: float = 1e-6) -> torch.Tensor:
    """Per-sample min over symmetry rotations G of the geodesic angle between
    R(z6) and G R(q_true) (acos of the clamped (trace - 1) / 2)."""
    R = rot6d_matrix(z6)
    Rt = dio.qu2om(q_true.to(z6.dtype))
    G = dio.qu2om(ops.to(z6.dtype))
    T = torch.einsum("gij,bjk->bgik", G, Rt)
    tr = (R[:, None] * T).sum((-1, -2))
    c = ((tr - 1) / 2).clamp(-1 + clamp_eps, 1 - clamp_eps)
    return torch.acos(c).min(1).values


# --------------------------------------------------------------------------
# Network
# --------------------------------------------------------------------------


def net_forward(params: Dict[str, torch.Tensor], x: torch.Tensor, shape: Tuple[int, int],
                n_conv: int = 4) -> torch.Tensor:
    """x (B, H*W) zero-mean unit-norm -> z (B, 6): input scaled by sqrt(H W),
    n_conv x [conv 3x3 / stride 2 / padding 1 + ReLU], NHWC flatten,
    fc1 + ReLU, fc2."""
    H, W = shape
    h = x.view(-1, 1, H, W) * math.sqrt(H * W)
    for l in range(n_conv):
        h = F.relu(F.conv2d(h, params[f"conv{l}.w"], params[f"conv{l}.b"], stride=2, padding=1))
    h = h.permute(0, 2, 3, 1).reshape(h.shape[0], -1)
    h = F.relu(F.linear(h, params["fc1.w"], params["fc1.b"]))
    return F.linear(h, params["fc2.w"], params["fc2.b"])


def net_loss_and_grads(params: Dict[str, torch.Tensor], x: torch.Tensor, q_true: torch.Tensor,
                       ops: torch.Tensor, shape: Tuple[int, int], dtype=torch.float64):
    """(mean loss, z, {name: gradient}) by CPU autograd in `dtype`."""
    p = {k: v.detach().to(dtype).clone().requires_grad_(True) for k, v in params.items()}
    z = net_forward(p, x.to(dtype), shape)
    loss = geodesic_sym_loss(z, q_true.to(dtype), ops.to(dtype)).mean()
    loss.backward()
    return float(loss.detach()), z.detach(), {k: v.grad.detach() for k, v in p.items()}


def conv_bruteforce(x: torch.Tensor, w: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Scalar loops: 3x3 / stride 2 / padding 1 convolution of NCHW x."""
    B, C, H, W = x.shape
    Co = w.shape[0]
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    out = torch.zeros(B, Co, Ho, Wo, dtype=x.dtype)
    for n in range(B):
        for o in range(Co):
            for oy in range(Ho):
                for ox in range(Wo):
                    s = float(b[o])
                    for c in range(C):
                        for ky in range(3):
                            for kx in range(3):
                                y, xx = 2 * oy + ky - 1, 2 * ox + kx - 1
                                if 0 <= y < H and 0 <= xx < W:
                                    s += float(w[o, c, ky, kx]) * float(x[n, c, y, xx])
                    out[n, o, oy, ox] = s
    return out


# --------------------------------------------------------------------------
# Local-dictionary refinement
# --------------------------------------------------------------------------


def local_scores(mp: torch.Tensor, rays: torch.Tensor, x: torch.Tensor, centers: torch.Tensor,
                 grid: torch.Tensor) -> torch.Tensor:
    """(B, G) float64 cosines between each query x[i] (zero-mean unit-norm
    float) and the zero-mean projections of qu_prod(grid[j], centers[i])."""
    mp64 = mp.double()
    rays64 = rays.double()
    out = torch.empty(x.shape[0], grid.shape[0], dtype=torch.float64)
    for i in range(x.shape[0]):
        q = grid.double()
        c = centers[i].double()
        w = q[:, 0] * c[0] - q[:, 1] * c[1] - q[:, 2] * c[2] - q[:, 3] * c[3]
        xx = q[:, 0] * c[1] + q[:, 1] * c[0] + q[:, 2] * c[3] - q[:, 3] * c[2]
        yy = q[:, 0] * c[2] - q[:, 1] * c[3] + q[:, 2] * c[0] + q[:, 3] * c[1]
        zz = q[:, 0] * c[3] + q[:, 1] * c[2] - q[:, 2] * c[1] + q[:, 3] * c[0]
        cq = torch.stack([w, xx, yy, zz], 1)
        cq = cq / cq.norm(dim=1, keepdim=True)
        R = dio.qu2om(cq)                                   # (G, 3, 3)
        v = torch.einsum("gij,dj->gdi", R, rays64)
        p = dio.interpolate(mp64, v, normalize_coords=False)  # (G, D)
        p = p - p.mean(1, keepdim=True)
        xi = x[i].double()
        out[i] = (p @ xi) / (p.norm(dim=1) * xi.norm())
    return out


def topk_total_order(scores: torch.Tensor, k: int) -> Tuple[torch.Tensor, torch.Tensor]:
    """Best k per row by (higher score, then lower index), brute force."""
    vals, idxs = [], []
    for row in scores:
        order = sorted(range(row.numel()), key=lambda j: (-float(row[j]), j))[:k]
        idxs.append(order)
        vals.append([float(row[j]) for j in order])
    return torch.tensor(vals, dtype=scores.dtype), torch.tensor(idxs, dtype=torch.int64)
