"""Learned orientation proposal with local-dictionary refinement.

A small network (4 convolutions 3x3 / stride 2 / padding 1 with ReLU, then a
2-layer MLP) maps a zero-mean unit-norm pattern to a 6D rotation
representation (two rows of a rotation matrix, completed by Gram-Schmidt as
orientations.py:940-963 zh2om). It is trained only on patterns simulated
from the master pattern on the GPU -- uniform random orientations, a
jittered pattern centre, Poisson + Gaussian detector noise -- with the
symmetric geodesic loss min_G angle(R, G R_true) over the Laue group's proper
rotations, so any of the symmetry-equivalent answers is correct.

Every step runs on the library's kernels: projection with per-pattern
pattern-centre shifts, counter-based noise, im2col + MFMA GEMMs in BF16 (or
FP32) with FP32 accumulation, bias/ReLU, col2im, split-K weight gradients,
the loss with its analytic gradient, and Adam on FP32 master weights that
refreshes the BF16 operand copy in the same pass. Activations are NHWC
float32; torch only allocates memory and draws the per-step random
orientations, pattern-centre shifts and count levels.

Indexing (LearnedIndexer): proposal -> local dictionary around it (a ball
of ~2,000 orientations within 3 degrees, ebsd/refining.py:13-68's grid
composed on the left as ebsd/refining.py:114) scored on the fly by a fused
kernel -> the plain full-dictionary search for every pattern whose best
local cosine falls below a threshold or lands on the ball's outer shell.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import kernels
from ._lib import DTYPE_F32
from .orientations import laue_elements, local_ball_grid_quats, ori_to_fz_laue, qu_prod

COMPUTE_DTYPES = {"bf16": torch.bfloat16, "fp32": torch.float32}


def _pad(n: int, m: int) -> int:
    return ((n + m - 1) // m) * m


def _conv_out(n: int) -> int:
    return (n + 1) // 2  # 3x3, stride 2, padding 1


@dataclass
class _Param:
    name: str
    shape: Tuple[int, int]      # (rows, padded cols) of the stored matrix / (n,) bias
    offset: int
    numel: int


class ProposalNet:
    """Conv + MLP orientation proposal network on the library's kernels.

    detector_shape (H, W); channels of the 4 convolutions; hidden width of
    the MLP. compute "bf16" runs every GEMM on BF16 operands (FP32 master
    weights, FP32 accumulation and FP32 activations in between); "fp32" runs
    them on FP32 operands. Parameters live in one flat FP32 buffer so one
    Adam launch updates all of them.

    Weight layouts: convolution l is (C_out, K_pad) with column
    (ky * 3 + kx) * C_in + c (zero beyond 9 C_in); fc1 is (hidden, F) with F
    the NHWC flattening (y, x, c) of the last feature map; fc2 is (6, hidden).
    """

    def __init__(self, detector_shape: Tuple[int, int], device: torch.device, laue_id: int = 11,
                 compute: str = "bf16", channels: Sequence[int] = (32, 64, 128, 128), hidden: int = 512,
                 seed: int = 0):
        if compute not in COMPUTE_DTYPES:
            raise ValueError(f"compute must be one of {sorted(COMPUTE_DTYPES)}")
        if device.type != "cuda":
            raise ValueError("the proposal network runs on the GPU")
        self.device = device
        self.compute = compute
        self.cdtype = COMPUTE_DTYPES[compute]
        self.pad = kernels.pad_multiple(kernels._DT_OF[self.cdtype])
        self.H, self.W = int(detector_shape[0]), int(detector_shape[1])
        self.n_pix = self.H * self.W
        self.channels = tuple(int(c) for c in channels)
        self.hidden = int(hidden)
        # (H_in, W_in, C_in, H_out, W_out, C_out) per convolution
        self.conv_shapes = []
        h, w, c = self.H, self.W, 1
        for co in self.channels:
            self.conv_shapes.append((h, w, c, _conv_out(h), _conv_out(w), co))
            h, w, c = _conv_out(h), _conv_out(w), co
        self.feat = h * w * c
        if self.feat % self.pad or self.hidden % self.pad:
            raise ValueError("flattened features and hidden width must be multiples of the GEMM padding")
        # parameter table
        self.params: Dict[str, _Param] = {}
        off = 0

        def add(name, shape):
            nonlocal off
            numel = shape[0] * (shape[1] if len(shape) > 1 else 1)
            self.params[name] = _Param(name, shape, off, numel)
            off += _pad(numel, 64)

        for l, (_, _, ci, _, _, co) in enumerate(self.conv_shapes):
            add(f"conv{l}.w", (co, _pad(9 * ci, self.pad)))
            add(f"conv{l}.b", (co,))
        add("fc1.w", (self.hidden, self.feat))
        add("fc1.b", (self.hidden,))
        add("fc2.w", (6, self.hidden))
        add("fc2.b", (6,))
        self.n_flat = off
        self.flat = torch.zeros(off, dtype=torch.float32, device=device)
        self.grad = torch.zeros(off, dtype=torch.float32, device=device)
        self.adam_m = torch.zeros(off, dtype=torch.float32, device=device)
        self.adam_v = torch.zeros(off, dtype=torch.float32, device=device)
        self.flat_lp = torch.zeros(off, dtype=torch.bfloat16, device=device) if compute == "bf16" else None
        self.step_count = 0
        self._acts = self._h1 = None
        self.set_laue(laue_id)
        self._init_weights(seed)

    # -- parameters ---------------------------------------------------------
    def view(self, name: str, buf: Optional[torch.Tensor] = None) -> torch.Tensor:
        p = self.params[name]
        b = self.flat if buf is None else buf
        return b.narrow(0, p.offset, p.numel).view(p.shape)

    def weight_lp(self, name: str) -> torch.Tensor:
        """GEMM operand copy of a weight in the compute dtype."""
        return self.view(name, self.flat_lp if self.flat_lp is not None else self.flat)

    def _init_weights(self, seed: int):
        g = torch.Generator().manual_seed(seed)
        host = torch.zeros(self.n_flat, dtype=torch.float32)
        for l, (_, _, ci, _, _, co) in enumerate(self.conv_shapes):
            w = torch.randn(co, 9 * ci, generator=g) * math.sqrt(2.0 / (9 * ci))
            p = self.params[f"conv{l}.w"]
            host[p.offset:p.offset + p.numel].view(p.shape)[:, :9 * ci] = w
        p = self.params["fc1.w"]
        host[p.offset:p.offset + p.numel] = (torch.randn(p.shape, generator=g) * math.sqrt(2.0 / self.feat)).view(-1)
        p = self.params["fc2.w"]
        host[p.offset:p.offset + p.numel] = (torch.randn(p.shape, generator=g) * math.sqrt(1.0 / self.hidden)).view(-1)
        self.flat.copy_(host)
        self._refresh_lp()

    def _refresh_lp(self):
        if self.flat_lp is not None:
            self.flat_lp.copy_(self.flat)

    def set_laue(self, laue_id: int):
        """Symmetry group of the loss (proper rotations of the Laue group)."""
        self.laue_id = int(laue_id)
        self.ops = laue_elements(laue_id).to(device=self.device, dtype=torch.float32).contiguous()

    def state_dict(self) -> Dict[str, torch.Tensor]:
        """FP32 master weights, Adam moments and step (checkpoint)."""
        return {"flat": self.flat.detach().cpu().clone(), "adam_m": self.adam_m.cpu().clone(),
                "adam_v": self.adam_v.cpu().clone(), "step": torch.tensor(self.step_count)}

    def load_state_dict(self, sd: Dict[str, torch.Tensor]):
        if sd["flat"].numel() != self.n_flat:
            raise ValueError("checkpoint does not match this network's shape")
        self.flat.copy_(sd["flat"])
        self.adam_m.copy_(sd["adam_m"])
        self.adam_v.copy_(sd["adam_v"])
        self.step_count = int(sd["step"])
        self._refresh_lp()

    # -- forward / backward -------------------------------------------------
    def _operand(self, src: torch.Tensor, cols_pad: int) -> torch.Tensor:
        """src (rows, cols) f32 -> contiguous (rows, cols_pad) in the compute dtype."""
        if self.cdtype == torch.float32 and src.is_contiguous() and src.shape[1] == cols_pad:
            return src
        out = torch.empty((src.shape[0], cols_pad), dtype=self.cdtype, device=self.device)
        return kernels.cast_pad(src, out)

    def _transposed(self, src: torch.Tensor) -> torch.Tensor:
        """src (rows, cols) f32 -> (cols, pad(rows)) in the compute dtype."""
        out = torch.empty((src.shape[1], _pad(src.shape[0], self.pad)), dtype=self.cdtype, device=self.device)
        return kernels.transpose_pad(src, out)

    def _splits(self, k_len: int) -> int:
        stages = k_len * torch.tensor([], dtype=self.cdtype).element_size() // 128
        return max(1, min(256, stages // 16))

    def forward(self, x: torch.Tensor, keep: bool = False):
        """x (B, H*W) f32 zero-mean unit-norm patterns -> z (B, 6) f32. With
        keep=True the activations are kept for backward()."""
        if x.dim() != 2 or x.shape[1] != self.n_pix or x.dtype != torch.float32:
            raise ValueError(f"input must be float32 (B, {self.n_pix})")
        x = x.contiguous()
        B = x.shape[0]
        acts = [x.view(B, self.H, self.W, 1)]
        for l, (h, w, ci, ho, wo, co) in enumerate(self.conv_shapes):
            scale = math.sqrt(self.n_pix) if l == 0 else 1.0
            if ci == 1 and co % 4 == 0 and co <= 256:
                # one input channel: 9 taps per output, done directly
                y = kernels.conv_c1(acts[-1], self.weight_lp(f"conv{l}.w"), self.view(f"conv{l}.b"), scale)
            elif scale == 1.0 and ci % (4 if self.cdtype == torch.float32 else 8) == 0:
                # implicit GEMM: the im2col operand is built in LDS, never stored
                y = kernels.conv_gemm(acts[-1], self.weight_lp(f"conv{l}.w"), self.view(f"conv{l}.b"))
            else:
                kp = self.params[f"conv{l}.w"].shape[1]
                col = torch.empty((B * ho * wo, kp), dtype=self.cdtype, device=self.device)
                kernels.im2col(acts[-1], col, scale=scale)
                y = torch.empty((B * ho * wo, co), dtype=torch.float32, device=self.device)
                kernels.matmul_nt_bias_act(col, self.weight_lp(f"conv{l}.w"), y, self.view(f"conv{l}.b"), True)
                del col
            acts.append(y.view(B, ho, wo, co))
        flat = acts[-1].view(B, self.feat)
        h1 = torch.empty((B, self.hidden), dtype=torch.float32, device=self.device)
        kernels.matmul_nt_bias_act(self._operand(flat, self.feat), self.weight_lp("fc1.w"), h1,
                                   self.view("fc1.b"), True)
        z = torch.empty((B, 6), dtype=torch.float32, device=self.device)
        kernels.matmul_nt_bias_act(self._operand(h1, self.hidden), self.weight_lp("fc2.w"), z,
                                   self.view("fc2.b"), False)
        if keep:
            self._acts = acts
            self._h1 = h1
        return z

    def backward(self, dz: torch.Tensor):
        """Gradients of the loss w.r.t. every parameter into self.grad, from
        dz = d loss / d z (B, 6) and the activations of forward(keep=True)."""
        if self._acts is None:
            raise RuntimeError("backward() needs the activations of a forward(x, keep=True) call")
        acts, h1 = self._acts, self._h1
        B = dz.shape[0]
        if dz.shape != (acts[0].shape[0], 6) or dz.dtype != torch.float32 or not dz.is_contiguous():
            raise ValueError("dz must be contiguous float32 (B, 6) for the batch of the kept forward")
        g = self.grad
        # fc2
        kernels.matmul_nt(self._transposed(dz), self._transposed(h1), self.view("fc2.w", g),
                          splits=self._splits(_pad(B, self.pad)))
        kernels.colsum(dz, 6, self.view("fc2.b", g))
        p6 = _pad(6, self.pad)
        w2t = torch.empty((self.hidden, p6), dtype=self.cdtype, device=self.device)
        kernels.transpose_pad(self.view("fc2.w"), w2t)
        dh1 = torch.empty((B, self.hidden), dtype=torch.float32, device=self.device)
        kernels.matmul_nt(self._operand(dz, p6), w2t, dh1)
        kernels.relu_bwd(dh1, h1)
        # fc1
        flat = acts[-1].reshape(B, self.feat)
        kernels.matmul_nt(self._transposed(dh1), self._transposed(flat), self.view("fc1.w", g),
                          splits=self._splits(_pad(B, self.pad)))
        kernels.colsum(dh1, self.hidden, self.view("fc1.b", g))
        w1t = torch.empty((self.feat, self.hidden), dtype=self.cdtype, device=self.device)
        kernels.transpose_pad(self.view("fc1.w"), w1t)
        dx = torch.empty((B, self.feat), dtype=torch.float32, device=self.device)
        kernels.matmul_nt(self._operand(dh1, self.hidden), w1t, dx)
        kernels.relu_bwd(dx, flat)
        # convolutions, last to first
        for l in range(len(self.conv_shapes) - 1, -1, -1):
            h, w, ci, ho, wo, co = self.conv_shapes[l]
            rows = B * ho * wo
            dy = dx.view(rows, co)
            rp = _pad(rows, self.pad)
            colt = torch.empty((9 * ci, rp), dtype=self.cdtype, device=self.device)
            kernels.im2col(acts[l], colt, scale=math.sqrt(self.n_pix) if l == 0 else 1.0, transposed=True)
            kernels.matmul_nt(self._transposed(dy), colt, self.view(f"conv{l}.w", g), splits=self._splits(rp))
            del colt
            kernels.colsum(dy, co, self.view(f"conv{l}.b", g))
            if l == 0:
                break
            kp = self.params[f"conv{l}.w"].shape[1]
            cop = _pad(co, self.pad)
            wt = torch.empty((kp, cop), dtype=self.cdtype, device=self.device)
            kernels.transpose_pad(self.view(f"conv{l}.w"), wt)
            dcol = torch.empty((rows, kp), dtype=torch.float32, device=self.device)
            kernels.matmul_nt(self._operand(dy, cop), wt, dcol)
            dxi = kernels.col2im(dcol, (B, h, w, ci))
            del dcol
            kernels.relu_bwd(dxi, acts[l])
            dx = dxi.view(B, h * w * ci)
        self._acts = self._h1 = None

    def loss_and_grad(self, z: torch.Tensor, q_true: torch.Tensor):
        """Per-sample symmetric geodesic loss (B,) and d(mean)/dz (B, 6)."""
        B = z.shape[0]
        loss = torch.empty(B, dtype=torch.float32, device=self.device)
        dz = torch.empty((B, 6), dtype=torch.float32, device=self.device)
        kernels.geodesic_loss(z, q_true, self.ops, loss, dz)
        return loss, dz

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
             weight_decay: float = 0.0):
        self.step_count += 1
        kernels.adam_step(self.flat, self.grad, self.adam_m, self.adam_v, lr, beta1, beta2, eps, self.step_count,
                          weight_decay, self.flat_lp)

    def propose(self, x: torch.Tensor, chunk: int = 8192) -> torch.Tensor:
        """(B, H*W) zero-mean unit-norm patterns -> proposals (B, 4), w >= 0
        (in chunks of `chunk` patterns, which bounds the im2col buffers)."""
        out = torch.empty((x.shape[0], 4), dtype=torch.float32, device=self.device)
        for s in range(0, x.shape[0], chunk):
            out[s:s + chunk] = kernels.rot6d_to_quat(self.forward(x[s:s + chunk]))
        return out

    # -- reference layout ---------------------------------------------------
    def torch_parameters(self) -> Dict[str, torch.Tensor]:
        """The weights in torch's layouts (CPU float32): conv (C_out, C_in, 3,
        3), fc1 (hidden, F) over the NHWC flattening, fc2 (6, hidden)."""
        out = {}
        for l, (_, _, ci, _, _, co) in enumerate(self.conv_shapes):
            w = self.view(f"conv{l}.w")[:, :9 * ci].detach().cpu().view(co, 3, 3, ci)
            out[f"conv{l}.w"] = w.permute(0, 3, 1, 2).contiguous()
            out[f"conv{l}.b"] = self.view(f"conv{l}.b").detach().cpu().clone()
        for n in ("fc1.w", "fc1.b", "fc2.w", "fc2.b"):
            out[n] = self.view(n).detach().cpu().clone()
        return out

    def gradients_torch_layout(self) -> Dict[str, torch.Tensor]:
        out = {}
        for l, (_, _, ci, _, _, co) in enumerate(self.conv_shapes):
            gw = self.view(f"conv{l}.w", self.grad)[:, :9 * ci].detach().cpu().view(co, 3, 3, ci)
            out[f"conv{l}.w"] = gw.permute(0, 3, 1, 2).contiguous()
            out[f"conv{l}.b"] = self.view(f"conv{l}.b", self.grad).detach().cpu().clone()
        for n in ("fc1.w", "fc1.b", "fc2.w", "fc2.b"):
            out[n] = self.view(n, self.grad).detach().cpu().clone()
        return out


class PatternSimulator:
    """On-the-fly training data: uniform random orientations, a pattern
    centre jittered uniformly by +-pc_jitter per pattern, Poisson noise at a
    mean count drawn uniformly from mean_counts, Gaussian read noise sigma."""

    def __init__(self, master_pattern: torch.Tensor, geometry, mean_counts=(100.0, 400.0), sigma: float = 0.05,
                 pc_jitter: float = 0.01, seed: int = 0, device: Optional[torch.device] = None):
        self.device = device if device is not None else master_pattern.device
        self.mp = master_pattern.to(self.device, torch.float32).contiguous()
        self.coords = geometry.get_coords_sample_frame((1, 1)).to(self.device, torch.float32).contiguous()
        self.jac = geometry.pc_jacobian().to(self.device, torch.float32)
        self.n_pix = self.coords.shape[0]
        self.mean_counts = (float(mean_counts[0]), float(mean_counts[1]))
        self.sigma = float(sigma)
        self.pc_jitter = float(pc_jitter)
        self.seed = int(seed)
        self.gen = torch.Generator(device=self.device).manual_seed(self.seed)
        self.calls = 0

    def batch(self, n: int, quats: Optional[torch.Tensor] = None):
        """(patterns (n, D) f32 zero-mean unit-norm, true orientations (n, 4))."""
        dev = self.device
        if quats is None:
            q = torch.randn((n, 4), generator=self.gen, device=dev)
            q = q / q.norm(dim=1, keepdim=True)
            q = torch.where(q[:, :1] < 0, -q, q).contiguous()
        else:
            q = quats.to(dev, torch.float32).contiguous()
        dpc = (torch.rand((n, 3), generator=self.gen, device=dev) * 2.0 - 1.0) * self.pc_jitter
        offsets = (dpc @ self.jac.T).contiguous()
        lo, hi = self.mean_counts
        mc = (lo + (hi - lo) * torch.rand(n, generator=self.gen, device=dev)).contiguous()
        clean = kernels.project_shifted(self.mp, self.coords, q, offsets, DTYPE_F32, zero_mean=False)
        noisy = kernels.poisson_gaussian_noise(clean.data, self.n_pix, mc, self.sigma, seed=self.seed,
                                               offset=self.calls)
        self.calls += 1
        x = kernels.prep_patterns(noisy, DTYPE_F32, zero_mean=True, unit_norm=True)
        return x.data[:, :self.n_pix], q


@dataclass
class TrainLog:
    steps: int = 0
    patterns: int = 0
    seconds: float = 0.0
    loss_curve: List[Tuple[int, float]] = field(default_factory=list)


def train_proposal_net(net: ProposalNet, sim: PatternSimulator, n_patterns: int, batch: int = 512,
                       lr: float = 1e-3, lr_final: float = 1e-5, warmup: int = 100, log_every: int = 100) -> TrainLog:
    """Train on n_patterns freshly simulated patterns (cosine learning-rate
    decay from lr to lr_final after a linear warm-up)."""
    import time

    steps = max(1, n_patterns // batch)
    log = TrainLog()
    torch.cuda.synchronize(net.device)
    t0 = time.perf_counter()
    for s in range(steps):
        if s < warmup:
            cur = lr * (s + 1) / warmup
        else:
            t = (s - warmup) / max(1, steps - warmup)
            cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * t))
        x, q = sim.batch(batch)
        z = net.forward(x, keep=True)
        loss, dz = net.loss_and_grad(z, q)
        net.backward(dz)
        net.adam(cur)
        if log_every and (s % log_every == 0 or s == steps - 1):
            log.loss_curve.append((s, float(loss.mean())))
    torch.cuda.synchronize(net.device)
    log.seconds = time.perf_counter() - t0
    log.steps = steps
    log.patterns = steps * batch
    return log


def _morton3(ijk: torch.Tensor) -> torch.Tensor:
    """Interleaved-bit (Z-order) code of non-negative integer triples < 1024."""
    code = torch.zeros(ijk.shape[0], dtype=torch.int64)
    for b in range(10):
        for a in range(3):
            code |= ((ijk[:, a].long() >> b) & 1) << (3 * b + (2 - a))
    return code


class LearnedIndexer:
    """Proposal -> local dictionary -> full-dictionary fallback.

    net: a trained ProposalNet; rays (D, 3) unit sample-frame directions of
    the indexing geometry; grid (G, 4) local rotations (default: the ball of
    2,197 orientations within 3 degrees); fallback: a DictionaryIndexer over
    the full dictionary built with dict_unit_norm=True (scores are cosines,
    like the local ones), or None. A pattern falls back when its best local
    cosine is below `threshold` or when the best local orientation lies on
    the grid's outer shell (the optimum is probably outside the ball).
    """

    def __init__(self, net: ProposalNet, master_pattern: torch.Tensor, rays: torch.Tensor, laue_id: int,
                 fallback=None, threshold: float = 0.0, grid: Optional[torch.Tensor] = None,
                 radius_deg: float = 3.0, radius_steps: int = 6, shell_fallback: bool = True,
                 binning: int = 1, local_rays: Optional[torch.Tensor] = None,
                 quad_dtype: torch.dtype = torch.float32, rescore_k: int = 8):
        self.net = net
        # The local refinement may score virtually binned patterns
        # (average-pooled binning x binning, as ebsd/refining.py's
        # virtual_binning with get_patterns(binning=...)) against the binned
        # detector's rays; the network and the fallback see full patterns.
        # With binning, the best rescore_k binned candidates are rescored on
        # the full patterns and the best full-resolution cosine is kept (0:
        # keep the best binned candidate and its binned cosine).
        self.binning = int(binning)
        self.rescore_k = int(rescore_k)
        if not 0 <= self.rescore_k <= 16:
            raise ValueError("rescore_k must be in [0, 16] (the local search keeps at most 16 candidates)")
        if self.binning > 1:
            if local_rays is None:
                raise ValueError("binning > 1 needs the binned detector's rays (geometry.get_rays((b, b)))")
            if net.H % self.binning or net.W % self.binning:
                raise ValueError("the binning must divide the detector shape")
            if local_rays.shape[0] != (net.H // self.binning) * (net.W // self.binning):
                raise ValueError("local_rays do not match the binned detector")
        self.local_rays = (rays if self.binning == 1 else local_rays).to(net.device, torch.float32).contiguous()
        self.device = net.device
        self.mp = master_pattern.to(self.device, torch.float32).contiguous()
        self.rays = rays.to(self.device, torch.float32).contiguous()
        self.laue_id = int(laue_id)
        self.fallback = fallback
        self.threshold = float(threshold)
        self.shell_fallback = bool(shell_fallback)
        if grid is None:
            grid = local_ball_grid_quats(radius_deg, radius_steps)
            r = radius_steps
            t = torch.arange(2 * r + 1) - r
            ijk = torch.stack(torch.meshgrid(t, t, t, indexing="ij"), -1).reshape(-1, 3)
            shell = ijk.abs().amax(-1) == r
            # Morton (Z-order) sort of the lattice points: consecutive runs of
            # 64 (one wave of the refinement kernel) are compact blocks, whose
            # samples share master-pattern cache lines
            order = torch.argsort(_morton3(ijk + r))
            grid, shell = grid[order], shell[order]
        else:
            shell = torch.zeros(grid.shape[0], dtype=torch.bool)
        self.grid = grid.to(self.device, torch.float32).contiguous()
        self.on_shell = shell.to(self.device)
        # float16 taps halve the gather bytes of the local search (relative
        # error of the master-pattern values <= 2^-11); float32 is exact
        self.quads = kernels.mp_quad_table(self.mp, quad_dtype)
        if fallback is not None and fallback.n_pix != self.rays.shape[0]:
            raise ValueError("the fallback dictionary must use the same detector pixels")
        if fallback is not None and not getattr(fallback, "unit_norm", False):
            # the better of the local cosine and the fallback's score is kept,
            # which compares like with like only when both are cosines
            raise ValueError("the fallback DictionaryIndexer must be built with dict_unit_norm=True")

    def _local_queries(self, x: torch.Tensor) -> kernels.PatternMatrix:
        if self.binning == 1:
            return kernels.PatternMatrix(x.contiguous(), None, DTYPE_F32, x.shape[1])
        b = self.binning
        pooled = torch.nn.functional.avg_pool2d(x.contiguous().view(-1, 1, self.net.H, self.net.W), b)
        pooled = pooled.view(x.shape[0], -1)
        return kernels.prep_patterns(pooled, DTYPE_F32, zero_mean=True, unit_norm=True)

    def local(self, x: torch.Tensor):
        """Proposal + local refinement of prepared patterns x (B, D):
        (orientation (B, 4), cosine (B,), grid index (B,), proposal (B, 4))."""
        q0 = self.net.propose(x)
        xm = self._local_queries(x)
        if self.binning > 1 and self.rescore_k > 0:
            k = min(self.rescore_k, self.grid.shape[0])
            _, cand = kernels.local_refine(self.quads, self.local_rays, xm, q0, self.grid, topk=k)
            full = kernels.PatternMatrix(x.contiguous(), None, DTYPE_F32, x.shape[1])
            v, j, _ = kernels.local_rescore(self.quads, self.rays, full, q0, self.grid, cand)
            # -1 only when no candidate was found (score -inf: falls back)
            j = j.long().clamp_min(0)
        else:
            val, idx = kernels.local_refine(self.quads, self.local_rays, xm, q0, self.grid, topk=1)
            v, j = val[:, 0], idx[:, 0].long()
        q = qu_prod(self.grid[j], q0)
        q = q / q.norm(dim=1, keepdim=True)
        return q, v, j, q0

    def index(self, patterns: torch.Tensor, chunk: int = 8192):
        """patterns (N, D) float32 on the GPU -> dict with orientations (N, 4)
        in the fundamental zone, scores (N,), fallback mask (N,), proposals
        (N, 4) and local scores (N,)."""
        n = patterns.shape[0]
        dev = self.device
        ori = torch.empty((n, 4), dtype=torch.float32, device=dev)
        score = torch.empty(n, dtype=torch.float32, device=dev)
        local_score = torch.empty(n, dtype=torch.float32, device=dev)
        prop = torch.empty((n, 4), dtype=torch.float32, device=dev)
        fb = torch.zeros(n, dtype=torch.bool, device=dev)
        for s in range(0, n, chunk):
            p = patterns[s:s + chunk].to(dev, torch.float32)
            b = p.shape[0]
            x = kernels.prep_patterns(p, DTYPE_F32, zero_mean=True, unit_norm=True)
            q, v, j, q0 = self.local(x.data[:, :p.shape[1]])
            need = v < self.threshold
            if self.shell_fallback:
                need |= self.on_shell[j]
            ori[s:s + b] = q
            score[s:s + b] = v
            local_score[s:s + b] = v
            prop[s:s + b] = q0
            fb[s:s + b] = need
        # one full-dictionary pass over every pattern that needs it (the
        # similarity kernel is efficient only on large query batches)
        if self.fallback is not None:
            sel = torch.nonzero(fb).flatten()
            for s in range(0, sel.numel(), 1 << 16):
                part = sel[s:s + (1 << 16)]
                fv, fi = self._full(patterns.index_select(0, part.to(patterns.device)).to(dev, torch.float32))
                # keep the better of the two matches (the local grid is finer
                # than the dictionary where both found the same basin)
                better = fv > local_score[part]
                take = part[better]
                ori[take] = self.fallback.orientations[fi[better].long()]
                score[take] = fv[better]
        ori = ori_to_fz_laue(ori, self.laue_id, left=True)
        return {"orientations": ori, "scores": score, "fallback": fb, "proposals": prop,
                "local_scores": local_score}

    def _full(self, p: torch.Tensor):
        fbk = self.fallback
        q, q32 = fbk.prepare_pair(p, unit_norm=True)
        val, idx = fbk.index(q, 1, queries_f32=q32)
        return val[:, 0], idx[:, 0]


def calibrate_threshold(indexer: LearnedIndexer, sim: PatternSimulator, n: int = 32768,
                        max_error_deg: float = 1.0, precision: float = 0.9999) -> float:
    """Smallest local-cosine threshold t such that, on n simulated patterns
    with known orientations, at least `precision` of the patterns kept
    (score >= t, not on the shell) are within max_error_deg of the truth.

    The defaults are strict on purpose: the rare wrong proposals (a wrong
    basin, tens of degrees off) are few and scattered in score, so a 99.9%
    target on a few thousand patterns rests on a handful of misses and the
    threshold it returns varies widely between calibration sets; at 99.99% on
    32,768 patterns it lands above them for a small increase in fallbacks."""
    from .orientations import disori_angle_laue

    x, q_true = sim.batch(n)
    q, v, j, _ = indexer.local(x)
    err = torch.rad2deg(disori_angle_laue(q.double(), q_true.double(), indexer.laue_id))
    keep = ~indexer.on_shell[j] if indexer.shell_fallback else torch.ones_like(v, dtype=torch.bool)
    v, err = v[keep], err[keep]
    order = torch.argsort(v, descending=True)
    good = (err[order] <= max_error_deg).double()
    frac = torch.cumsum(good, 0) / torch.arange(1, good.numel() + 1, device=good.device)
    ok = torch.nonzero(frac >= precision).flatten()
    if ok.numel() == 0:
        return float("inf")
    # the longest prefix (lowest threshold) that still meets the precision
    last = int(ok.max())
    return float(v[order][last])
