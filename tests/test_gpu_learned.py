"""GPU parity of the learned-proposal kernels against the CPU oracle:
noise, shifted projection, im2col/col2im, split-K GEMM, loss and its
gradient, 6D -> quaternion, Adam, the whole network's forward and backward
(FP32 and BF16 operands) and the fused local-dictionary refinement."""

import math

import numpy as np
import pytest
import torch

from oracle import di_oracle as O
from oracle import learned_oracle as L

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _kernels():
    from ebsdtorch_amd import kernels

    return kernels


def _rand_quats(n, seed):
    g = torch.Generator().manual_seed(seed)
    q = torch.randn(n, 4, generator=g, dtype=torch.float64)
    q = q / q.norm(dim=1, keepdim=True)
    return torch.where(q[:, :1] < 0, -q, q).float()


@pytest.fixture(scope="module")
def small_mp():
    from ebsdtorch_amd.synthetic import band_master_pattern

    return band_master_pattern(101, laue_id=11, n_families=10, sigma_deg=3.0, seed=3)


# ------------------------------------------------------------------ simulation


def test_noise_matches_oracle():
    K = _kernels()
    g = torch.Generator().manual_seed(0)
    clean = 0.1 + 0.9 * torch.rand(4, 700, generator=g)
    # the last pattern's mean count 0 switches the Poisson stage off
    mc = torch.tensor([3.0, 40.0, 300.0, 0.0])
    print("noise 4 x 700", flush=True)
    out = K.poisson_gaussian_noise(clean.to(DEV), 700, mc.to(DEV), 0.05, seed=1234, offset=7).cpu().numpy()
    _, ref = L.poisson_gaussian_noise(clean.numpy(), mc.numpy(), 0.05, 1234, 7)
    match = np.abs(out - ref) < 1e-5
    # float32 rounding of the rate (GPU block sum vs float64 mean) may flip a
    # rare accept/reject decision; everything else is the same random stream
    assert match.mean() > 0.995, match.mean()
    assert np.isfinite(out).all()
    np.testing.assert_allclose(out[3], ref[3], atol=1e-5, rtol=0)


def test_project_shifted_matches_oracle(small_mp):
    K = _kernels()
    from ebsdtorch_amd.geometry import EBSDGeometry

    geo = EBSDGeometry((20, 20), proj_center=(0.5, 0.5, 0.6))
    coords = geo.get_coords_sample_frame((1, 1))
    J = geo.pc_jacobian()
    q = _rand_quats(8, 1)
    dpc = (torch.rand(8, 3, generator=torch.Generator().manual_seed(2)) * 2 - 1) * 0.01
    off = dpc @ J.T
    print("project_shifted 8 x 400", flush=True)
    got = K.project_shifted(small_mp.to(DEV), coords.to(DEV), q.to(DEV), off.to(DEV)).data[:, :400].cpu()
    for n in range(8):
        rays = coords + off[n]
        rays = rays / rays.norm(dim=1, keepdim=True)
        ref = O.project_dictionary(small_mp, rays, q[n:n + 1])[0]
        torch.testing.assert_close(got[n], ref, atol=5e-5, rtol=0)


# ------------------------------------------------------------------ network pieces


def test_im2col_matches_unfold_and_transposed():
    K = _kernels()
    g = torch.Generator().manual_seed(0)
    x = torch.randn(2, 9, 7, 3, generator=g)
    ref = torch.nn.functional.unfold(x.permute(0, 3, 1, 2), 3, padding=1, stride=2)  # (B, C*9, L)
    ref = ref.view(2, 3, 9, -1).permute(0, 3, 2, 1).reshape(-1, 27)          # rows (b, oy, ox), cols (k, c)
    col = torch.full((ref.shape[0], 32), 7.0, device=DEV)
    print("im2col 2x9x7x3", flush=True)
    K.im2col(x.to(DEV), col, scale=2.0)
    torch.testing.assert_close(col[:, :27].cpu(), 2.0 * ref, atol=0, rtol=0)
    assert float(col[:, 27:].abs().max()) == 0.0
    colt = torch.full((27, 64), 7.0, device=DEV, dtype=torch.bfloat16)
    K.im2col(x.to(DEV), colt, scale=2.0, transposed=True)
    torch.testing.assert_close(colt[:, :ref.shape[0]].float().cpu(), (2.0 * ref).T.bfloat16().float(), atol=0, rtol=0)
    assert float(colt[:, ref.shape[0]:].float().abs().max()) == 0.0


@pytest.mark.parametrize("dtype,ld", [(torch.float32, 144), (torch.bfloat16, 192), (torch.float32, 150)])
def test_im2col_channel_vectors_match_unfold(dtype, ld):
    """C a multiple of 8 (whole 8-channel runs of one tap per thread), on
    padded and unpadded rows; ld 150 (not a multiple of 8) takes the generic
    kernel."""
    K = _kernels()
    g = torch.Generator().manual_seed(5)
    x = torch.randn(3, 11, 10, 16, generator=g)
    ref = torch.nn.functional.unfold(x.permute(0, 3, 1, 2), 3, padding=1, stride=2)
    ref = ref.view(3, 16, 9, -1).permute(0, 3, 2, 1).reshape(-1, 144)
    col = torch.full((ref.shape[0], ld), 7.0, device=DEV, dtype=dtype)
    print("im2col C=16", dtype, ld, flush=True)
    K.im2col(x.to(DEV), col, scale=0.5)
    torch.testing.assert_close(col[:, :144].float().cpu(), (0.5 * ref).to(dtype).float(), atol=0, rtol=0)
    if ld > 144:
        assert float(col[:, 144:].float().abs().max()) == 0.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_conv_c1_matches_im2col_gemm(dtype):
    """The direct first convolution equals im2col + GEMM + bias/ReLU up to the
    order of the 9-term fp32 sum (bf16 taps and weights: exact products), and
    the fused GEMM epilogue equals GEMM then bias_act bit for bit."""
    K = _kernels()
    g = torch.Generator().manual_seed(6)
    x = torch.randn(3, 13, 10, 1, generator=g)
    co = 32
    kp = 64 if dtype == torch.bfloat16 else 32
    wt = torch.zeros(co, kp)
    wt[:, :9] = torch.randn(co, 9, generator=g)
    bias = 0.3 * torch.randn(co, generator=g)
    wd = wt.to(dtype).to(DEV)
    print("conv_c1", dtype, flush=True)
    y = K.conv_c1(x.to(DEV), wd, bias.to(DEV), scale=1.7).cpu().double()
    col = torch.empty((3 * 7 * 5, kp), dtype=dtype, device=DEV)
    K.im2col(x.to(DEV), col, scale=1.7)
    ref = torch.empty((col.shape[0], co), device=DEV)
    K.matmul_nt(col, wd, ref)
    K.bias_act(ref, bias.to(DEV), co, relu=True)
    torch.testing.assert_close(y, ref.cpu().double(), atol=2e-6, rtol=1e-6)
    # against float64 on the rounded operands
    taps = torch.nn.functional.unfold((1.7 * x).to(dtype).double().permute(0, 3, 1, 2), 3, padding=1, stride=2)
    taps = taps.permute(0, 2, 1).reshape(-1, 9)
    ref64 = torch.relu(taps @ wt[:, :9].to(dtype).double().T + bias.double())
    torch.testing.assert_close(y, ref64, atol=1e-5, rtol=1e-5)
    fused = torch.empty_like(ref)
    K.matmul_nt_bias_act(col, wd, fused, bias.to(DEV), relu=True)
    assert torch.equal(fused, ref)
    lin = torch.empty_like(ref)
    K.matmul_nt_bias_act(col, wd, lin, bias.to(DEV), relu=False)
    plain = torch.empty_like(ref)
    K.matmul_nt(col, wd, plain)
    K.bias_act(plain, bias.to(DEV), co, relu=False)
    assert torch.equal(lin, plain)


@pytest.mark.parametrize("dtype,c,c_out,b,h,w", [(torch.float32, 4, 6, 3, 13, 10), (torch.float32, 12, 130, 2, 9, 16),
                                                 (torch.bfloat16, 8, 64, 5, 11, 11),
                                                 (torch.bfloat16, 32, 130, 4, 16, 16)])
def test_conv_gemm_equals_im2col_gemm(dtype, c, c_out, b, h, w):
    """The implicit-GEMM convolution builds the same operand stages as im2col
    and runs the same MFMA sequence: equal bit for bit; and within rounding of
    float64 on the rounded operands."""
    K = _kernels()
    g = torch.Generator().manual_seed(c + c_out)
    x = torch.relu(torch.randn(b, h, w, c, generator=g))
    kp = K.padded_width(9 * c, K.MATCH_DTYPES[dtype])
    wt = torch.zeros(c_out, kp)
    wt[:, :9 * c] = torch.randn(c_out, 9 * c, generator=g) / math.sqrt(9 * c)
    bias = 0.1 * torch.randn(c_out, generator=g)
    wd, bd, xd = wt.to(dtype).to(DEV), bias.to(DEV), x.to(DEV)
    print("conv_gemm", dtype, c, c_out, b, h, w, flush=True)
    y = K.conv_gemm(xd, wd, bd)
    col = torch.empty((b * ((h + 1) // 2) * ((w + 1) // 2), kp), dtype=dtype, device=DEV)
    K.im2col(xd, col)
    ref = torch.empty((col.shape[0], c_out), device=DEV)
    K.matmul_nt_bias_act(col, wd, ref, bd, relu=True)
    assert torch.equal(y, ref)
    taps = torch.nn.functional.unfold(x.to(dtype).double().permute(0, 3, 1, 2), 3, padding=1, stride=2)
    taps = taps.view(b, c, 9, -1).permute(0, 3, 2, 1).reshape(-1, 9 * c)
    ref64 = torch.relu(taps @ wt[:, :9 * c].to(dtype).double().T + bias.double())
    torch.testing.assert_close(y.cpu().double(), ref64, atol=1e-4, rtol=1e-4)


def test_col2im_is_adjoint_of_im2col():
    K = _kernels()
    g = torch.Generator().manual_seed(1)
    shape = (3, 11, 8, 5)
    x = torch.randn(shape, generator=g)
    rows = 3 * 6 * 4
    y = torch.randn(rows, 64, generator=g)
    col = torch.empty((rows, 64), device=DEV)
    K.im2col(x.to(DEV), col)
    dx = K.col2im(y.to(DEV), shape).cpu()
    lhs = float((col.cpu()[:, :45].double() * y[:, :45].double()).sum())
    rhs = float((x.double() * dx.double()).sum())
    assert abs(lhs - rhs) < 1e-4 * max(1.0, abs(lhs))


@pytest.mark.parametrize("qdt", [torch.float32, torch.float16])
def test_quad_table(qdt):
    K = _kernels()
    g = torch.Generator().manual_seed(9)
    mp = torch.rand(5, 7, generator=g)
    q = K.mp_quad_table(mp.to(DEV), qdt).cpu()
    m = mp.to(qdt)
    for y in range(5):
        for x in range(7):
            x1, y1 = min(x + 1, 6), min(y + 1, 4)
            assert q[y, x].tolist() == [m[y, x].item(), m[y, x1].item(), m[y1, x].item(), m[y1, x1].item()]


def test_cast_transpose_bias_relu_colsum():
    K = _kernels()
    g = torch.Generator().manual_seed(2)
    a = torch.randn(37, 21, generator=g)
    out = torch.full((37, 64), 3.0, device=DEV, dtype=torch.bfloat16)
    K.cast_pad(a.to(DEV), out)
    torch.testing.assert_close(out[:, :21].float().cpu(), a.bfloat16().float(), atol=0, rtol=0)
    assert float(out[:, 21:].float().abs().max()) == 0.0
    t = torch.full((21, 64), 3.0, device=DEV)
    K.transpose_pad(a.to(DEV), t)
    torch.testing.assert_close(t[:, :37].cpu(), a.T, atol=0, rtol=0)
    assert float(t[:, 37:].abs().max()) == 0.0
    y = a.to(DEV).clone()
    b = torch.randn(21, generator=g)
    K.bias_act(y, b.to(DEV), 21, relu=True)
    torch.testing.assert_close(y.cpu(), torch.relu(a + b), atol=0, rtol=0)
    dy = torch.randn(37, 21, generator=g)
    d = dy.to(DEV).clone()
    K.relu_bwd(d, y)
    torch.testing.assert_close(d.cpu(), dy * (torch.relu(a + b) > 0), atol=0, rtol=0)
    s = torch.empty(21, device=DEV)
    big = torch.randn(1000, 21, generator=g)
    K.colsum(big.to(DEV), 21, s)
    torch.testing.assert_close(s.cpu().double(), big.double().sum(0), atol=1e-4, rtol=0)
    # more rows than 65,535 partial sums of 256 rows: wider partials, exact here
    # (every partial and running sum is a multiple of 64 below 2^25)
    tall = torch.ones(17_000_000, 1, device=DEV)
    s1 = torch.empty(1, device=DEV)
    print("colsum 17M x 1", flush=True)
    K.colsum(tall, 1, s1)
    assert float(s1.item()) == 17_000_000.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_split_k_gemm_matches_single_pass(dtype):
    K = _kernels()
    g = torch.Generator().manual_seed(3)
    a = torch.randn(100, 4096, generator=g).to(dtype)
    b = torch.randn(70, 4096, generator=g).to(dtype)
    c1 = torch.empty((100, 70), device=DEV)
    c2 = torch.empty((100, 70), device=DEV)
    print("gemm split-k", dtype, flush=True)
    K.matmul_nt(a.to(DEV), b.to(DEV), c1)
    K.matmul_nt(a.to(DEV), b.to(DEV), c2, splits=8)
    ref = a.double() @ b.double().T
    torch.testing.assert_close(c1.cpu().double(), ref, atol=1e-3, rtol=1e-5)
    torch.testing.assert_close(c2.cpu().double(), ref, atol=1e-3, rtol=1e-5)


def test_geodesic_loss_kernel_matches_oracle():
    K = _kernels()
    from ebsdtorch_amd.orientations import laue_elements

    g = torch.Generator().manual_seed(4)
    z = torch.randn(300, 6, generator=g)
    q = _rand_quats(300, 5)
    ops = laue_elements(11).float()
    loss = torch.empty(300, device=DEV)
    dz = torch.empty((300, 6), device=DEV)
    print("geodesic loss 300", flush=True)
    K.geodesic_loss(z.to(DEV), q.to(DEV), ops.to(DEV), loss, dz)
    zz = z.double().requires_grad_(True)
    ref = L.geodesic_sym_loss(zz, q.double(), ops.double())
    ref.mean().backward()
    torch.testing.assert_close(loss.cpu().double(), ref.detach(), atol=2e-5, rtol=0)
    torch.testing.assert_close(dz.cpu().double(), zz.grad, atol=2e-6, rtol=2e-3)


def test_rot6d_to_quat_matches_gram_schmidt():
    K = _kernels()
    g = torch.Generator().manual_seed(6)
    z = torch.randn(500, 6, generator=g)
    q = K.rot6d_to_quat(z.to(DEV)).cpu()
    assert bool((q[:, 0] >= 0).all())
    torch.testing.assert_close(q.norm(dim=1), torch.ones(500), atol=1e-6, rtol=0)
    torch.testing.assert_close(O.qu2om(q.double()), L.rot6d_matrix(z.double()), atol=2e-6, rtol=0)


def test_adam_matches_torch():
    K = _kernels()
    g = torch.Generator().manual_seed(7)
    w0 = torch.randn(5000, generator=g)
    grads = [torch.randn(5000, generator=g) for _ in range(3)]
    w = w0.to(DEV).clone()
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    wb = torch.empty(5000, device=DEV, dtype=torch.bfloat16)
    p = torch.nn.Parameter(w0.clone())
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for s, gr in enumerate(grads, 1):
        K.adam_step(w, gr.to(DEV), m, v, 1e-2, 0.9, 0.999, 1e-8, s, 0.0, wb)
        p.grad = gr.clone()
        opt.step()
    torch.testing.assert_close(w.cpu(), p.detach(), atol=1e-6, rtol=1e-6)
    torch.testing.assert_close(wb.float().cpu(), w.cpu().bfloat16().float(), atol=0, rtol=0)


# ------------------------------------------------------------------ whole network


def _net_case(compute, shape, channels, hidden, batch, seed):
    from ebsdtorch_amd.learned import ProposalNet

    net = ProposalNet(shape, DEV, laue_id=11, compute=compute, channels=channels, hidden=hidden, seed=seed)
    g = torch.Generator().manual_seed(seed + 1)
    x = torch.randn(batch, shape[0] * shape[1], generator=g)
    x = x - x.mean(1, keepdim=True)
    x = x / x.norm(dim=1, keepdim=True)
    q = _rand_quats(batch, seed + 2)
    print("network", compute, shape, channels, hidden, batch, flush=True)
    z = net.forward(x.to(DEV), keep=True)
    loss, dz = net.loss_and_grad(z, q.to(DEV))
    net.backward(dz)
    params = net.torch_parameters()
    ref_loss, ref_z, ref_g = L.net_loss_and_grads(params, x, q, net.ops.cpu(), shape)
    return net, z.cpu(), float(loss.mean()), net.gradients_torch_layout(), ref_z, ref_loss, ref_g


def _rel(a, b):
    return float((a.double() - b.double()).norm() / max(b.double().norm(), 1e-30))


@pytest.mark.parametrize("shape,channels,hidden,batch", [
    ((32, 32), (8, 16, 32, 32), 64, 8),
    ((64, 64), (32, 64, 128, 128), 512, 4),
    ((60, 60), (32, 64, 128, 128), 512, 3),
])
def test_network_fp32_matches_autograd(shape, channels, hidden, batch):
    net, z, loss, grads, rz, rloss, rg = _net_case("fp32", shape, channels, hidden, batch, 10)
    assert _rel(z, rz) < 1e-4
    assert abs(loss - rloss) < 1e-4
    for k in rg:
        assert _rel(grads[k], rg[k]) < 2e-3, k


def test_network_bf16_matches_autograd():
    net, z, loss, grads, rz, rloss, rg = _net_case("bf16", (64, 64), (32, 64, 128, 128), 512, 8, 20)
    assert _rel(z, rz) < 5e-2
    assert abs(loss - rloss) < 5e-2
    for k in rg:
        cos = float((grads[k].double().flatten() @ rg[k].double().flatten())
                    / (grads[k].double().norm() * rg[k].double().norm() + 1e-30))
        assert cos > 0.98, (k, cos)


# ------------------------------------------------------------------ local refinement


@pytest.mark.parametrize("det,qdt", [(20, torch.float32), (70, torch.float32), (20, torch.float16)])
def test_local_refine_matches_oracle(small_mp, det, qdt):
    K = _kernels()
    from ebsdtorch_amd.orientations import local_ball_grid_quats, qu_prod

    rays = O.normalized_rays((det, det), (0.5, 0.5, 0.6))
    truth = _rand_quats(4, 30 + det)
    pats = O.project_dictionary(small_mp, rays, truth)
    g = torch.Generator().manual_seed(det)
    pats = pats + 0.02 * torch.randn(pats.shape, generator=g)
    x = pats - pats.mean(1, keepdim=True)
    x = (x / x.norm(dim=1, keepdim=True)).float()
    # proposals about 1 degree off the truth
    pert = torch.nn.functional.normalize(torch.tensor([[1.0, 0.006, -0.004, 0.005]]).expand(4, 4), dim=1)
    centers = qu_prod(pert, truth)
    centers = (centers / centers.norm(dim=1, keepdim=True)).float()
    grid = local_ball_grid_quats(3.0, 3)
    xm = K.PatternMatrix(x.to(DEV).contiguous(), None, 0, det * det)
    print("local_refine", det, flush=True)
    quads = K.mp_quad_table(small_mp.to(DEV), qdt)
    val, idx = K.local_refine(quads, rays.to(DEV), xm, centers.to(DEV), grid.to(DEV), topk=4)
    val, idx = val.cpu(), idx.cpu().long()
    # float16 taps: master-pattern values within 2^-11 relative, scores within ~1e-3
    tol = 2e-5 if qdt == torch.float32 else 2e-3
    ref = L.local_scores(small_mp, rays, x.double(), centers, grid)
    for i in range(4):
        # every reported score is the oracle's score of that candidate
        torch.testing.assert_close(val[i].double(), ref[i, idx[i]], atol=tol, rtol=0)
        best = float(ref[i].max())
        assert float(ref[i, idx[i, 0]]) >= best - tol
        # sorted descending
        assert bool((val[i, 1:] <= val[i, :-1]).all())


@pytest.mark.parametrize("det,k,qdt", [(20, 1, torch.float32), (20, 5, torch.float32), (35, 13, torch.float32),
                                       (20, 64, torch.float32), (20, 8, torch.float16)])
def test_local_rescore_matches_oracle(small_mp, det, k, qdt):
    K = _kernels()
    from ebsdtorch_amd.orientations import local_ball_grid_quats, qu_prod

    n = 6
    rays = O.normalized_rays((det, det), (0.5, 0.5, 0.6))
    truth = _rand_quats(n, 70 + det + k)
    pats = O.project_dictionary(small_mp, rays, truth)
    g = torch.Generator().manual_seed(det + k)
    pats = pats + 0.02 * torch.randn(pats.shape, generator=g)
    x = pats - pats.mean(1, keepdim=True)
    x = (x / x.norm(dim=1, keepdim=True)).float()
    pert = torch.nn.functional.normalize(torch.tensor([[1.0, 0.006, -0.004, 0.005]]).expand(n, 4), dim=1)
    centers = qu_prod(pert, truth)
    centers = (centers / centers.norm(dim=1, keepdim=True)).float()
    grid = local_ball_grid_quats(3.0, 3)
    ref = L.local_scores(small_mp, rays, x.double(), centers, grid)
    # random candidate lists with repeats; row 1 holds -1 entries, row 2 only -1
    cand = torch.randint(0, grid.shape[0], (n, k), generator=g)
    cand[1, ::2] = -1
    cand[2] = -1
    if k == 1:
        cand[1] = -1
    # query rows padded to a longer row stride
    ld = det * det + 7
    xp = torch.zeros(n, ld)
    xp[:, :det * det] = x
    xm = K.PatternMatrix(xp.to(DEV), None, 0, det * det)
    print("local_rescore", det, k, qdt, flush=True)
    quads = K.mp_quad_table(small_mp.to(DEV), qdt)
    bv, bi, sc = K.local_rescore(quads, rays.to(DEV), xm, centers.to(DEV), grid.to(DEV), cand.to(DEV),
                                 all_scores=True)
    bv, bi, sc = bv.cpu().double(), bi.cpu().long(), sc.cpu().double()
    tol = 2e-5 if qdt == torch.float32 else 2e-3
    for i in range(n):
        valid = cand[i] >= 0
        assert bool(torch.isinf(sc[i][~valid]).all()) and bool((sc[i][~valid] < 0).all())
        if valid.any():
            torch.testing.assert_close(sc[i][valid], ref[i, cand[i][valid]], atol=tol, rtol=0)
            # the best is the first listed candidate with the maximal score
            pos = int(torch.nonzero(sc[i] == sc[i].max())[0])
            assert int(bi[i]) == int(cand[i, pos]) and float(bv[i]) == float(sc[i, pos])
            assert float(ref[i, bi[i]]) >= float(ref[i, cand[i][valid]].max()) - tol
        else:
            assert int(bi[i]) == -1 and float(bv[i]) == float("-inf")


def test_binned_local_search_rescored_at_full_resolution():
    """With binning, LearnedIndexer.local returns the best full-resolution
    cosine among the binned search's top rescore_k candidates."""
    from ebsdtorch_amd.learned import LearnedIndexer, ProposalNet
    from ebsdtorch_amd.geometry import EBSDGeometry
    from ebsdtorch_amd.orientations import local_ball_grid_quats
    from ebsdtorch_amd.synthetic import band_master_pattern

    K = _kernels()
    geo = EBSDGeometry((32, 32), proj_center=(0.5, 0.5, 0.6))
    mp = band_master_pattern(41, laue_id=11, n_families=4, seed=1)
    net = ProposalNet((32, 32), DEV, laue_id=11, compute="fp32", channels=(8, 16, 32, 32), hidden=64)
    grid = local_ball_grid_quats(3.0, 2)
    rays = geo.get_rays()
    li = LearnedIndexer(net, mp.to(DEV), rays.to(DEV), 11, grid=grid, binning=2,
                        local_rays=geo.get_rays((2, 2)).to(DEV), rescore_k=6)
    truth = _rand_quats(5, 91)
    pats = O.project_dictionary(mp, rays, truth)
    x = pats - pats.mean(1, keepdim=True)
    x = (x / x.norm(dim=1, keepdim=True)).float().to(DEV)
    print("binned rescore", flush=True)
    q, v, j, q0 = li.local(x)
    xm = li._local_queries(x)
    _, cand = K.local_refine(li.quads, li.local_rays, xm, q0, li.grid, topk=6)
    ref = L.local_scores(mp, rays, x.cpu().double(), q0.cpu(), grid)
    cand, j, v = cand.cpu().long(), j.cpu().long(), v.cpu().double()
    for i in range(5):
        assert int(j[i]) in cand[i].tolist()
        torch.testing.assert_close(v[i], ref[i, j[i]], atol=2e-5, rtol=0)
        assert float(ref[i, j[i]]) >= float(ref[i, cand[i]].max()) - 2e-5


def test_binned_local_queries_match_avg_pool():
    from ebsdtorch_amd.learned import LearnedIndexer, ProposalNet
    from ebsdtorch_amd.geometry import EBSDGeometry
    from ebsdtorch_amd.synthetic import band_master_pattern

    geo = EBSDGeometry((32, 32), proj_center=(0.5, 0.5, 0.6))
    mp = band_master_pattern(41, laue_id=11, n_families=4, seed=1).to(DEV)
    net = ProposalNet((32, 32), DEV, laue_id=11, compute="fp32", channels=(8, 16, 32, 32), hidden=64)
    li = LearnedIndexer(net, mp, geo.get_rays().to(DEV), 11, binning=2, local_rays=geo.get_rays((2, 2)).to(DEV))
    g = torch.Generator().manual_seed(4)
    x = torch.randn(5, 1024, generator=g)
    got = li._local_queries(x.to(DEV)).data.cpu().double()
    ref = torch.nn.functional.avg_pool2d(x.double().view(5, 1, 32, 32), 2).view(5, -1)
    ref = ref - ref.mean(1, keepdim=True)
    ref = ref / ref.norm(dim=1, keepdim=True)
    torch.testing.assert_close(got, ref, atol=1e-6, rtol=0)
