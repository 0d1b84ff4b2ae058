"""GPU parity of each kernel against the CPU oracle (or a float64 brute
force), called through the C ABI."""

import math

import numpy as np
import pytest
import torch

import ebsdtorch_amd as E
from ebsdtorch_amd import kernels, synthetic
from ebsdtorch_amd._lib import DTYPE_BF16, DTYPE_F16, DTYPE_F32, DTYPE_FP8_E4M3
from oracle import di_oracle as O

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def mp101():
    return synthetic.band_master_pattern(101, 11, 40, 2.0, seed=0)


def _rays(shape=(60, 60), pc=(0.5, 0.5, 0.6)):
    return O.normalized_rays(shape, pc)


def test_sample_dirs_matches_oracle_interpolate(golden):
    _say("sample_dirs vs golden interpolate (33x66 MP, 3000 points)")
    mp = E.MasterPattern(torch.tensor(golden["mp33"]).to(DEV), laue_group=11)
    pts = torch.tensor(golden["interp_pts"])
    got = mp.interpolate(pts.to(DEV), normalize_coords=False).cpu()
    ref = torch.tensor(golden["interp_out"])
    assert (got - ref).abs().max().item() < 2e-5


@pytest.mark.parametrize("shape,n", [((60, 60), 300), ((8, 8), 50), ((80, 80), 40), ((120, 120), 12)])
def test_projection_matches_oracle(mp101, shape, n):
    _say(f"projection {shape} n={n}")
    rays = _rays(shape)
    q = synthetic.random_orientations(n, seed=11)
    ref = O.project_dictionary(mp101, rays, q)  # zero-mean, reference op order
    got = kernels.project(mp101.to(DEV), rays.to(DEV), q.to(DEV))
    assert got.ld % 32 == 0 and got.ld >= rays.shape[0]
    g = got.data.cpu()
    assert torch.all(g[:, rays.shape[0]:] == 0)
    assert (g[:, : rays.shape[0]] - ref).abs().max().item() < 2e-5


@pytest.mark.parametrize("shape,dtype", [((60, 60), DTYPE_F32), ((60, 60), DTYPE_FP8_E4M3), ((120, 120), DTYPE_F32),
                                         ((20, 20), DTYPE_BF16)])
def test_projection_processing_order_changes_nothing(mp101, shape, dtype):
    """The processing order of ebsd_project_quads (view_order's sort, any
    permutation, or none) only schedules rows: the output is bit for bit the
    same, and equals the plain ebsd_project's."""
    _say(f"projection order {shape} dtype={dtype}")
    mp = mp101.to(DEV)
    rays = _rays(shape).to(DEV)
    q = synthetic.random_orientations(3000, seed=13).to(DEV)
    quads = kernels.mp_quad_table(mp)
    plain = kernels.project(mp, rays, q, dtype)
    base = kernels.project(mp, rays, q, dtype, quads=quads, order=None)
    order = kernels.view_order(q, rays)
    assert order.dtype == torch.int32 and torch.equal(torch.sort(order.long()).values,
                                                      torch.arange(3000, device=DEV))
    perm = torch.randperm(3000, generator=torch.Generator().manual_seed(4)).to(DEV, torch.int32)
    for o in (order, perm):
        got = kernels.project(mp, rays, q, dtype, quads=quads, order=o)
        assert torch.equal(got.data, base.data)
        if dtype == DTYPE_FP8_E4M3:
            assert torch.equal(got.scale, base.scale)
    assert torch.equal(base.data, plain.data)


def test_projection_raw_unit_norm_and_renorm(mp101):
    _say("projection flags")
    rays = _rays()
    coords = O.detector_coords((60, 60), proj_center=(0.5, 0.5, 0.6))
    q = synthetic.random_orientations(64, seed=12)
    raw = kernels.project(mp101.to(DEV), rays.to(DEV), q.to(DEV), zero_mean=False).to_float().cpu()
    ref_raw = O.interpolate(mp101, O.qu_apply(q[:, None], rays[None]), normalize_coords=False)
    assert (raw - ref_raw).abs().max().item() < 2e-5
    # normalize_coords=True path of PCA DI: un-normalised coordinates
    rn = kernels.project(mp101.to(DEV), coords.to(DEV), q.to(DEV), zero_mean=False, renorm_dirs=True).to_float().cpu()
    ref_rn = O.interpolate(mp101, O.qu_apply(q[:, None], coords[None]), normalize_coords=True)
    assert (rn - ref_rn).abs().max().item() < 2e-5
    un = kernels.project(mp101.to(DEV), rays.to(DEV), q.to(DEV), unit_norm=True).to_float().cpu().double()
    zm = ref_raw.double() - ref_raw.double().mean(1, keepdim=True)
    assert torch.allclose(un, zm / zm.norm(dim=1, keepdim=True), atol=1e-6)


@pytest.mark.parametrize("dtype", [DTYPE_BF16, DTYPE_F16, DTYPE_FP8_E4M3])
def test_projection_low_precision_outputs(mp101, dtype):
    _say(f"projection dtype {dtype}")
    rays = _rays()
    q = synthetic.random_orientations(100, seed=13)
    f32 = kernels.project(mp101.to(DEV), rays.to(DEV), q.to(DEV)).to_float()
    lp = kernels.project(mp101.to(DEV), rays.to(DEV), q.to(DEV), dtype)
    assert lp.ld % kernels.pad_multiple(dtype) == 0
    if dtype == DTYPE_FP8_E4M3:
        # the scale is the smallest power of two 2^k with amax / 2^k <= 448,
        # and the stored bytes are the OCP e4m3 rounding of value / 2^k
        m, e = torch.frexp(f32.abs().amax(1) / 448.0)
        k = torch.where(m == 0.5, e - 1, e)
        assert torch.equal(lp.scale, torch.ldexp(torch.ones_like(m), k))
        amax = f32.abs().amax(1)
        assert torch.all(amax / lp.scale <= 448) and torch.all(amax / lp.scale > 224)
        expect = torch.ldexp(f32, -k[:, None]).to(torch.float8_e4m3fn).view(torch.uint8)
        assert torch.equal(lp.data[:, :3600], expect)
        assert torch.all(lp.data[:, 3600:] == 0)
    else:
        tdt = torch.bfloat16 if dtype == DTYPE_BF16 else torch.float16
        assert torch.equal(lp.data[:, :3600], f32.to(tdt))


@pytest.mark.parametrize("p", [3600, 14400, 5000, 3597])
def test_prep_patterns_mean_mask_and_norm(p):
    """Rows of up to 4,096 pixels take 256-thread workgroups, longer rows
    1,024 threads (8 or 16 values per thread); widths that are multiples of
    4 are read and written 16 bytes per access, others one value at a
    time."""
    _say(f"prep patterns p={p}")
    g = torch.Generator().manual_seed(4 + p)
    x = torch.rand(37, p, generator=g) * 3 + 1
    got = kernels.prep_patterns(x.to(DEV)).to_float().cpu()
    assert torch.allclose(got, x - x.mean(1, keepdim=True), atol=2e-6)
    idx = torch.nonzero(torch.rand(p, generator=g) > 0.3).flatten()
    got = kernels.prep_patterns(x.to(DEV), pix_idx=idx.to(DEV)).to_float().cpu()
    xs = x[:, idx]
    assert got.shape[1] == len(idx)
    assert torch.allclose(got, xs - xs.mean(1, keepdim=True), atol=2e-6)
    got = kernels.prep_patterns(x.to(DEV), unit_norm=True).to_float().cpu().double()
    z = x.double() - x.double().mean(1, keepdim=True)
    assert torch.allclose(got, z / z.norm(dim=1, keepdim=True), atol=1e-6)
    # FP8 rows: power-of-two scale, the largest magnitude in (224, 448]
    q8 = kernels.prep_patterns(x.to(DEV), DTYPE_FP8_E4M3)
    amax = q8.data[:, :p].view(torch.float8_e4m3fn).float().abs().amax(1)
    assert torch.all(amax > 224) and torch.all(amax <= 448)
    assert torch.allclose(q8.to_float().cpu().double(), z, rtol=0.07, atol=0.07 * z.abs().max().item() / 16)


@pytest.mark.parametrize("dtype2", [DTYPE_BF16, DTYPE_F16])
@pytest.mark.parametrize("p,masked,unit_norm", [(3600, False, False), (3600, True, True), (100, False, True),
                                                (14400, True, False), (3597, False, False)])
def test_prep_patterns_pair_equals_two_preps(dtype2, p, masked, unit_norm):
    """ebsd_prep_patterns_pair writes bit for bit what two ebsd_prep_patterns
    calls (FP32 and dtype2) write, padding included."""
    _say(f"prep pair dtype2={dtype2} p={p} masked={masked} unit_norm={unit_norm}")
    g = torch.Generator().manual_seed(p + dtype2)
    x = (torch.rand(45, p, generator=g) * 3 + 1).to(DEV)
    idx = torch.nonzero(torch.rand(p, generator=g) > 0.3).flatten().to(DEV) if masked else None
    f, h = kernels.prep_patterns_pair(x, dtype2, unit_norm=unit_norm, pix_idx=idx)
    f_ref = kernels.prep_patterns(x, DTYPE_F32, unit_norm=unit_norm, pix_idx=idx)
    h_ref = kernels.prep_patterns(x, dtype2, unit_norm=unit_norm, pix_idx=idx)
    assert f.ld == f_ref.ld and h.ld == h_ref.ld
    assert torch.equal(f.data, f_ref.data) and torch.equal(h.data.view(torch.int16), h_ref.data.view(torch.int16))


def _check_topk(val, idx, q, d, k, tol=1e-4):
    """val/idx against float64 brute force, allowing swaps only between
    near-equal scores."""
    full = q.double() @ d.double().T
    ref_v, ref_i = torch.topk(full, k, dim=1)
    assert torch.allclose(val.double(), ref_v, atol=tol * ref_v.abs().max().item())
    mism = idx != ref_i
    if mism.any():
        gv = full.gather(1, idx.clamp_min(0).long())
        assert torch.allclose(gv[mism], ref_v[mism], atol=tol * ref_v.abs().max().item())
    # rows sorted descending
    assert torch.all(val[:, 1:] <= val[:, :-1])


@pytest.mark.parametrize("m,n,d,k", [(300, 5000, 200, 1), (300, 5000, 200, 4), (129, 1000, 3600, 16),
                                     (1, 130, 64, 16), (1000, 20000, 96, 8)])
def test_sim_topk_f32_brute_force(m, n, d, k):
    _say(f"sim_topk f32 m={m} n={n} d={d} k={k}")
    g = torch.Generator().manual_seed(m + n + d + k)
    q = torch.randn(m, d, generator=g)
    dd = torch.randn(n, d, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), zero_mean=False)
    val = torch.empty(m, k, device=DEV)
    idx = torch.empty(m, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, val, idx, merge=False)
    _check_topk(val.cpu(), idx.cpu(), q, dd, k)


def test_sim_topk_streaming_merge_equals_one_pass():
    _say("sim_topk streaming merge")
    g = torch.Generator().manual_seed(5)
    q = torch.randn(200, 128, generator=g)
    dd = torch.randn(7000, 128, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), zero_mean=False)
    k = 16
    v1 = torch.empty(200, k, device=DEV)
    i1 = torch.empty(200, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, v1, i1, merge=False)
    v2 = torch.full((200, k), float("-inf"), device=DEV)
    i2 = torch.full((200, k), -1, dtype=torch.int32, device=DEV)
    for s in range(0, 7000, 999):
        kernels.sim_topk(Q, D.narrow(s, min(999, 7000 - s)), k, v2, i2, index_base=s, merge=True)
    assert torch.equal(i1, i2) and torch.equal(v1, v2)
    # the ChunkedKNN wrapper: chunked data and queries, global indices
    kn = E.ChunkedKNN(7000, 200, 5, DEV, match_dtype=torch.float32)
    for s in range(0, 7000, 1500):
        kn.set_data_chunk(dd[s:s + 1500].to(DEV))
        for qs in range(0, 200, 64):
            kn.query_chunk(q[qs:qs + 64].to(DEV), qs)
    ki, kv = kn.retrieve_topk()
    assert ki.dtype == torch.int64
    assert torch.equal(ki.cpu(), i1[:, :5].long().cpu())


def test_sim_topk_fewer_rows_than_k_and_ties():
    _say("sim_topk ties")
    q = torch.ones(3, 32)
    dd = torch.ones(5, 32)  # all scores equal: lower index wins
    Q = kernels.prep_patterns(q.to(DEV), zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), zero_mean=False)
    val = torch.empty(3, 8, device=DEV)
    idx = torch.empty(3, 8, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, 8, val, idx, merge=False)
    assert idx.cpu().tolist() == [[0, 1, 2, 3, 4, -1, -1, -1]] * 3
    assert torch.all(val[:, :5] == 32) and torch.all(torch.isinf(val[:, 5:]))


def test_rescore_exact_fp32():
    _say("rescore")
    g = torch.Generator().manual_seed(8)
    q = torch.randn(100, 300, generator=g)
    dd = torch.randn(900, 300, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), zero_mean=False)
    cand = torch.randint(0, 900, (100, 16), generator=g, dtype=torch.int32)
    cand[:, -1] = -1
    v, i = kernels.rescore(Q, D, cand.to(DEV), 4)
    full = (q.double() @ dd.double().T).gather(1, cand.clamp_min(0).long())
    full[:, -1] = -math.inf
    rv, ri = torch.topk(full, 4, dim=1)
    assert torch.allclose(v.cpu().double(), rv, atol=1e-4)
    assert torch.equal(i.cpu().long(), cand.long().gather(1, ri))


@pytest.mark.parametrize("dtype", [DTYPE_BF16, DTYPE_F16, DTYPE_FP8_E4M3])
@pytest.mark.parametrize("m,n,d,k", [(300, 5000, 200, 4), (129, 3000, 3600, 16), (1000, 9000, 96, 1)])
def test_sim_topk_low_precision_exact_on_stored_values(dtype, m, n, d, k):
    """Scores must equal the float64 dot products of the stored (rounded)
    values, scaled: products of BF16/FP16/FP8 are exact in FP32, so only the
    FP32 accumulation rounds."""
    _say(f"sim_topk dtype={dtype} m={m} n={n} d={d} k={k}")
    g = torch.Generator().manual_seed(m * 7 + n + d + k + dtype)
    q = torch.randn(m, d, generator=g)
    dd = torch.randn(n, d, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), dtype, zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), dtype, zero_mean=False)
    val = torch.empty(m, k, device=DEV)
    idx = torch.empty(m, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, val, idx, merge=False)
    _check_topk(val.cpu(), idx.cpu(), Q.to_float().cpu(), D.to_float().cpu(), k, tol=2e-5)


@pytest.mark.parametrize("dtype", [DTYPE_BF16, DTYPE_F16, DTYPE_FP8_E4M3])
@pytest.mark.parametrize("m,n,d,k", [(5, 256, 384, 16), (1000, 5000, 400, 16), (777, 300, 1000, 1),
                                     (130, 70000, 520, 8), (4096, 2048, 3600, 16), (300, 192, 256, 16),
                                     (257, 1000, 128, 4), (2000, 20000, 640, 16)])
def test_sim_topk_large_tile_kernel(dtype, m, n, d, k):
    """Shapes that take the 192 x 256 low-precision kernel (N >= 192, at
    least four 64-byte K stages): ragged M and N, a single tile, the minimum
    stage count (FP8 d=256, BF16/FP16 d=128), one and many parts."""
    _say(f"sim_topk large-tile dtype={dtype} m={m} n={n} d={d} k={k}")
    g = torch.Generator().manual_seed(m + n + d + k + dtype)
    q = torch.randn(m, d, generator=g)
    dd = torch.randn(n, d, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), dtype, zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), dtype, zero_mean=False)
    val = torch.empty(m, k, device=DEV)
    idx = torch.empty(m, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, val, idx, merge=False)
    _check_topk(val.cpu(), idx.cpu(), Q.to_float().cpu(), D.to_float().cpu(), k, tol=2e-5)


@pytest.mark.parametrize("m,n,k", [(300, 193, 16), (1000, 383, 16), (257, 5000 + 7, 16), (4096, 20000 + 100, 4),
                                   (129, 45000 + 191, 1)])
def test_sim_topk_two_workgroup_kernel_partial_last_tile(m, n, k):
    """FP8 rows of 256 bytes (the two-workgroup kernel): the last, partial
    dictionary tile is loaded as the dictionary's last 192 rows, so its
    leading rows are read a second time. Rows planted in that overlap with
    the largest scores of every query must be listed exactly once, in their
    place; lists equal the float64 brute force of the stored values."""
    _say(f"sim_topk two-workgroup partial tile m={m} n={n} k={k}")
    plan = kernels.sim_schedule(m, n, k, DTYPE_FP8_E4M3, 256, counts=False)
    assert plan["bt"] == 2
    g = torch.Generator().manual_seed(m + n + k)
    d = 256
    q = torch.randn(m, d, generator=g)
    dd = torch.randn(n, d, generator=g)
    lo, hi = max(n - 192, 0), (n // 192) * 192  # rows the last tile re-reads
    if hi > lo:
        plant = torch.arange(lo, hi, max((hi - lo) // 3, 1))[:3]
        dd[plant] *= 10.0  # in the lists of a good share of the queries
        dd[plant[0]] = 20.0 * q[0]  # first of query 0's list
    Q = kernels.prep_patterns(q.to(DEV), DTYPE_FP8_E4M3, zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), DTYPE_FP8_E4M3, zero_mean=False)
    val = torch.empty(m, k, device=DEV)
    idx = torch.empty(m, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, val, idx, merge=False)
    ix = idx.long().cpu()
    for r in range(m):
        assert len(set(ix[r].tolist())) == k, (r, ix[r].tolist())
    _check_topk(val.cpu(), idx.cpu(), Q.to_float().cpu(), D.to_float().cpu(), k, tol=2e-5)


@pytest.mark.parametrize("dtype,d,k", [(DTYPE_FP8_E4M3, 256, 16), (DTYPE_BF16, 128, 16), (DTYPE_FP8_E4M3, 384, 16),
                                       (DTYPE_FP8_E4M3, 256, 1), (DTYPE_F16, 256, 4)])
def test_sim_topk_large_tile_exact_ties_over_many_tiles(dtype, d, k):
    """A dictionary of identical row pairs (2j, 2j + 1), so every score ties
    exactly with its partner, and 65,536 queries, so that each work unit
    streams all 469 dictionary tiles (rows of at most 256 bytes: one part per
    query tile) or half of them (longer rows). Equal scores rank by smaller
    index: a pair in the list is adjacent with the smaller index first, and
    a pair cut by the K-th place keeps only its smaller index. Values equal
    the float64 products of the stored values (checked on 1,024 queries)."""
    _say(f"sim_topk exact ties dtype={dtype} d={d} k={k}")
    m, n = 65536, 2 * 45000
    g = torch.Generator(device=DEV).manual_seed(d + k + dtype)
    q = torch.randn(m, d, generator=g, device=DEV)
    half = torch.randn(n // 2, d, generator=g, device=DEV)
    dd = half.repeat_interleave(2, dim=0)
    Q = kernels.prep_patterns(q, dtype, zero_mean=False)
    D = kernels.prep_patterns(dd, dtype, zero_mean=False)
    val = torch.empty(m, k, device=DEV)
    idx = torch.empty(m, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, val, idx, merge=False)
    sel = torch.arange(0, m, 64, device=DEV)
    full = Q.to_float()[sel].double() @ D.to_float().double().T
    ref_v, _ = torch.topk(full, k, dim=1)
    tol = 2e-5 * ref_v.abs().max().item()
    assert torch.allclose(val[sel].double(), ref_v, atol=tol)
    ix = idx[sel].long()
    assert torch.allclose(full.gather(1, ix), ref_v, atol=tol)
    assert torch.all(val[:, 1:] <= val[:, :-1])
    ix_c = idx.long().cpu()
    for r in range(m):
        row = ix_c[r].tolist()
        present = set(row)
        for pos, i in enumerate(row):
            partner = i ^ 1
            if partner in present:
                # both in: adjacent, smaller index first
                j = row.index(partner)
                assert abs(j - pos) == 1 and (row[min(j, pos)] == min(i, partner)), (r, row)
            else:
                # cut by the K-th place: only the smaller index may stay
                assert i % 2 == 0, (r, row)


@pytest.mark.parametrize("dtype", [DTYPE_BF16, DTYPE_F16, DTYPE_FP8_E4M3])
def test_sim_topk_low_precision_streaming_mixes_kernels(dtype):
    """Chunks below 192 rows take the 128 x 128 kernel, larger ones the
    192 x 256 kernel; the streamed result equals one pass bit for bit."""
    _say(f"sim_topk low-precision streaming dtype={dtype}")
    g = torch.Generator().manual_seed(17 + dtype)
    q = torch.randn(300, 512, generator=g)
    dd = torch.randn(3000, 512, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), dtype, zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), dtype, zero_mean=False)
    k = 16
    v1 = torch.empty(300, k, device=DEV)
    i1 = torch.empty(300, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, v1, i1, merge=False)
    v2 = torch.full((300, k), float("-inf"), device=DEV)
    i2 = torch.full((300, k), -1, dtype=torch.int32, device=DEV)
    for s, c in ((0, 100), (100, 700), (800, 150), (950, 191), (1141, 192), (1333, 1667)):
        kernels.sim_topk(Q, D.narrow(s, c), k, v2, i2, index_base=s, merge=True)
    assert torch.equal(i1, i2) and torch.equal(v1, v2)


@pytest.mark.parametrize("dtype", [DTYPE_F32, DTYPE_BF16, DTYPE_FP8_E4M3])
@pytest.mark.parametrize("shape,renorm", [((60, 60), False), ((33, 47), True)])
def test_projection_through_quad_table_is_bit_identical(mp101, dtype, shape, renorm):
    _say(f"projection quad table dtype={dtype} shape={shape}")
    rays = O.normalized_rays(shape, (0.5, 0.45, 0.6))
    q = O.sample_ori_fz_laue(11, 500)[:300]
    mp = mp101.to(DEV)
    a = kernels.project(mp, rays.to(DEV), q.to(DEV), dtype, renorm_dirs=renorm)
    quads = kernels.mp_quad_table(mp)
    b = kernels.project(mp, rays.to(DEV), q.to(DEV), dtype, renorm_dirs=renorm, quads=quads)
    assert torch.equal(a.data, b.data)
    if a.scale is not None:
        assert torch.equal(a.scale, b.scale)


@pytest.mark.parametrize("n", [150, 2000])
def test_sim_topk_fp8_uses_the_exponent_of_row_scales(n):
    """FP8 row scales act as E8M0 block scales: only their power of two
    counts, in both similarity kernels (N < 192: 128 x 128, else 192 x 256).
    Scaling every scale by 1.5 (same exponent) must not change a bit of the
    result; doubling them doubles the scores exactly."""
    _say(f"sim_topk fp8 scale exponent n={n}")
    g = torch.Generator().manual_seed(21 + n)
    q = torch.randn(300, 520, generator=g)
    dd = torch.randn(n, 520, generator=g)
    Q = kernels.prep_patterns(q.to(DEV), DTYPE_FP8_E4M3, zero_mean=False)
    D = kernels.prep_patterns(dd.to(DEV), DTYPE_FP8_E4M3, zero_mean=False)
    assert torch.equal(torch.frexp(D.scale)[0], torch.full_like(D.scale, 0.5))  # powers of two
    k = 8
    v1 = torch.empty(300, k, device=DEV)
    i1 = torch.empty(300, k, dtype=torch.int32, device=DEV)
    kernels.sim_topk(Q, D, k, v1, i1, merge=False)
    D15 = type(D)(D.data, D.scale * 1.5, D.dtype, D.n_pix)
    v2 = torch.empty_like(v1)
    i2 = torch.empty_like(i1)
    kernels.sim_topk(Q, D15, k, v2, i2, merge=False)
    assert torch.equal(v1, v2) and torch.equal(i1, i2)
    D2 = type(D)(D.data, D.scale * 2.0, D.dtype, D.n_pix)
    kernels.sim_topk(Q, D2, k, v2, i2, merge=False)
    assert torch.equal(v2, 2.0 * v1) and torch.equal(i1, i2)
