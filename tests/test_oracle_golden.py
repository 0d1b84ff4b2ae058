"""Pin the CPU oracle against the reference's golden vectors and against
closed forms / invariants that a plausible restatement mistake would break."""

import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import di_oracle as O

T = torch.tensor


# ---------------------------------------------------------------- sampler


def test_cubochoric_grid_matches_reference(golden):
    got = O.so3_cubochoric_grid(6).numpy()
    np.testing.assert_array_equal(got, golden["grid_edge6"])


@pytest.mark.parametrize("key,laue,n", [("fz_m3m_2000", 11, 2000), ("fz_6mmm_2000", 9, 2000)])
def test_fz_sampler_matches_reference(golden, key, laue, n):
    got = O.sample_ori_fz_laue(laue, n).numpy()
    np.testing.assert_array_equal(got, golden[key])


def test_fz_angle_sampler_matches_reference(golden):
    np.testing.assert_array_equal(O.sample_ori_fz_laue_angle(11, 8.0).numpy(), golden["fz_angle_m3m_8deg"])


@pytest.mark.parametrize("laue", range(1, 12))
def test_fz_membership_matches_reference(golden, laue):
    q = T(golden["fz_in_q"])
    np.testing.assert_array_equal(O.ori_in_fz_laue(q, laue).numpy(), golden[f"fz_in_{laue}"])
    np.testing.assert_array_equal(O.laue_elements(laue).numpy(), golden[f"laue_ops_{laue}"])


@pytest.mark.parametrize("laue", [3, 4, 5, 9, 10, 11])
def test_fz_membership_equals_brute_force(laue):
    # closed-form FZ test == "identity gives the largest |w|" (laue_fz_ori.py:108-141)
    g = torch.Generator().manual_seed(laue)
    q = torch.randn(4000, 4, generator=g, dtype=torch.float64)
    q = O.qu_std(q / q.norm(dim=-1, keepdim=True))
    np.testing.assert_array_equal(O.ori_in_fz_laue(q, laue).numpy(), O.ori_in_fz_laue_brute(q, laue).numpy())


@pytest.mark.parametrize("laue", range(1, 12))
def test_laue_groups_are_closed(laue):
    ops = O.laue_elements(laue)
    assert torch.allclose(ops.norm(dim=-1), torch.ones(len(ops), dtype=torch.float64))
    prods = O.qu_prod(ops[:, None], ops[None, :]).reshape(-1, 1, 4)
    # every product is +-some element
    d = (prods - ops[None]).abs().sum(-1).min(dim=-1).values
    d2 = (prods + ops[None]).abs().sum(-1).min(dim=-1).values
    assert torch.all(torch.minimum(d, d2) < 1e-6)
    assert len(ops) == O.get_laue_mult(laue) // 2


def test_sampler_invariants_and_count():
    s = O.sample_ori_fz_laue(11, 10000)
    assert torch.allclose(s.norm(dim=-1), torch.ones(len(s)), atol=2e-6)
    assert torch.all(s[:, 0] >= 0)
    assert torch.all(O.ori_in_fz_laue(s, 11))
    edge = int((10000 * 0.5 * 48) ** (1 / 3))
    assert edge == 62
    # FZ holds 2/48 of SO(3) (inclusive boundaries add a few points)
    assert abs(len(s) - edge ** 3 * 2 / 48) / len(s) < 0.02


def test_cubochoric_grid_is_haar_uniform():
    # the cubochoric map is volume preserving, so grid moments approach the
    # Haar moments of SO(3): E[cos w] = -1/2, E[cos 2w] = 0, E[q_i^2] = 1/4
    q = O.so3_cubochoric_grid(40).double()
    w = 2 * torch.arccos(q[:, 0].clamp(-1, 1))
    assert abs(torch.cos(w).mean().item() + 0.5) < 2e-3
    assert abs(torch.cos(2 * w).mean().item()) < 2e-3
    assert torch.allclose((q[:, 1:] ** 2).mean(0), torch.full((3,), 0.25, dtype=torch.float64), atol=1e-3)
    assert torch.allclose(q.norm(dim=-1), torch.ones(len(q), dtype=torch.float64), atol=1e-6)


def test_disorientation_matches_reference(golden):
    got = O.disori_angle_laue(T(golden["dis_q1"]), T(golden["dis_q2"]), 11).numpy()
    np.testing.assert_allclose(got, golden["dis_m3m"], atol=2e-6)
    assert np.all(got <= math.radians(62.81) + 1e-6)


# ---------------------------------------------------------------- geometry


def test_geometry_matches_reference(golden):
    got = O.detector_coords((60, 60), proj_center=(0.5, 0.5, 0.6)).numpy()
    np.testing.assert_array_equal(got, golden["geom_60_coords"])
    for b in (1, 2):
        got = O.detector_coords((12, 16), (3.0, 68.0, 5.0), (0.45, 0.55, 0.7), (b, b)).numpy()
        np.testing.assert_array_equal(got, golden[f"geom_12x16_b{b}_coords"])


def test_geometry_backprojection_round_trip():
    c = O.detector_coords((12, 16), (3.0, 68.0, 5.0), (0.45, 0.55, 0.7)).double()
    pix = O.backproject_to_detector(c.float(), (12, 16), (3.0, 68.0, 5.0), (0.45, 0.55, 0.7)).double()
    i = torch.linspace(0, 12, 12, dtype=torch.float64)
    j = torch.linspace(0, 16, 16, dtype=torch.float64)
    expect = torch.stack(torch.meshgrid(i, j, indexing="ij"), -1).view(-1, 2)
    assert torch.allclose(pix, expect, atol=1e-3)


def test_geometry_pattern_centre_closed_form():
    # the ray through the pattern centre pixel (pcx W, pcy H) is the detector
    # normal scaled by pcz H: sample-frame distance = pcz * H
    H, W, pc = 60, 60, (0.5, 0.5, 0.6)
    rot, tr = O.detector_to_sample_transform((H, W), (0.0, 70.0, 0.0), pc)
    p = torch.tensor([pc[1] * H, pc[0] * W, 0.0])
    s = (rot[0] @ p + tr[0]).double()
    # closed form: detector at distance pcz*H from the source along the
    # detector normal, which is (cos 20deg, 0, -sin 20deg) rotated by the tilt
    assert abs(s.norm().item() - pc[2] * H) < 0.5


# ---------------------------------------------------------------- rotation


def test_qu_apply_matches_reference(golden):
    got = O.qu_apply(T(golden["qa_q"]), T(golden["qa_p"])).numpy()
    np.testing.assert_array_equal(got, golden["qa_out"])


def test_qu_apply_is_rotation_matrix_product():
    g = torch.Generator().manual_seed(3)
    q = torch.randn(100, 4, generator=g, dtype=torch.float64)
    q = q / q.norm(dim=-1, keepdim=True)
    p = torch.randn(100, 3, generator=g, dtype=torch.float64)
    expect = (O.qu2om(q) @ p[..., None])[..., 0]
    assert torch.allclose(O.qu_apply(q, p), expect, atol=1e-12)
    # 90 deg about z maps x -> y
    qz = torch.tensor([math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)], dtype=torch.float64)
    out = O.qu_apply(qz, torch.tensor([1.0, 0, 0], dtype=torch.float64))
    assert torch.allclose(out, torch.tensor([0, 1.0, 0], dtype=torch.float64), atol=1e-15)


# ---------------------------------------------------------------- Lambert


def test_lambert_matches_reference(golden):
    got = O.rosca_lambert_side_by_side(T(golden["lam_pts"])).numpy()
    np.testing.assert_array_equal(got, golden["lam_out"])


def test_lambert_special_points():
    pts = torch.tensor([[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0]], dtype=torch.float64)
    out = O.rosca_lambert_side_by_side(pts)
    expect = torch.tensor([[-0.5, 0.0], [0.5, 0.0], [0.0, 0.0], [-0.5, 1.0]], dtype=torch.float64)
    assert torch.allclose(out, expect, atol=1e-12)


def test_lambert_inverse_round_trip_and_equal_area():
    g = torch.Generator().manual_seed(5)
    v = torch.randn(200000, 3, generator=g, dtype=torch.float64)
    v = v / v.norm(dim=-1, keepdim=True)
    nh = v[v[:, 2] > 0]
    sq = O.rosca_lambert_side_by_side(nh)
    # undo the side-by-side shift: u' = u/2 - 0.5 for the north half
    flat = torch.stack([(sq[:, 0] + 0.5) * 2, sq[:, 1]], -1)
    back = O.inv_rosca_lambert(flat)
    assert torch.allclose(back, nh, atol=1e-9)
    # equal area: uniform sphere points give a uniform square histogram
    h, _, _ = np.histogram2d(flat[:, 0].numpy(), flat[:, 1].numpy(), bins=8, range=[[-1, 1], [-1, 1]])
    assert h.std() / h.mean() < 0.03


# ---------------------------------------------------------------- sampling


def _brute_bilinear_border(img, gx, gy):
    H, W = img.shape
    out = np.empty(len(gx))
    for n in range(len(gx)):
        ix = min(max(((gx[n] + 1) * W - 1) / 2, 0.0), W - 1)
        iy = min(max(((gy[n] + 1) * H - 1) / 2, 0.0), H - 1)
        x0, y0 = int(math.floor(ix)), int(math.floor(iy))
        acc = 0.0
        for yy, wy in ((y0, 1 - (iy - y0)), (y0 + 1, iy - y0)):
            for xx, wx in ((x0, 1 - (ix - x0)), (x0 + 1, ix - x0)):
                if 0 <= xx < W and 0 <= yy < H:
                    acc += wy * wx * img[yy, xx]
        out[n] = acc
    return out


def test_grid_sample_semantics_brute_force():
    g = torch.Generator().manual_seed(7)
    img = torch.rand(5, 10, generator=g, dtype=torch.float64)
    pts = torch.rand(400, 2, generator=g, dtype=torch.float64) * 2.4 - 1.2  # includes out-of-range
    out = F.grid_sample(img[None, None], pts.view(1, 1, -1, 2), mode="bilinear", padding_mode="border",
                        align_corners=False).view(-1).numpy()
    brute = _brute_bilinear_border(img.numpy(), pts[:, 0].numpy(), pts[:, 1].numpy())
    np.testing.assert_allclose(out, brute, atol=1e-12)


def test_interpolate_matches_reference(golden):
    got = O.interpolate(T(golden["mp33"]), T(golden["interp_pts"]), normalize_coords=False).numpy()
    np.testing.assert_array_equal(got, golden["interp_out"])


MODES = [(m, p, a) for m in ("bilinear", "nearest", "bicubic") for p in ("zeros", "border", "reflection")
         for a in (False, True)]


@pytest.mark.parametrize("mode,pad,align", MODES)
def test_interpolate_modes_match_reference(golden, mode, pad, align):
    got = O.interpolate(T(golden["interp_modes_mp"]), T(golden["interp_modes_pts"]), normalize_coords=True,
                        mode=mode, padding_mode=pad, align_corners=align).numpy()
    np.testing.assert_array_equal(got, golden[f"interp_{mode}_{pad}_{int(align)}"])


def _source_index(g, size, align):
    return (g + 1) / 2 * (size - 1) if align else ((g + 1) * size - 1) / 2


def _pad_index(x, size, pad, align):
    """Continuous source coordinate after padding (None: zeros padding)."""
    if pad == "border":
        return min(max(x, 0.0), size - 1)
    if pad == "reflection":
        lo, hi = (0.0, size - 1.0) if align else (-0.5, size - 0.5)
        span = hi - lo
        if span == 0:
            return 0.0
        d = abs(x - lo) % (2 * span)
        x = lo + (d if d <= span else 2 * span - d)
        return min(max(x, 0.0), size - 1)
    return x


def _tap(img, xi, yi):
    H, W = img.shape
    return img[yi, xi] if 0 <= xi < W and 0 <= yi < H else 0.0


def _cubic(t, A=-0.75):
    t = abs(t)
    if t <= 1:
        return ((A + 2) * t - (A + 3)) * t * t + 1
    if t < 2:
        return ((A * t - 5 * A) * t + 8 * A) * t - 4 * A
    return 0.0


def _brute_grid_sample(img, gx, gy, mode, pad, align):
    """grid_sample from its definitions: source index, padding of the
    coordinate (bilinear, nearest) or of each tap (bicubic), taps outside the
    image read as 0, Keys' cubic convolution with A = -0.75."""
    H, W = img.shape
    out = np.empty(len(gx))
    for n in range(len(gx)):
        x = _source_index(gx[n], W, align)
        y = _source_index(gy[n], H, align)
        if mode == "bicubic":
            x0, y0 = math.floor(x), math.floor(y)
            acc = 0.0
            for j in range(-1, 3):
                for i in range(-1, 3):
                    xi, yi = _pad_index(x0 + i, W, pad, align), _pad_index(y0 + j, H, pad, align)
                    acc += _cubic(x - (x0 + i)) * _cubic(y - (y0 + j)) * _tap(img, int(xi), int(yi)) \
                        if float(xi).is_integer() and float(yi).is_integer() else 0.0
            out[n] = acc
            continue
        x, y = _pad_index(x, W, pad, align), _pad_index(y, H, pad, align)
        if mode == "nearest":
            out[n] = _tap(img, int(np.rint(x)), int(np.rint(y)))
            continue
        x0, y0 = math.floor(x), math.floor(y)
        acc = 0.0
        for yy, wy in ((y0, 1 - (y - y0)), (y0 + 1, y - y0)):
            for xx, wx in ((x0, 1 - (x - x0)), (x0 + 1, x - x0)):
                acc += wy * wx * _tap(img, xx, yy)
        out[n] = acc
    return out


@pytest.mark.parametrize("mode,pad,align", MODES)
def test_grid_sample_modes_brute_force(mode, pad, align):
    """The library routine the oracle calls, in every mode, against the
    definitions (float64; points up to 20% outside [-1, 1])."""
    g = torch.Generator().manual_seed(11)
    img = torch.rand(5, 10, generator=g, dtype=torch.float64)
    pts = torch.rand(300, 2, generator=g, dtype=torch.float64) * 2.4 - 1.2
    if mode == "nearest":  # keep clear of rounding ties
        pts = pts + 1e-7
    out = F.grid_sample(img[None, None], pts.view(1, 1, -1, 2), mode=mode, padding_mode=pad,
                        align_corners=align).view(-1).numpy()
    brute = _brute_grid_sample(img.numpy(), pts[:, 0].numpy(), pts[:, 1].numpy(), mode, pad, align)
    np.testing.assert_allclose(out, brute, atol=1e-12)


# ---------------------------------------------------------------- top-k


def test_chunked_knn_matches_reference(golden):
    data, query = T(golden["knn_data"]), T(golden["knn_query"])
    kn = O.ChunkedKNN(300, 50, 4)
    for s in range(0, 300, 64):
        kn.set_data_chunk(data[s:s + 64])
        for qs in range(0, 50, 16):
            kn.query_chunk(query[qs:qs + 16], qs)
    idx, val = kn.retrieve_topk()
    np.testing.assert_array_equal(idx.numpy(), golden["knn_idx"])
    np.testing.assert_array_equal(val.numpy(), golden["knn_val"])


def test_chunked_knn_equals_float64_brute_force_and_is_chunk_invariant():
    g = torch.Generator().manual_seed(11)
    data = torch.randn(513, 24, generator=g)
    query = torch.randn(77, 24, generator=g)
    full = (query.double() @ data.double().T)
    ref_val, ref_idx = torch.topk(full, 5, dim=1)
    for dcs, qcs in ((1000, 1000), (100, 16), (9, 5)):
        kn = O.ChunkedKNN(513, 77, 5)
        for s in range(0, 513, dcs):
            kn.set_data_chunk(data[s:s + dcs])
            if qcs > 77:
                kn.query_all(query)
            else:
                for qs in range(0, 77, qcs):
                    kn.query_chunk(query[qs:qs + qcs], qs)
        idx, val = kn.retrieve_topk()
        assert torch.equal(idx, ref_idx)
        assert torch.allclose(val.double(), ref_val, atol=1e-4)


def test_chunked_knn_int8_path_finds_planted_matches_within_quantisation_error():
    """The reference's CPU default (int8 quantize_dynamic of each data chunk,
    knn.py:52-72): queries that are noisy copies of data rows find their
    source row, global indices survive the chunked merge, and every reported
    score is the float score of that pair up to the int8 steps (weights: the
    chunk's max|w| / 127; activations: the query row range / 255)."""
    g = torch.Generator().manual_seed(12)
    data = torch.randn(300, 64, generator=g)
    src = torch.randperm(300, generator=g)[:40]
    query = data[src] + 0.05 * torch.randn(40, 64, generator=g)
    kn = O.ChunkedKNN(300, 40, 3, quantized=True)
    for s in range(0, 300, 128):
        kn.set_data_chunk(data[s:s + 128])
        for qs in range(0, 40, 16):
            kn.query_chunk(query[qs:qs + 16], qs)
    idx, val = kn.retrieve_topk()
    assert torch.equal(idx[:, 0], src)
    assert bool((val[:, :-1] >= val[:, 1:]).all())
    exact = (query.double() @ data.double().T).gather(1, idx)
    step_w = float(data.abs().max()) / 127
    step_x = float((query.max() - query.min())) / 255
    bound = 64 * (float(query.abs().max()) * step_w + float(data.abs().max()) * step_x + step_w * step_x)
    assert float((val.double() - exact).abs().max()) < bound
    # the quantised scores are not the float ones (the int8 path really ran)
    assert float((val.double() - exact).abs().max()) > 1e-4


def test_query_before_data_raises():
    kn = O.ChunkedKNN(10, 4, 1)
    with pytest.raises(ValueError):
        kn.query_all(torch.zeros(4, 3))


# ---------------------------------------------------------------- covariance


def test_online_cov_matches_reference_and_numpy(golden):
    x = T(golden["cov_x"])
    oc = O.OnlineCovMatrix(20)
    for s in range(0, 300, 100):
        oc.update(x[s:s + 100])
    np.testing.assert_array_equal(oc.get_covmat().numpy(), golden["cov_out"])
    np.testing.assert_allclose(oc.get_covmat().numpy(), np.cov(golden["cov_x"].astype(np.float64).T), atol=2e-5)


def test_eigh_residuals():
    g = torch.Generator().manual_seed(2)
    a = torch.randn(200, 30, generator=g, dtype=torch.float64)
    oc = O.OnlineCovMatrix(30, covmat_dtype=torch.float64, delta_dtype=torch.float64)
    oc.update(a)
    c = oc.get_covmat()
    ev = oc.get_eigenvectors()
    lam = torch.linalg.eigvalsh(c)
    assert torch.allclose(c @ ev, ev * lam, atol=1e-10)
    assert torch.allclose(ev.T @ ev, torch.eye(30, dtype=torch.float64), atol=1e-10)
    assert torch.all(lam[1:] >= lam[:-1])


# ---------------------------------------------------------------- end to end


def _rays8():
    return O.normalized_rays((8, 8), (0.5, 0.5, 0.6))


@pytest.mark.parametrize("tag,ecs", [("chunked", 5), ("all", 64)])
def test_dictionary_index_matches_reference(golden, tag, ecs):
    dico = T(golden[f"di_{tag}_dict"])
    exp = T(golden["di_exp"]).view(16, -1)
    idx, val = O.dictionary_index(T(golden["mp33"]), _rays8(), exp, dico, top_k=3,
                                  dictionary_chunk_size=37, experiment_chunk_size=ecs)
    np.testing.assert_array_equal(dico[idx].numpy(), golden[f"di_{tag}_quats"])
    np.testing.assert_array_equal(val.numpy(), golden[f"di_{tag}_metrics"])


def test_pca_dictionary_index_matches_reference(golden):
    dico = T(golden["pca_dict"])
    exp = T(golden["di_exp"]).view(16, -1)
    coords = O.detector_coords((8, 8), proj_center=(0.5, 0.5, 0.6))
    idx, val = O.pca_dictionary_index(T(golden["mp33"]), coords, exp, dico, dico, 20, top_k=2,
                                      dictionary_chunk_size=37, experiment_chunk_size=5, reference_slice=True)
    np.testing.assert_array_equal(dico[idx].numpy(), golden["pca_quats"])
    np.testing.assert_allclose(val.numpy(), golden["pca_metrics"], rtol=1e-4, atol=1e-5)
