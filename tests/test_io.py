"""The EMsoft master-pattern reader (ebsdtorch_amd.io, io/read_master_pattern.py)
on CPU: files written by HDF5's own h5import from known arrays
(tests/golden/make_emsoft_fixture.py) read back exactly; the energy
weighting against the oracle's restatement and against closed forms."""

import os

import numpy as np
import pytest
import torch

from ebsdtorch_amd import io as EIO
from ebsdtorch_amd.master_pattern import space_group_to_laue
from oracle import io_oracle as IO

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = [("emsoft_mp_m3m", 225, 11), ("emsoft_mp_6mmm_1e", 191, 9)]


def _stored(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name,sg,laue", FIX)
def test_reader_returns_the_stored_arrays(name, sg, laue):
    got = EIO.read_emsoft_arrays(os.path.join(GOLD, name + ".h5"))
    ref = _stored(name)
    for k in ("accum_e", "mLPNH", "mLPSH"):
        assert got[k].dtype == np.float32 and got[k].shape == ref[k].shape
        np.testing.assert_array_equal(got[k], ref[k].astype(np.float32))
    assert got["space_group"] == sg


@pytest.mark.parametrize("name,sg,laue", FIX)
def test_read_master_pattern_matches_oracle(name, sg, laue):
    mp = EIO.read_master_pattern(os.path.join(GOLD, name + ".h5"))
    ref = _stored(name)
    want = IO.master_pattern_from_arrays(ref["accum_e"], ref["mLPNH"], ref["mLPSH"])
    got = mp.master_pattern.numpy()
    assert got.dtype == np.float32 and got.shape == want.shape == (ref["mLPNH"].shape[-2], 2 * ref["mLPNH"].shape[-1])
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7)
    assert mp.laue_group == laue == IO.laue_of_space_group(sg)


def test_space_group_to_laue_all_groups():
    for sg in range(1, 231):
        assert space_group_to_laue(sg) == IO.laue_of_space_group(sg)
    # the reference's two ranges that differ from the International Tables
    assert space_group_to_laue(194) == 10 and space_group_to_laue(191) == 9
    assert [space_group_to_laue(g) for g in (148, 149, 155, 156)] == [6, 6, 6, 7]
    with pytest.raises(ValueError):
        space_group_to_laue(0)
    with pytest.raises(ValueError):
        space_group_to_laue(231)


def test_energy_weighting_closed_forms():
    rng = np.random.default_rng(0)
    a = rng.random((8, 8), dtype=np.float32)
    b = rng.random((8, 8), dtype=np.float32)
    c = rng.random((8, 8), dtype=np.float32)
    # equal patterns at every energy: the pattern itself (pdf sums to one)
    same = np.stack([a, a, a])[None]
    hist = rng.integers(1, 50, size=(4, 4, 3)).astype(np.float32)
    out = EIO.energy_weighted_master_pattern(hist, same, same * 2)
    np.testing.assert_allclose(out[:, :8], a, rtol=1e-6)
    np.testing.assert_allclose(out[:, 8:], 2 * a, rtol=1e-6)  # south half on the right
    # a histogram with counts in one energy only selects that energy
    one = np.zeros((3, 5, 3), np.float32)
    one[..., 1] = 7
    out = EIO.energy_weighted_master_pattern(one, np.stack([a, b, c])[None], np.stack([c, a, b])[None])
    np.testing.assert_array_equal(out, np.concatenate([b, a], 1))
    # explicit weights: counts 1 and 3 over the (H, W) histogram -> 0.25 a + 0.75 b
    h2 = np.zeros((2, 2, 2), np.float32)
    h2[0, 0] = (1, 3)
    out = EIO.energy_weighted_master_pattern(h2, np.stack([a, b])[None], np.stack([b, a])[None])
    np.testing.assert_allclose(out, np.concatenate([0.25 * a + 0.75 * b, 0.25 * b + 0.75 * a], 1), rtol=1e-6)
    # one energy stored as (1, 1, H, W): the pattern itself
    out = EIO.energy_weighted_master_pattern(hist[..., :1], a[None, None], b[None, None])
    np.testing.assert_allclose(out, np.concatenate([a, b], 1), rtol=1e-6)


def test_missing_file_and_dataset_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        EIO.read_master_pattern(str(tmp_path / "nope.h5"))
    bad = tmp_path / "bad.h5"
    bad.write_bytes(b"not an hdf5 file")
    with pytest.raises(OSError):
        EIO.read_master_pattern(str(bad))
    with EIO._File(os.path.join(GOLD, "emsoft_mp_m3m.h5")) as f:
        with pytest.raises(KeyError):
            f.read("EMData/EBSDmaster/nothing")
        with pytest.raises(KeyError):
            f.read("NoGroup/x")


def test_master_pattern_tensor_device_default_cpu():
    mp = EIO.read_master_pattern(os.path.join(GOLD, "emsoft_mp_m3m.h5"))
    assert mp.master_pattern.device == torch.device("cpu")
