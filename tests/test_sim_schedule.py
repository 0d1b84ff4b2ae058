"""The persistent similarity kernels' work plan (ebsd_sim_schedule), walked
on the host with the kernels' own block -> work mapping: every (query tile,
dictionary tile) pair is scored by exactly one workgroup and every partial
top-k list that the merge reads is written exactly once, for all three
kernels, both groupings of the 192 x 256 one and the shapes of the benchmark
configurations. No GPU is needed."""

import ctypes

import numpy as np
import pytest

from ebsdtorch_amd import _lib
from ebsdtorch_amd.kernels import padded_width, sim_schedule

F32, BF16, F16, FP8 = _lib.DTYPE_F32, _lib.DTYPE_BF16, _lib.DTYPE_F16, _lib.DTYPE_FP8_E4M3

CASES = [
    # (queries, dictionary rows, pattern pixels, dtype, top-k)
    (1, 1, 64, F32, 1),
    (1, 1, 64, F32, 16),
    (17, 129, 100, F32, 16),
    (128, 995_277, 3600, F32, 16),        # one query tile: the XCDs split the dictionary
    (1000, 50_000, 3600, F32, 16),        # fewer query groups than XCDs
    (65_536, 995_277, 3600, F32, 16),     # full-pattern FP32 at 60x60
    (65_536, 995_277, 6400, F32, 16),     # 80x80
    (4096, 333_227, 3600, F32, 16),       # cubic 2.4 degree dictionary
    (300, 191, 3600, FP8, 16),            # below one 192-row tile: 128 x 128 kernel
    (300, 192, 3600, FP8, 16),            # one 192-row tile
    (300, 5000, 128, FP8, 16),            # 128-byte rows: too few stages, 128 x 128 kernel
    (65_536, 995_277, 256, FP8, 16),      # 256 principal coordinates: two workgroups per CU
    (300, 5000, 256, FP8, 16),            # ... with fewer query groups than XCDs
    (65_536, 995_277, 128, BF16, 16),     # 256-byte BF16 rows: resident queries, one workgroup per CU
    (65_536, 995_277, 448, FP8, 16),
    (65_536, 995_277, 3600, FP8, 16),     # full-pattern FP8 at 60x60
    (65_536, 995_277, 3600, BF16, 16),
    (65_536, 995_277, 6400, FP8, 1),
    (65_536, 986_945, 14_400, FP8, 16),   # 120x120
    (5, 995_277, 3600, F16, 16),
    (2049, 7777, 1000, BF16, 16),
]


@pytest.mark.parametrize("m,n,n_pix,dtype,topk", CASES)
def test_every_tile_pair_and_partial_list_once(m, n, n_pix, dtype, topk):
    ld = padded_width(n_pix, dtype)
    s = sim_schedule(m, n, topk, dtype, ld)
    eb = {F32: 4, BF16: 2, F16: 2, FP8: 1}[dtype]
    # which kernel: a large-tile one for low precision with >= 192 rows and
    # >= 4 stages of 64 bytes; FP8 rows of exactly 256 bytes take the
    # 192 x 128 kernel with two workgroups per CU (2), the others the
    # 192 x 256 one (1)
    large = dtype != F32 and n >= 192 and ld * eb // 64 >= 4
    two = large and dtype == FP8 and ld * eb == 256
    assert s["bt"] == (2 if two else 1 if large else 0)
    assert (s["q_rows"], s["d_rows"]) == ((128, 192) if two else (256, 192) if large else (128, 128))
    assert s["n_qtiles"] == -(-m // s["q_rows"]) and s["n_dtiles"] == -(-n // s["d_rows"])
    # the grid is 8 XCDs x (A x B) slots
    assert s["grid"] == 8 * s["A"] * s["B"]
    assert s["n_parts"] == s["B"] * s["R"]
    assert s["k_list"] == (1 if topk == 1 else 16)
    assert np.all(s["tile_visits"] == 1), np.argwhere(s["tile_visits"] != 1)[:5]
    assert np.all(s["list_writes"] == 1), np.argwhere(s["list_writes"] != 1)[:5]
    # the partial lists fit the workspace the library asks for
    nbytes = ctypes.c_size_t()
    _lib.check(_lib.load().ebsd_sim_topk_workspace(m, n, topk, dtype, ctypes.byref(nbytes)), "workspace")
    assert s["n_parts"] * s["k_list"] * m * 8 <= nbytes.value


def test_schedule_argument_errors():
    lib = _lib.load()
    info = (ctypes.c_int * _lib.EBSD_SCHEDULE_INFO)()
    assert lib.ebsd_sim_schedule(10, 10, 16, F32, 64, None, None, 0, None, 0) == _lib.EBSD_ERR_NULL
    assert lib.ebsd_sim_schedule(0, 10, 16, F32, 64, info, None, 0, None, 0) == _lib.EBSD_ERR_SHAPE
    assert lib.ebsd_sim_schedule(10, 10, 17, F32, 64, info, None, 0, None, 0) == _lib.EBSD_ERR_TOPK
    assert lib.ebsd_sim_schedule(10, 10, 16, 9, 64, info, None, 0, None, 0) == _lib.EBSD_ERR_DTYPE
    assert lib.ebsd_sim_schedule(10, 10, 16, F32, 33, info, None, 0, None, 0) == _lib.EBSD_ERR_SHAPE
    short = (ctypes.c_int * 1)()
    assert lib.ebsd_sim_schedule(1000, 1000, 16, F32, 64, info, short, 1, None, 0) == _lib.EBSD_ERR_WORKSPACE
