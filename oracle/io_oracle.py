"""Test infrastructure only (the checker of ebsdtorch_amd.io; never imported
by the product): a CPU restatement of io/read_master_pattern.py:18-52 on
arrays already read from the file.

The energy pdf is the Monte-Carlo histogram accum_e summed over its first two
axes and normalised (:37-40); the per-energy hemispheres mLPNH / mLPSH,
(1, n_energies, H, W) squeezed, are weighted by it and summed over energies
(:43-44), in float32 as the reference's astype(np.float32) arrays; north and
south are concatenated along the last axis (:48), and the space group number
(:33) picks the Laue group through symmetry_classes.py:4-30. Written as an
explicit accumulation over energies (not the reference's broadcast sum) so
that it is a second expression of the same arithmetic.

Parity pinned by tests/test_io.py: the stored-array round trip through
h5import-written files, and closed forms (equal energies give the pattern
itself, a one-hot histogram selects one energy, explicit two-energy weights,
hemisphere order, the single-energy squeeze).
"""

from __future__ import annotations

import numpy as np


def laue_of_space_group(sg: int) -> int:
    """symmetry_classes.py:4-30: the reference's table of Laue groups by
    space-group number, range by range. Two of its ranges differ from the
    International Tables: 149-155 (point group 32) map to Laue group 6
    instead of 7, and 194 (P6_3/mmc) to 10 instead of 9; the drop-in keeps
    the reference's values."""
    ranges = [(1, 2, 1), (3, 15, 2), (16, 74, 3), (75, 88, 4), (89, 142, 5), (143, 155, 6), (156, 167, 7),
              (168, 176, 8), (177, 193, 9), (194, 206, 10), (207, 230, 11)]
    for lo, hi, lid in ranges:
        if lo <= sg <= hi:
            return lid
    raise ValueError("space group outside 1-230")


def master_pattern_from_arrays(accum_e, mlpnh, mlpsh):
    e = np.asarray(accum_e, dtype=np.float32)
    p = np.zeros(e.shape[-1], dtype=np.float32)
    for i in range(e.shape[0]):
        for j in range(e.shape[1]):
            p += e[i, j]
    pdf = p / p.sum()
    north = np.asarray(mlpnh, dtype=np.float32).reshape(-1, *np.asarray(mlpnh).shape[-2:])
    south = np.asarray(mlpsh, dtype=np.float32).reshape(-1, *np.asarray(mlpsh).shape[-2:])
    nh = np.zeros(north.shape[1:], dtype=np.float32)
    sh = np.zeros(south.shape[1:], dtype=np.float32)
    for k in range(north.shape[0]):
        nh += pdf[k] * north[k]
        sh += pdf[k] * south[k]
    return np.concatenate([nh, sh], axis=1)
