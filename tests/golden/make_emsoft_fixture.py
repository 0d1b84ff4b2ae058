"""Write small EMsoft-layout master-pattern files for the HDF5 reader tests.

    python tests/golden/make_emsoft_fixture.py

Each file holds the four datasets io/read_master_pattern.py reads, with the
shapes and types EMsoft writes: EMData/MCOpenCL/accum_e (int32 counts,
(H, W, n_energies) as the reference's comment has it, summed over the first
two axes), EMData/EBSDmaster/mLPNH and mLPSH (float32, (1, n_energies, S, S)
square-Lambert hemispheres) and CrystalData/SpaceGroupNumber (int32, (1,)).
The HDF5 files are written by the HDF5 distribution's own h5import tool from
raw little-endian arrays, never by this package's reader; the same arrays
are saved next to them (npz) so the tests can compare what the reader returns
with what was stored. Per-energy patterns are Laue-symmetric Gaussian band
patterns (the package's synthetic generator) whose band width and contrast
change with energy, so the energy weighting is visible in the result.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

H5IMPORT = shutil.which("h5import") or "/opt/conda/bin/h5import"

FIXTURES = {
    # name: (space group, Laue group of the synthetic bands, hemisphere size, energies, seed)
    "emsoft_mp_m3m": (225, 11, 65, 5, 3),
    "emsoft_mp_6mmm_1e": (191, 9, 49, 1, 4),
}


def _dataset(tmp, arr, path, cls):
    raw = os.path.join(tmp, path.replace("/", "_") + ".bin")
    cfg = raw + ".cfg"
    arr = np.ascontiguousarray(arr)
    arr.astype("<f4" if cls == "FP" else "<i4").tofile(raw)
    with open(cfg, "w") as f:
        f.write(f"PATH {path}\nINPUT-CLASS {cls}\nINPUT-SIZE 32\nINPUT-BYTE-ORDER LE\nRANK {arr.ndim}\n"
                f"DIMENSION-SIZES {' '.join(str(d) for d in arr.shape)}\nOUTPUT-CLASS {cls}\nOUTPUT-SIZE 32\n"
                f"OUTPUT-ARCHITECTURE {'IEEE' if cls == 'FP' else 'STD'}\nOUTPUT-BYTE-ORDER LE\n")
    return [raw, "-c", cfg]


def main():
    import torch

    from ebsdtorch_amd import synthetic

    for name, (sg, laue, size, n_e, seed) in FIXTURES.items():
        rng = np.random.default_rng(seed)
        nh, sh = [], []
        for e in range(n_e):
            mp = synthetic.band_master_pattern(size, laue, 12, 3.0 + 1.5 * e, seed=seed).numpy()
            mp = (0.2 + 0.15 * e) + (1.0 - 0.1 * e) * mp
            nh.append(mp[:, :size])
            sh.append(mp[:, size:])
        mlpnh = np.stack(nh)[None].astype(np.float32)
        mlpsh = np.stack(sh)[None].astype(np.float32)
        accum = rng.integers(0, 400, size=(9, 9, n_e)).astype(np.int32)
        accum[..., 0] += 50  # the lowest-energy bin is the most populated
        group = np.array([sg], dtype=np.int32)
        out = os.path.join(HERE, name + ".h5")
        if os.path.exists(out):
            os.remove(out)
        with tempfile.TemporaryDirectory() as tmp:
            args = [H5IMPORT]
            args += _dataset(tmp, accum, "EMData/MCOpenCL/accum_e", "IN")
            args += _dataset(tmp, mlpnh, "EMData/EBSDmaster/mLPNH", "FP")
            args += _dataset(tmp, mlpsh, "EMData/EBSDmaster/mLPSH", "FP")
            args += _dataset(tmp, group, "CrystalData/SpaceGroupNumber", "IN")
            subprocess.run(args + ["-o", out], check=True)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), accum_e=accum, mLPNH=mlpnh, mLPSH=mlpsh,
                            space_group=group)
        print(f"wrote {out} ({os.path.getsize(out)} bytes)")
    del torch


if __name__ == "__main__":
    main()
