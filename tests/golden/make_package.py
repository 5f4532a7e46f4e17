"""Write the small dataset package under tests/golden/package_blobs_u8_44x40x30_bs8/.

A 44 x 40 x 30 u8 volume of seeded Gaussian blobs (ragged against the 8-voxel
brick: the grid is 6 x 5 x 4), with every voxel below 24 set to 0 so that some
bricks hold only zeros and the chebyshev thumbnail has distances above 1, in the
layout hyd_renderer_create_package reads (package_meta.json, level-0 voxels,
chebyshev and mean thumbnails; include/hydrant_renderer.h).

    python tests/golden/make_package.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from hydrant_amd import dataset as D  # noqa: E402

NAME = "package_blobs_u8_44x40x30_bs8"


def volume() -> np.ndarray:
    vol = D.gen_blobs((44, 40, 30), n_blobs=3, seed=77, sig=(3.0, 6.0), amp=(0.6, 1.0))
    vol[vol < 24] = 0
    return vol


def main():
    meta = D.write_package(HERE / NAME, volume(), block_size=8, padding=1)
    print(NAME, meta["sample_levels"][0]["dim"])


if __name__ == "__main__":
    main()
