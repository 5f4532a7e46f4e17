"""Shared helpers: build oracle scenes from workloads and run both sides."""
from __future__ import annotations

import numpy as np

from hydrant_amd import hip
from hydrant_amd.workloads import Scene
from oracle import oracle as O


def oracle_volume(sc: Scene, schedule=None, vaddr=None, pool_values=None, exact_mapping=False) -> O.VolumeScene:
    b = sc.bricked
    return O.VolumeScene(sc.W, sc.H, sc.bbox_min, sc.bbox_max, sc.geom.step, sc.geom.du, sc.tf, b.cheb,
                         b.vaddr if vaddr is None else vaddr, b.page_k, b.page_b,
                         b.pool_values() if pool_values is None else pool_values, b.pad,
                         max_steps=sc.max_steps, mode=sc.mode, rank=sc.rank, density=sc.density,
                         schedule=sc.schedule if schedule is None else schedule, ert=sc.ert,
                         lowest_blkcnt=b.lowest_blkcnt, map_bs=b.bs if exact_mapping else 0)


def exact_pool_values(sc: Scene) -> np.ndarray:
    """The pool's bricks re-packed from the source u8 volume: float(v)/255 (ground truth
    for reduced-precision storage)."""
    from hydrant_amd import _lib as L
    from hydrant_amd import dataset as D

    b = sc.bricked
    return D.normalised(D.pack(sc.extra["volume"], b.bs, b.apron, L.FMT_U8, b.order[:b.n_pages]))


def error_pixels(W: int, H: int) -> np.ndarray:
    """Pixels compared against the FP64 oracle: the full frame up to 512x512, else a
    centred 64x64 crop plus every 7th pixel (row-major)."""
    if W * H <= 512 * 512:
        ys, xs = np.mgrid[0:H, 0:W]
        return np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    cx, cy = W // 2 - 32, H // 2 - 32
    ys, xs = np.mgrid[cy:cy + 64, cx:cx + 64]
    crop = np.stack([xs.ravel(), ys.ravel()], 1)
    idx = np.arange(0, W * H, 7)
    every = np.stack([idx % W, idx // W], 1)
    return np.unique(np.concatenate([crop, every]), axis=0).astype(np.int32)


def oracle_view(sc: Scene) -> O.OView:
    """The product's binary32 view handed to the oracle (tests/test_camera_parity.py checks the
    library's camera against the oracle's, bit for bit)."""
    return O.view_from_dict(hip.view_to_dict(sc.view()))


def film_fields(film: np.ndarray) -> dict:
    """Split a (H, W, 16) float32 film into named fields (int fields reinterpreted)."""
    f = np.ascontiguousarray(film, np.float32)
    ints = f.view(np.int32)
    return {"o": f[..., 0:3], "d": f[..., 3:6], "nsteps": ints[..., 6], "step_size": ints[..., 7],
            "v": f[..., 8:12], "theta": f[..., 12:15], "phi": f[..., 15]}


def oracle_film_fields(film: np.ndarray) -> dict:
    return {"o": film[..., 0:3].astype(np.float32), "d": film[..., 3:6].astype(np.float32),
            "nsteps": film[..., 6].astype(np.int32), "step_size": film[..., 7].astype(np.int32),
            "v": film[..., 8:12].astype(np.float32), "theta": film[..., 12:15].astype(np.float32),
            "phi": film[..., 15].astype(np.float32)}


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) if np.size(a) else 0.0
