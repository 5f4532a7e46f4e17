"""Shared helpers: build oracle scenes from workloads and run both sides."""
from __future__ import annotations

import numpy as np

from hydrant_amd import hip
from hydrant_amd.workloads import Scene
from oracle import oracle as O


def oracle_volume(sc: Scene, schedule=None, vaddr=None) -> O.VolumeScene:
    b = sc.bricked
    return O.VolumeScene(sc.W, sc.H, sc.bbox_min, sc.bbox_max, sc.geom.step, sc.geom.du, sc.tf, b.cheb,
                         b.vaddr if vaddr is None else vaddr, b.page_k, b.page_b, b.pool_values(), b.pad,
                         max_steps=sc.max_steps, mode=sc.mode, density=sc.density,
                         schedule=sc.schedule if schedule is None else schedule, ert=sc.ert,
                         lowest_blkcnt=b.lowest_blkcnt)


def oracle_view(sc: Scene) -> O.OView:
    """The product's binary32 view handed to the oracle (the camera parity test pins that both agree)."""
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
