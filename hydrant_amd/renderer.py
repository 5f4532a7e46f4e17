"""The JSON-configured renderer interface (include/hydrant_renderer.h) from Python.

``Renderer(ctx, volume, config)`` is ``RendererFactory(dataset).create(config)``
(include/hydrant/core/renderer.hpp:53-69 of the reference): the C++ front end
bricks the volume, derives the empty-space grids, makes bricks resident and
runs the HIP passes; ``offline_render(camera)`` returns the RGB8 image as the
reference's ``IRenderer::offline_render`` does.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib as L

_FMT_OF_DTYPE = {np.dtype(np.uint8): L.FMT_U8, np.dtype(np.float16): L.FMT_F16, np.dtype(np.float32): L.FMT_F32}


def _text(cfg) -> bytes:
    return (cfg if isinstance(cfg, str) else json.dumps(cfg)).encode()


def candidates() -> list[str]:
    """RendererFactory::list_candidates: the registered renderer names."""
    n = C.c_int32(0)
    L.call("hyd_renderer_candidates", None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    L.call("hyd_renderer_candidates", buf, n.value + 1, C.byref(n))
    return [s for s in buf.value.decode().split(",") if s]


def parse_config(cfg) -> tuple[str, int, int]:
    """(renderer name, width, height) of a RendererConfig or whole config (no GPU)."""
    name = C.create_string_buffer(256)
    w, h = C.c_int32(0), C.c_int32(0)
    L.call("hyd_renderer_parse_config", _text(cfg), name, 256, C.byref(w), C.byref(h))
    return name.value.decode(), w.value, h.value


def parse_camera(cfg) -> L.Camera:
    """The hyd_camera of a CameraConfig or whole config (no GPU)."""
    cam = L.Camera()
    L.call("hyd_renderer_parse_camera", _text(cfg), C.byref(cam))
    return cam


class Renderer:
    """One renderer over a raw volume (numpy (z, y, x) of uint8 / float16 / float32),
    or over a dataset package on disk (``Renderer.from_package``)."""

    @classmethod
    def from_package(cls, ctx, dataset_path, config) -> "Renderer":
        """RendererFactory(dataset_path).create(config): package_meta.json, its level-0
        voxels and thumbnails (hydrant_amd.dataset.write_package writes one)."""
        self = cls.__new__(cls)
        self._vol = None
        h = C.c_void_p()
        L.call("hyd_renderer_create_package", ctx.h, str(dataset_path).encode(), _text(config), C.byref(h))
        self.h = h
        self._resolution()
        return self

    def _resolution(self) -> None:
        w, hh = C.c_int32(0), C.c_int32(0)
        L.call("hyd_renderer_resolution", self.h, C.byref(w), C.byref(hh))
        self.W, self.H = w.value, hh.value

    def __init__(self, ctx, volume: np.ndarray, config, block_size: int = 64, apron: int = 1):
        vol = np.ascontiguousarray(volume)
        if vol.dtype not in _FMT_OF_DTYPE or vol.ndim != 3:
            raise TypeError("volume must be a 3-D uint8, float16 or float32 array")
        self._vol = vol
        d = L.Dataset()
        d.dims[:] = [int(x) for x in vol.shape[::-1]]
        d.fmt = _FMT_OF_DTYPE[vol.dtype]
        d.voxels = vol.ctypes.data
        d.block_size, d.apron = int(block_size), int(apron)
        h = C.c_void_p()
        L.call("hyd_renderer_create", ctx.h, C.byref(d), _text(config), C.byref(h))
        self.h = h
        self._resolution()

    def update(self, params) -> None:
        L.call("hyd_renderer_update", self.h, _text(params))

    def offline_render(self, camera) -> np.ndarray:
        out = np.empty((self.H, self.W, 3), np.uint8)
        L.call("hyd_renderer_offline_render", self.h, _text(camera), out.ctypes.data)
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            L.call("hyd_renderer_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
