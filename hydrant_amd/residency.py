"""View-driven brick residency through the C ABI (include/hydrant_residency.h).

``cull_bricks`` restates the reference's frustum culler (OctreeCuller::cull,
include/hydrant/octree_culler.hpp:121-161); ``Residency`` its realtime paging
server (RtBlockPagingServerImpl, src/slave/paging/rt_block_paging.cc:249-326)
with the caller in the role of the decode thread: ``update`` returns the
missing bricks nearest first, ``commit`` the pool slot each one goes to.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _grid(cheb: np.ndarray):
    """(x, y, z) dims and a contiguous int32 copy of a [z][y][x] chebyshev grid."""
    cheb = np.ascontiguousarray(cheb, np.int32)
    if cheb.ndim != 3:
        raise ValueError("chebyshev grid must be 3-D [z][y][x]")
    return L.i3(cheb.shape[2], cheb.shape[1], cheb.shape[0]), cheb


def cull_bricks(view: L.View, cheb: np.ndarray, limit: int | None = None, bbox=None) -> np.ndarray:
    """Linear indices (x fastest, increasing) of the non-empty bricks the view's frustum
    contains, the `limit` nearest ones.  bbox: ((x0, y0, z0), (x1, y1, z1)) half-open."""
    dims, cheb = _grid(cheb)
    lo, hi = (bbox if bbox is not None else ((0, 0, 0), tuple(dims)))
    out = np.empty(cheb.size, np.int32)
    n = C.c_int64(0)
    L.call("hyd_cull_bricks", C.byref(view), dims, L.ptr(cheb, C.c_int32), L.i3(*lo), L.i3(*hi),
           int(limit if limit is not None else cheb.size), L.ptr(out, C.c_int32), C.byref(n))
    return out[:n.value].copy()


class Residency:
    """Pool-slot bookkeeping for a brick grid: slots first_slot .. first_slot + max_pages - 1."""

    def __init__(self, grid_shape_zyx, max_pages: int, first_slot: int = 0, basic_vaddr: np.ndarray | None = None):
        nz, ny, nx = (int(v) for v in grid_shape_zyx)
        self.shape = (nz, ny, nx)
        self.dims = L.i3(nx, ny, nz)
        self.max_pages = int(max_pages)
        basic = None
        if basic_vaddr is not None:
            basic = np.ascontiguousarray(basic_vaddr, np.int32)
            if basic.shape != self.shape:
                raise ValueError(f"basic_vaddr must have shape {self.shape}")
        h = C.c_void_p()
        L.call("hyd_residency_create", self.dims, int(first_slot), self.max_pages,
               L.ptr(basic, C.c_int32) if basic is not None else None, C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            L.call("hyd_residency_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, view: L.View, cheb: np.ndarray) -> np.ndarray:
        """Missing bricks for this view, nearest first."""
        _, cheb = _grid(cheb)
        if cheb.shape != self.shape:
            raise ValueError(f"chebyshev grid must have shape {self.shape}")
        out = np.empty(max(self.max_pages, 1), np.int32)
        n = C.c_int64(0)
        L.call("hyd_residency_update", self.h, C.byref(view), L.ptr(cheb, C.c_int32), L.ptr(out, C.c_int32),
               C.byref(n))
        return out[:n.value].copy()

    def commit(self, idx: int) -> int:
        """Pool slot for brick idx, or -1 (already present, or no slot can be freed)."""
        s = C.c_int32(0)
        L.call("hyd_residency_commit", self.h, int(idx), C.byref(s))
        return s.value

    def vaddr(self) -> np.ndarray:
        out = np.empty(self.shape, np.int32)
        L.call("hyd_residency_vaddr", self.h, L.ptr(out, C.c_int32))
        return out

    def stats(self):
        n = C.c_int64(0)
        s = C.c_int32(0)
        L.call("hyd_residency_stats", self.h, C.byref(n), C.byref(s))
        return n.value, s.value
