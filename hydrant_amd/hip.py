"""Python handle on a libhydrant_hip context.

torch is used only for device memory and streams (outputs are torch tensors
on ``cuda:<device>``); every kernel runs through the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .dataset import Bricked


def _f3(v):
    return L.f3(*[float(x) for x in v])


def orbit_camera(arm, center=(0.0, 0.0, 0.0), fovy: float = 60.0) -> L.Camera:
    """Camera from CameraOrbit (scene.hpp:68-81) computed by the library."""
    cam = L.Camera()
    arm_a = np.asarray(arm, np.float32)
    ctr = np.asarray(center, np.float32)
    L.call("hyd_orbit_camera", L.ptr(arm_a, C.c_float), L.ptr(ctr, C.c_float), float(fovy), C.byref(cam))
    return cam


def ptu_camera(position, target=(0, 0, 0), up=(0, 1, 0), fovy: float = 60.0) -> L.Camera:
    cam = L.Camera()
    cam.position = _f3(position)
    cam.target = _f3(target)
    cam.up = _f3(up)
    cam.fovy = float(fovy)
    return cam


def make_view(cam: L.Camera, ex_center, ex_size) -> L.View:
    v = L.View()
    c = np.asarray(ex_center, np.float32)
    s = np.asarray(ex_size, np.float32)
    L.call("hyd_make_view", C.byref(cam), L.ptr(c, C.c_float), L.ptr(s, C.c_float), C.byref(v))
    return v


def view_to_dict(v: L.View) -> dict:
    return {"trans": np.array(v.trans[:], np.float32), "ctg_fovy_2": np.float32(v.ctg_fovy_2),
            "ray_o": np.array(v.ray_o[:], np.float32)}


def volume_params(bbox_min, bbox_max, step, du, max_steps=4_000_000, mode=L.MODE_DEFAULT, rank=0.0, density=1.0,
                  schedule=L.SCHED_REFERENCE, ert=0.0) -> L.VolumeParams:
    p = L.VolumeParams()
    p.bbox_min = _f3(bbox_min)
    p.bbox_max = _f3(bbox_max)
    p.step = float(step)
    p.du = float(du)
    p.max_steps = int(max_steps)
    p.mode = int(mode)
    p.rank = float(rank)
    p.density = float(density)
    p.schedule = int(schedule)
    p.ert = float(ert)
    return p


class Context:
    """One libhydrant_hip context on ``cuda:<device>``; work runs on torch's current stream."""

    def __init__(self, device: int = 0):
        import torch

        self.torch = torch
        self.device = device
        self.dev = torch.device("cuda", device)
        h = C.c_void_p()
        L.call("hyd_create", device, C.byref(h))
        self.h = h
        self.sync_stream()

    def sync_stream(self):
        """Issue library work on torch's current stream of this device."""
        s = self.torch.cuda.current_stream(self.dev).cuda_stream
        L.call("hyd_set_stream", self.h, C.c_void_p(s))

    def close(self):
        if self.h:
            L.call("hyd_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- state ----
    def upload_pool(self, fmt: int, pad: int, pool: np.ndarray):
        pool = np.ascontiguousarray(pool)
        L.call("hyd_upload_pool", self.h, fmt, pad, pool.shape[0], pool.ctypes.data)

    def write_pool_pages(self, first_page: int, pages: np.ndarray):
        """Overwrite pool pages [first_page, first_page + len(pages)) with host bricks."""
        pages = np.ascontiguousarray(pages)
        L.call("hyd_write_pool_pages", self.h, int(first_page), pages.shape[0], pages.ctypes.data)

    def bind_pool_device(self, fmt: int, pad: int, n_pages: int, ptr: int):
        L.call("hyd_bind_pool_device", self.h, fmt, pad, n_pages, C.c_void_p(ptr))

    def set_page_table(self, vaddr: np.ndarray, page_k: np.ndarray, page_b: np.ndarray, lowest_blkcnt: int = 0):
        vaddr = np.ascontiguousarray(vaddr, np.int32)
        dims = np.array(vaddr.shape[::-1], np.int32)
        page_k = np.ascontiguousarray(page_k, np.float32)
        page_b = np.ascontiguousarray(page_b, np.float32)
        L.call("hyd_set_page_table", self.h, L.ptr(dims, C.c_int32), L.ptr(vaddr, C.c_int32), page_k.shape[0],
               L.ptr(page_k, C.c_float), L.ptr(page_b, C.c_float), lowest_blkcnt)

    def set_empty_grid(self, cheb: np.ndarray):
        cheb = np.ascontiguousarray(cheb, np.int32)
        dims = np.array(cheb.shape[::-1], np.int32)
        L.call("hyd_set_empty_grid", self.h, L.ptr(dims, C.c_int32), L.ptr(cheb, C.c_int32))

    def set_mean_grid(self, mean: np.ndarray):
        mean = np.ascontiguousarray(mean, np.float32)
        dims = np.array(mean.shape[::-1], np.int32)
        L.call("hyd_set_mean_grid", self.h, L.ptr(dims, C.c_int32), L.ptr(mean, C.c_float))

    def set_tf(self, tf: np.ndarray):
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        L.call("hyd_set_tf", self.h, tf.shape[0], L.ptr(tf, C.c_float))

    def load(self, b: Bricked, tf: np.ndarray | None = None):
        """Make a bricked dataset resident: pool, page table, empty grid, mean grid (and TF)."""
        self.upload_pool(b.fmt, b.pad, b.pool)
        self.set_page_table(b.vaddr, b.page_k, b.page_b, b.lowest_blkcnt)
        self.set_empty_grid(b.cheb)
        self.set_mean_grid(b.mean)
        if tf is not None:
            self.set_tf(tf)

    # --------------------------------------------------------- passes ----
    def render_volume(self, params: L.VolumeParams, view: L.View, W: int, H: int, film: bool = True,
                      count: bool = False, rgba=None):
        t = self.torch
        self.sync_stream()
        film_t = t.empty((H, W, 16), dtype=t.float32, device=self.dev) if film else None
        if rgba is not None:
            check_output(rgba, H * W * 4, "volume output")
        rgba_t = rgba if rgba is not None else t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        cnt = t.zeros(1, dtype=t.int64, device=self.dev) if count else None
        L.call("hyd_render_volume", self.h, C.byref(params), C.byref(view), W, H,
               C.c_void_p(film_t.data_ptr() if film else 0), C.c_void_p(rgba_t.data_ptr()),
               C.c_void_p(cnt.data_ptr() if count else 0))
        return {"film": film_t, "rgba": rgba_t, "samples": cnt}

    def resume_volume(self, params: L.VolumeParams, film, count: bool = False):
        t = self.torch
        H, W = check_image(film, 16, "film", self.dev)
        self.sync_stream()
        rgba_t = t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        cnt = t.zeros(1, dtype=t.int64, device=self.dev) if count else None
        L.call("hyd_resume_volume", self.h, C.byref(params), W, H, C.c_void_p(film.data_ptr()),
               C.c_void_p(rgba_t.data_ptr()), C.c_void_p(cnt.data_ptr() if count else 0))
        return {"film": film, "rgba": rgba_t, "samples": cnt}

    def pixel_pass(self, rgba, clear=(0, 0, 0)):
        t = self.torch
        H, W = check_image(rgba, 4, "rgba", self.dev)
        self.sync_stream()
        out = t.empty((H, W, 3), dtype=t.uint8, device=self.dev)
        cc = (C.c_uint8 * 3)(*clear)
        L.call("hyd_pixel_pass", self.h, C.c_void_p(rgba.data_ptr()), W, H, cc, C.c_void_p(out.data_ptr()))
        return out

    def fetch_pass(self, film):
        t = self.torch
        H, W = check_image(film, 16, "film", self.dev)
        self.sync_stream()
        out = t.empty((H, W, 8), dtype=t.float32, device=self.dev)
        L.call("hyd_fetch_pass", self.h, C.c_void_p(film.data_ptr()), W, H, C.c_void_p(out.data_ptr()))
        return out

    def composite(self, partials, z_order):
        """partials: (N, H, W, 8) fetch pixels; returns (rgba, rgb8)."""
        t = self.torch
        if partials.dim() != 4:
            raise ValueError(f"partials must have shape (N, H, W, 8); got {tuple(partials.shape)}")
        n = partials.shape[0]
        H, W = check_image(partials, 8, "partials", self.dev, lead=n)
        self.sync_stream()
        z = np.ascontiguousarray(z_order, np.int32)
        rgba = t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        rgb = t.empty((H, W, 3), dtype=t.uint8, device=self.dev)
        L.call("hyd_composite", self.h, n, C.c_void_p(partials.data_ptr()), L.ptr(z, C.c_int32), W, H,
               C.c_void_p(rgba.data_ptr()), C.c_void_p(rgb.data_ptr()))
        return rgba, rgb

    def iso_fetch_pass(self, px):
        """IsosurfaceShaderKernel::fetch of an (H, W, 6) isosurface output: (H, W, 8) uint8
        records (3 colour bytes, 1 pad byte, float32 depth bytes)."""
        t = self.torch
        H, W = check_image(px, 6, "isosurface pixels", self.dev)
        self.sync_stream()
        out = t.empty((H, W, 8), dtype=t.uint8, device=self.dev)
        L.call("hyd_iso_fetch_pass", self.h, C.c_void_p(px.data_ptr()), W, H, C.c_void_p(out.data_ptr()))
        return out

    def composite_iso(self, partials, z_order):
        """partials: (N, H, W, 8) uint8 isosurface fetch records; returns (rgb8, depth)."""
        t = self.torch
        if partials.dim() != 4 or partials.shape[-1] != 8 or partials.dtype != t.uint8 or \
                not partials.is_contiguous() or partials.device != self.dev:
            raise ValueError(f"partials must be a contiguous uint8 (N, H, W, 8) tensor on {self.dev}; "
                             f"got {partials.dtype} {tuple(partials.shape)} on {partials.device}")
        n, H, W = partials.shape[0], partials.shape[1], partials.shape[2]
        self.sync_stream()
        z = np.ascontiguousarray(z_order, np.int32)
        rgb = t.empty((H, W, 3), dtype=t.uint8, device=self.dev)
        depth = t.empty((H, W), dtype=t.float32, device=self.dev)
        L.call("hyd_composite_iso", self.h, n, C.c_void_p(partials.data_ptr()), L.ptr(z, C.c_int32), W, H,
               C.c_void_p(rgb.data_ptr()), C.c_void_p(depth.data_ptr()))
        return rgb, depth

    def render_sort_last(self, params: L.VolumeParams, view: L.View, W: int, H: int, batches, grid_dims,
                         count: bool = False):
        """Sort-last render over residency batches: batches is a list of (fmt, pad, pool, vaddr, page_k,
        page_b); each batch is made resident in turn and its bricks' per-ray partials are written,
        then the partials are composited on the GPU.  Returns (rgba, samples)."""
        t = self.torch
        self.sync_stream()
        max_slots = int(sum(grid_dims))
        slots = t.empty((max_slots, H, W, 4), dtype=t.float32, device=self.dev)
        nslots = t.zeros((H, W), dtype=t.int32, device=self.dev)
        cnt = t.zeros(1, dtype=t.int64, device=self.dev) if count else None
        L.call("hyd_segments_begin", self.h, W, H, max_slots, C.c_void_p(slots.data_ptr()))
        for fmt, pad, pool, vaddr, page_k, page_b in batches:
            self.upload_pool(fmt, pad, pool)
            self.set_page_table(vaddr, page_k, page_b)
            self.sync_stream()
            L.call("hyd_segments_pass", self.h, C.byref(params), C.byref(view), W, H, max_slots,
                   C.c_void_p(slots.data_ptr()), C.c_void_p(nslots.data_ptr()),
                   C.c_void_p(cnt.data_ptr() if count else 0))
        rgba = t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        L.call("hyd_segments_composite", self.h, W, H, max_slots, C.c_void_p(slots.data_ptr()),
               C.c_void_p(nslots.data_ptr()), C.c_void_p(rgba.data_ptr()))
        return rgba, cnt

    def render_sort_last_resident(self, params: L.VolumeParams, view: L.View, W: int, H: int, batches, grid_dims,
                                  slots=None, nslots=None, rgba=None, count: bool = False):
        """render_sort_last with every batch's pool already in device memory: batches is a
        list of (fmt, pad, pool_tensor, vaddr, page_k, page_b) with pool_tensor a device
        tensor of the batch's bricks.  Per batch the pool is bound (no copy) and its page
        table set; slots / nslots / rgba may be passed in to reuse their memory."""
        t = self.torch
        self.sync_stream()
        max_slots = int(sum(grid_dims))
        if slots is None:
            slots = t.empty((max_slots, H, W, 4), dtype=t.float32, device=self.dev)
        check_output(slots, max_slots * H * W * 4, "segment slots")
        if nslots is None:
            nslots = t.zeros((H, W), dtype=t.int32, device=self.dev)
        if rgba is None:
            rgba = t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        check_output(rgba, H * W * 4, "sort-last output")
        if nslots.dtype != t.int32 or not nslots.is_contiguous() or nslots.numel() < H * W:
            raise ValueError("nslots must be a contiguous int32 tensor of H * W entries")
        cnt = t.zeros(1, dtype=t.int64, device=self.dev) if count else None
        L.call("hyd_segments_begin", self.h, W, H, max_slots, C.c_void_p(slots.data_ptr()))
        for fmt, pad, pool_t, vaddr, page_k, page_b in batches:
            self.bind_pool_device(fmt, pad, int(pool_t.shape[0]), pool_t.data_ptr())
            self.set_page_table(vaddr, page_k, page_b)
            L.call("hyd_segments_pass", self.h, C.byref(params), C.byref(view), W, H, max_slots,
                   C.c_void_p(slots.data_ptr()), C.c_void_p(nslots.data_ptr()),
                   C.c_void_p(cnt.data_ptr() if count else 0))
        L.call("hyd_segments_composite", self.h, W, H, max_slots, C.c_void_p(slots.data_ptr()),
               C.c_void_p(nslots.data_ptr()), C.c_void_p(rgba.data_ptr()))
        return rgba, cnt

    def segment_set(self, batches) -> "SegmentSet":
        """A set of residency batches rendered by one segment launch: batches is a list of
        (fmt, pad, pool_tensor, vaddr, page_k, page_b) with pool_tensor a device tensor
        (render_sort_last_resident's form); uses the context's current empty-space grid."""
        return SegmentSet(self, batches)

    def render_sort_last_set(self, params: L.VolumeParams, view: L.View, W: int, H: int, sset: "SegmentSet",
                             slots=None, nslots=None, rgba=None, count: bool = False):
        """render_sort_last_resident with the set's batches rendered in one launch."""
        t = self.torch
        self.sync_stream()
        max_slots = int(sum(sset.grid_dims))
        if slots is None:
            slots = t.empty((max_slots, H, W, 4), dtype=t.float32, device=self.dev)
        check_output(slots, max_slots * H * W * 4, "segment slots")
        if nslots is None:
            nslots = t.zeros((H, W), dtype=t.int32, device=self.dev)
        if rgba is None:
            rgba = t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        check_output(rgba, H * W * 4, "sort-last output")
        if nslots.dtype != t.int32 or not nslots.is_contiguous() or nslots.numel() < H * W:
            raise ValueError("nslots must be a contiguous int32 tensor of H * W entries")
        cnt = t.zeros(1, dtype=t.int64, device=self.dev) if count else None
        L.call("hyd_segments_begin", self.h, W, H, max_slots, C.c_void_p(slots.data_ptr()))
        L.call("hyd_segments_pass_set", self.h, sset.h, C.byref(params), C.byref(view), W, H, max_slots,
               C.c_void_p(slots.data_ptr()), C.c_void_p(nslots.data_ptr()), C.c_void_p(cnt.data_ptr() if count else 0))
        L.call("hyd_segments_composite", self.h, W, H, max_slots, C.c_void_p(slots.data_ptr()),
               C.c_void_p(nslots.data_ptr()), C.c_void_p(rgba.data_ptr()))
        return rgba, cnt

    def trace_volume(self, params: L.VolumeParams, view: L.View, W: int, H: int, x: int, y: int,
                     max_recs: int = 8192) -> np.ndarray:
        recs = (L.TraceRec * max_recs)()
        n = C.c_int32(0)
        L.call("hyd_trace_volume", self.h, C.byref(params), C.byref(view), W, H, x, y, recs, max_recs, C.byref(n))
        arr = np.frombuffer(recs, dtype=np.int32).reshape(max_recs, 10)[: n.value].copy()
        return arr

    def render_blocks(self, params: L.BlocksParams, view: L.View, W: int, H: int):
        t = self.torch
        self.sync_stream()
        rgba = t.empty((H, W, 4), dtype=t.float32, device=self.dev)
        L.call("hyd_render_blocks", self.h, C.byref(params), C.byref(view), W, H, C.c_void_p(rgba.data_ptr()))
        return rgba

    def render_iso(self, params: L.IsoParams, view: L.View, W: int, H: int, count: bool = False, px=None):
        t = self.torch
        self.sync_stream()
        if px is not None:
            check_output(px, H * W * 6, "isosurface output")
        out = px if px is not None else t.empty((H, W, 6), dtype=t.float32, device=self.dev)
        cnt = t.zeros(1, dtype=t.int64, device=self.dev) if count else None
        L.call("hyd_render_iso", self.h, C.byref(params), C.byref(view), W, H, C.c_void_p(out.data_ptr()),
               C.c_void_p(cnt.data_ptr() if count else 0))
        return {"px": out, "samples": cnt}

    def synchronize(self):
        L.call("hyd_synchronize", self.h)

    # -------------------------------------------------------- batched ----
    def set_volumes(self, descs: list[dict]):
        """descs: dicts with grid (x, y, z), bs, apron, bbox_max, step, vaddr, cheb (numpy grids)."""
        arr = (L.VolumeDesc * len(descs))()
        keep = []
        for i, d in enumerate(descs):
            va = np.ascontiguousarray(d["vaddr"], np.int32)
            ch = np.ascontiguousarray(d["cheb"], np.int32)
            keep += [va, ch]
            arr[i].grid = L.i3(*[int(x) for x in d["grid"]])
            arr[i].bs, arr[i].apron = int(d["bs"]), int(d["apron"])
            arr[i].bbox_max = _f3(d["bbox_max"])
            arr[i].step = float(d["step"])
            arr[i].value_scale = float(d.get("value_scale", 1.0))
            arr[i].vaddr = L.ptr(va, C.c_int32)
            arr[i].chebyshev = L.ptr(ch, C.c_int32)
        L.call("hyd_set_volumes", self.h, len(descs), arr)

    def batch(self, views: list[tuple]) -> "Batch":
        """views: (hyd View, width, height, volume index) tuples, packed back to back in the output."""
        return Batch(self, views)


def check_output(out, n_floats: int, what: str):
    """Kernels write n_floats floats straight into a caller's `out`: refuse a buffer that
    cannot hold them."""
    import torch

    if out.dtype != torch.float32 or not out.is_contiguous() or out.numel() < n_floats:
        raise ValueError(f"{what} must be a contiguous float32 tensor of at least {n_floats} floats; "
                         f"got {out.dtype} {tuple(out.shape)}")
    if not out.is_cuda:
        raise ValueError(f"{what} must be a device tensor")


def srgb8_thresholds() -> np.ndarray:
    """The 256 sRGB byte thresholds hyd_composite uses (host computation)."""
    t = np.zeros(256, np.float32)
    L.call("hyd_srgb8_thresholds", L.ptr(t, C.c_float))
    return t


def check_image(x, per_px: int, what: str, dev, lead: int = 1):
    """Kernels read `x` as `lead` images of H x W pixels of `per_px` floats: refuse a tensor
    of another type, device, layout or size before it reaches one.  Returns (H, W)."""
    import torch

    shape = tuple(x.shape)
    if x.dim() < 3 or shape[-1] != per_px:
        raise ValueError(f"{what} must have shape (..., H, W, {per_px}); got {shape}")
    H, W = int(shape[-3]), int(shape[-2])
    if x.dtype != torch.float32 or not x.is_contiguous():
        raise ValueError(f"{what} must be a contiguous float32 tensor; got {x.dtype} {shape}")
    if not x.is_cuda or x.device != torch.device(dev):
        raise ValueError(f"{what} must live on device {dev}; got {x.device}")
    if H <= 0 or W <= 0 or x.numel() != lead * H * W * per_px:
        raise ValueError(f"{what} must hold {lead} x {H} x {W} x {per_px} floats; got {shape}")
    return H, W



class SegmentSet:
    """hyd_segment_set: residency batches (device pools and their page tables) that one
    segment launch renders together.  Keeps the pools referenced while it lives."""

    def __init__(self, ctx, batches):
        self._keep = []
        arr = (L.SegmentBatch * len(batches))()
        dims = None
        for i, (fmt, pad, pool_t, vaddr, page_k, page_b) in enumerate(batches):
            vaddr = np.ascontiguousarray(vaddr, np.int32)
            page_k = np.ascontiguousarray(page_k, np.float32)
            page_b = np.ascontiguousarray(page_b, np.float32)
            dims = vaddr.shape[::-1] if dims is None else dims
            if vaddr.shape[::-1] != dims:
                raise ValueError("every batch's page table must have the same grid")
            self._keep += [pool_t, vaddr, page_k, page_b]
            arr[i] = L.SegmentBatch(int(fmt), int(pad), int(page_k.shape[0]), C.c_void_p(pool_t.data_ptr()),
                                    L.ptr(vaddr, C.c_int32), L.ptr(page_k, C.c_float), L.ptr(page_b, C.c_float))
        self.grid_dims = tuple(int(d) for d in dims)
        h = C.c_void_p()
        L.call("hyd_segment_set_create", ctx.h, len(batches), arr, C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.call("hyd_segment_set_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

class Batch:
    """A batched-render plan: many views, one persistent launch (optionally a HIP graph)."""

    def __init__(self, ctx: Context, views: list[tuple]):
        self.ctx = ctx
        arr = (L.BatchView * len(views))()
        off = 0
        self.offsets = []
        for i, (v, w, h, vol) in enumerate(views):
            arr[i].view = v
            arr[i].width, arr[i].height, arr[i].volume = int(w), int(h), int(vol)
            arr[i].out_offset = off
            self.offsets.append((off, int(w), int(h)))
            off += int(w) * int(h)
        self.n_pixels = off
        h = C.c_void_p()
        L.call("hyd_batch_create", ctx.h, arr, len(views), C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            L.call("hyd_batch_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def new_output(self):
        t = self.ctx.torch
        return t.empty((self.n_pixels, 4), dtype=t.float32, device=self.ctx.dev)

    def render(self, cfg: L.BatchConfig, density: float, ert: float, out=None, count: bool = False):
        t = self.ctx.torch
        self.ctx.sync_stream()
        out = self.new_output() if out is None else out
        check_output(out, 4 * self.n_pixels, "batched output")
        cnt = t.zeros(1, dtype=t.int64, device=self.ctx.dev) if count else None
        L.call("hyd_batch_render", self.ctx.h, self.h, C.byref(cfg), float(density), float(ert),
               C.c_void_p(out.data_ptr()), C.c_void_p(cnt.data_ptr() if count else 0))
        return out, cnt

    def capture(self, cfg: L.BatchConfig, density: float, ert: float, out):
        check_output(out, 4 * self.n_pixels, "batched output")
        self.ctx.sync_stream()
        L.call("hyd_batch_capture", self.ctx.h, self.h, C.byref(cfg), float(density), float(ert),
               C.c_void_p(out.data_ptr()))

    def replay(self):
        self.ctx.sync_stream()
        L.call("hyd_batch_replay", self.ctx.h, self.h)

    def autotune(self, density: float, ert: float, out, reps: int = 2):
        check_output(out, 4 * self.n_pixels, "batched output")
        self.ctx.sync_stream()
        best = L.BatchConfig()
        ms = C.c_float(0)
        L.call("hyd_batch_autotune", self.ctx.h, self.h, float(density), float(ert), C.c_void_p(out.data_ptr()),
               int(reps), C.byref(best), C.byref(ms))
        return best, ms.value

    def autotune_schedule(self, density: float, max_error: float, out, reps: int = 2):
        """Launch shape and early-termination threshold: the fastest within max_error."""
        check_output(out, 4 * self.n_pixels, "batched output")
        self.ctx.sync_stream()
        best = L.BatchConfig()
        ert = C.c_float(0)
        ms = C.c_float(0)
        L.call("hyd_batch_autotune_schedule", self.ctx.h, self.h, float(density), float(max_error),
               C.c_void_p(out.data_ptr()), int(reps), C.byref(best), C.byref(ert), C.byref(ms))
        return best, ert.value, ms.value

    def view_image(self, out, i: int):
        off, w, h = self.offsets[i]
        return out[off:off + w * h].view(h, w, 4)


def batch_config(tile_w=8, unroll=1, waves_per_cu=8, group=1) -> L.BatchConfig:
    c = L.BatchConfig()
    c.tile_w, c.unroll, c.waves_per_cu, c.group = tile_w, unroll, waves_per_cu, group
    return c
