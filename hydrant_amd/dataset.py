"""Dataset preparation: synthetic volumes, padded bricks, thumbnails, page tables.

Stands in for the reference's varch tooling and paging server
(build_dataset.sh:79-124, src/slave/paging/lossless_block_paging.cc:15-83):
volumes are synthesised deterministically (include/hydrant_dataset.h), packed
into padded bricks in page order, and described by the same BlockPaging
structures the reference builds (vaddr grid, per-page {k, b} mapping,
chebyshev empty-space thumbnail, mean thumbnail).
"""
from __future__ import annotations

import ctypes as C
import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


# ----------------------------------------------------------- generators ----
def gen_blobs(n: int | tuple, n_blobs: int = 16, seed: int = 1, sig=(8.0, 40.0), sigma_scale: float = 1.0,
              amp=(0.3, 1.0)) -> np.ndarray:
    """Sum of seeded Gaussian blobs quantised to u8, shape (nz, ny, nx)."""
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    out = np.empty((nz, ny, nx), np.uint8)
    L.call("hyd_gen_blobs_u8", nx, ny, nz, n_blobs, seed, sig[0], sig[1], sigma_scale, amp[0], amp[1],
           out.ctypes.data)
    return out


def gen_perlin(n: int, octaves: int = 5, persistence: float = 0.5, base_period: float = 128.0, gain: float = 2.0,
               seed: int = 2) -> np.ndarray:
    """Seeded improved-Perlin fBm volume quantised to u8, shape (n, n, n)."""
    out = np.empty((n, n, n), np.uint8)
    L.call("hyd_gen_perlin_u8", n, octaves, persistence, base_period, gain, seed, out.ctypes.data)
    return out


def gen_sdf_spheres(n: int, n_spheres: int = 64, r=(10.0, 60.0), k: float = 8.0, seed: int = 3) -> np.ndarray:
    """Smooth-union SDF of seeded spheres as IEEE half bits (uint16), shape (n, n, n)."""
    out = np.empty((n, n, n), np.uint16)
    L.call("hyd_gen_sdf_spheres_f16", n, n_spheres, r[0], r[1], k, seed, out.ctypes.data)
    return out


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def normalised(vol: np.ndarray) -> np.ndarray:
    """Voxel values as the renderer reads them: float(v)/255 for u8, half->float for f16."""
    if vol.dtype == np.uint8:
        return vol.astype(np.float32) / np.float32(255.0)
    if vol.dtype == np.uint16:
        return vol.view(np.float16).astype(np.float32)
    return vol.astype(np.float32)


def _fmt_of(vol: np.ndarray) -> int:
    return {np.dtype(np.uint8): L.FMT_U8, np.dtype(np.uint16): L.FMT_F16, np.dtype(np.float32): L.FMT_F32}[vol.dtype]


_NP_OF_FMT = {L.FMT_U8: np.uint8, L.FMT_F16: np.uint16, L.FMT_F32: np.float32}


# ------------------------------------------------------ bricked dataset ----
@dataclass
class Bricked:
    """A bricked volume in the reference's BlockPaging terms."""
    dims: tuple            # raw volume dims (x, y, z)
    bs: int                # block size (brick edge without apron)
    apron: int
    grid: tuple            # brick grid (x, y, z)
    fmt: int               # pool storage format
    pool: np.ndarray       # (n_pages, pad, pad, pad) in page order
    vaddr: np.ndarray      # int32 grid (z, y, x); -1 = not resident
    page_k: np.ndarray     # float32 (n_pages,)
    page_b: np.ndarray     # float32 (n_pages, 3)
    cheb: np.ndarray       # int32 grid
    mean: np.ndarray       # float32 grid (interior mean, normalised)
    empty: np.ndarray      # uint8 grid
    order: np.ndarray      # non-empty bricks in page-assignment order
    n_nonempty: int
    lowest_blkcnt: int = 0
    bmin: np.ndarray = field(default=None)
    bmax: np.ndarray = field(default=None)
    raw_counts: bool = False   # f16/f32 pool holds 8-bit counts (renderer scale 1/255)

    @property
    def pad(self) -> int:
        return self.bs + 2 * self.apron

    @property
    def n_pages(self) -> int:
        return int(self.pool.shape[0])

    @property
    def value_scale(self) -> float:
        """Multiplier of interpolated pool values in the batched renderer."""
        return 1.0 / 255.0 if (self.raw_counts or self.fmt == L.FMT_U8) else 1.0

    def pool_values(self) -> np.ndarray:
        """Normalised float32 voxel values of the pool (what the samplers read)."""
        if self.raw_counts:
            return self.pool.view(np.float16).astype(np.float32) / np.float32(255) if self.fmt == L.FMT_F16 \
                else self.pool.astype(np.float32) / np.float32(255)
        return normalised(self.pool)


def mapping(bs: int, apron: int) -> tuple[np.float32, np.float32]:
    """BlockSamplerMapping k, b (lossless_block_paging.cc:21-22): k = bs/pad, b = (apron/bs)*k, binary32."""
    pad = bs + 2 * apron
    k = np.float32(bs) / np.float32(pad)
    b = (np.float32(apron) / np.float32(bs)) * k
    return np.float32(k), np.float32(b)


def brick_stats(vol: np.ndarray, bs: int, apron: int):
    dims = np.array(vol.shape[::-1], np.int32)
    grid = [(int(d) + bs - 1) // bs for d in dims]
    nb = grid[0] * grid[1] * grid[2]
    bmin = np.empty(nb, np.float32)
    bmax = np.empty(nb, np.float32)
    bmean = np.empty(nb, np.float32)
    vol = np.ascontiguousarray(vol)
    L.call("hyd_brick_stats", vol.ctypes.data, _fmt_of(vol), L.ptr(dims, C.c_int32), bs, apron,
           L.ptr(bmin, C.c_float), L.ptr(bmax, C.c_float), L.ptr(bmean, C.c_float))
    shape = (grid[2], grid[1], grid[0])
    return grid, bmin.reshape(shape), bmax.reshape(shape), bmean.reshape(shape)


def tf_alpha_zero_on(tf: np.ndarray, lo: float, hi: float) -> bool:
    tf = np.ascontiguousarray(tf, np.float32).reshape(-1)
    return bool(L.lib().hyd_tf_alpha_zero_on(L.ptr(tf, C.c_float), tf.size // 4, float(lo), float(hi)))


def chebyshev(empty: np.ndarray) -> np.ndarray:
    empty = np.ascontiguousarray(empty, np.uint8)
    dims = np.array(empty.shape[::-1], np.int32)
    out = np.empty(empty.shape, np.int32)
    L.call("hyd_chebyshev", L.ptr(dims, C.c_int32), L.ptr(empty, C.c_uint8), L.ptr(out, C.c_int32))
    return out


# ------------------------------------------------------- dataset package ----
_PKG_FORMATS = {np.dtype(np.uint8): "u8", np.dtype(np.float16): "f16", np.dtype(np.uint16): "f16",
                np.dtype(np.float32): "f32"}


def write_package(root, vol: np.ndarray, block_size: int, padding: int = 1, empty: np.ndarray | None = None) -> dict:
    """Write a dataset package that hyd_renderer_create_package reads (the layout is in
    include/hydrant_renderer.h): package_meta.json, the level-0 voxels and the chebyshev
    and mean thumbnails of sample_levels[0] (the files build_dataset.sh:79-124 makes with
    varch's tools).  A brick is empty when its apron-inclusive maximum is <= 0, the same
    for every transfer function, unless ``empty`` (bool, (dim.z, dim.y, dim.x)) says
    otherwise; the mean is that of its normalised interior voxels.  Returns the meta
    dictionary."""
    import json
    from pathlib import Path

    vol = np.ascontiguousarray(vol)
    if vol.ndim != 3 or vol.dtype not in _PKG_FORMATS:
        raise TypeError("vol must be a 3-D uint8, float16 (or uint16 half bits) or float32 array")
    if vol.dtype == np.float16:
        vol = vol.view(np.uint16)
    root = Path(root)
    root.mkdir(parents=True, exist_ok=True)
    grid, _, bmax, bmean = brick_stats(vol, block_size, padding)
    if empty is None:
        empty = bmax <= 0
    elif np.shape(empty) != bmax.shape:
        raise ValueError(f"empty must have shape {bmax.shape}")
    cheb = chebyshev(np.asarray(empty, np.uint8))
    raw = vol.shape[::-1]
    meta = {
        "block_size": int(block_size),
        "padding": int(padding),
        "sample_levels": [{
            "raw": {"x": int(raw[0]), "y": int(raw[1]), "z": int(raw[2])},
            "dim": {"x": int(grid[0]), "y": int(grid[1]), "z": int(grid[2])},
            "thumbnails": {"chebyshev": "thumb_chebyshev.i32", "mean": "thumb_mean.f32"},
            "voxels": {"path": "voxels.raw", "format": _PKG_FORMATS[vol.dtype]},
        }],
    }
    (root / "voxels.raw").write_bytes(vol.astype(vol.dtype.newbyteorder("<"), copy=False).tobytes())
    (root / "thumb_chebyshev.i32").write_bytes(cheb.astype("<i4").tobytes())
    (root / "thumb_mean.f32").write_bytes(bmean.astype("<f4").tobytes())
    (root / "package_meta.json").write_text(json.dumps(meta, indent=1) + "\n")
    return meta


def read_package_info(root) -> L.PackageInfo:
    """hyd_package_read_info: the package's sizes, after validating meta and thumbnails."""
    info = L.PackageInfo()
    L.call("hyd_package_read_info", str(root).encode(), C.byref(info))
    return info


def read_package_thumbnails(root) -> tuple[np.ndarray, np.ndarray | None]:
    """(chebyshev int32, mean float32 or None), each shaped (dim.z, dim.y, dim.x)."""
    info = read_package_info(root)
    shape = (info.dim[2], info.dim[1], info.dim[0])
    cheb = np.empty(shape, np.int32)
    mean = np.empty(shape, np.float32) if info.has_mean else None
    L.call("hyd_package_read_thumbnails", str(root).encode(), L.ptr(cheb, C.c_int32),
           L.ptr(mean, C.c_float) if mean is not None else None, cheb.size)
    return cheb, mean


def pack(vol: np.ndarray, bs: int, apron: int, fmt_out: int, bricks: np.ndarray | None = None,
         normalize: bool = True) -> np.ndarray:
    vol = np.ascontiguousarray(vol)
    dims = np.array(vol.shape[::-1], np.int32)
    grid = [(int(d) + bs - 1) // bs for d in dims]
    count = grid[0] * grid[1] * grid[2] if bricks is None else len(bricks)
    pad = bs + 2 * apron
    out = np.empty((count, pad, pad, pad), _NP_OF_FMT[fmt_out])
    if count == 0:
        return out
    bp = None
    if bricks is not None:
        bricks = np.ascontiguousarray(bricks, np.int32)
        bp = L.ptr(bricks, C.c_int32)
    L.call("hyd_pack_bricks", vol.ctypes.data, _fmt_of(vol), L.ptr(dims, C.c_int32), bs, apron, fmt_out,
           int(bool(normalize)), bp, count, out.ctypes.data)
    return out


@dataclass
class LowestLevel:
    """Fallback pages of a brick grid: one page per level-0 brick holding the block of the
    coarse level that covers it, with the mapping into that block."""
    pages: np.ndarray      # (n_bricks0, pad, pad, pad), level-0 brick order (x fastest)
    page_k: np.ndarray     # float32 (n_bricks0,)
    page_b: np.ndarray     # float32 (n_bricks0, 3)
    scale: int             # level-0 bricks per coarse block along an axis (2^level)


def lowest_level(vol: np.ndarray, bs: int, apron: int, fmt_out: int, level: int,
                 normalize: bool = True) -> LowestLevel:
    """The reference's lowest-sample-level fallback (RtBlockPagingServerImpl::
    unarchive_lowest_level, src/slave/paging/rt_block_paging.cc:119-196).

    The volume at level `level` is the level-0 volume averaged over 2^level boxes (edge
    voxels repeated to a multiple of 2^level): this build's filter, the archive's own
    being outside it.  It is packed into bricks of the same edge and apron; level-0 brick
    (x, y, z) gets a copy of the coarse block (x, y, z) // s, s = 2^level, with the
    reference's mapping k = bs / pad / s, b = apron / pad + (x, y, z) mod s * k
    (rt_block_paging.cc:143-144,189), which the page table carries per page.
    u8 bricks hold the rounded box average; FP16 / FP32 bricks hold it as float (divided
    by 255 for a u8 volume when `normalize`, as the level-0 pack stores it)."""
    s = 1 << int(level)
    v = np.asarray(vol, np.float64)
    nz, ny, nx = v.shape
    v = np.pad(v, ((0, (-nz) % s), (0, (-ny) % s), (0, (-nx) % s)), mode="edge")
    low = v.reshape(v.shape[0] // s, s, v.shape[1] // s, s, v.shape[2] // s, s).mean(axis=(1, 3, 5))
    if fmt_out == L.FMT_U8:
        blocks = pack(np.floor(low + 0.5).astype(np.uint8), bs, apron, L.FMT_U8)
    else:
        vals = low / 255.0 if (normalize and np.asarray(vol).dtype == np.uint8) else low
        blocks = pack(vals.astype(np.float32), bs, apron, fmt_out, normalize=False)
    gl = [(d + bs - 1) // bs for d in low.shape[::-1]]          # coarse grid (x, y, z)
    g0 = [(d + bs - 1) // bs for d in (nx, ny, nz)]            # level-0 grid (x, y, z)
    pad = bs + 2 * apron
    k = np.float32(np.float32(bs) / np.float32(pad)) / np.float32(s)
    b0 = np.float32(apron) / np.float32(pad)
    n0 = g0[0] * g0[1] * g0[2]
    pages = np.empty((n0,) + blocks.shape[1:], blocks.dtype)
    page_k = np.full(n0, k, np.float32)
    page_b = np.empty((n0, 3), np.float32)
    i = 0
    for z in range(g0[2]):
        for y in range(g0[1]):
            for x in range(g0[0]):
                pages[i] = blocks[((z // s) * gl[1] + y // s) * gl[0] + x // s]
                page_b[i] = [b0 + np.float32(c % s) * k for c in (x, y, z)]
                i += 1
    return LowestLevel(pages, page_k, page_b, s)


def make_bricked(vol: np.ndarray, bs: int, apron: int, fmt_out: int, empty_fn=None, resident: int | float = 1.0,
                 seed: int = 7, normalize: bool = True) -> Bricked:
    """Brick ``vol`` (shape (z, y, x)); ``empty_fn(bmin, bmax) -> bool array`` marks empty bricks.

    ``resident``: number (int) or fraction (float) of bricks made resident by the
    seeded page table (drawn from the non-empty bricks).
    """
    grid, bmin, bmax, bmean = brick_stats(vol, bs, apron)
    empty = np.zeros(bmin.shape, np.uint8) if empty_fn is None else np.asarray(empty_fn(bmin, bmax), np.uint8)
    cheb = chebyshev(empty)
    nb = bmin.size
    n_res = int(round(resident * nb)) if isinstance(resident, float) else int(resident)
    dims = np.array(grid, np.int32)
    vaddr = np.empty(bmin.shape, np.int32)
    order = np.empty(nb, np.int32)
    nn = C.c_int32(0)
    empty_c = np.ascontiguousarray(empty)
    L.call("hyd_seeded_page_table", L.ptr(dims, C.c_int32), L.ptr(empty_c, C.c_uint8), n_res, seed,
           L.ptr(vaddr, C.c_int32), L.ptr(order, C.c_int32), C.byref(nn))
    n_pages = min(n_res, nn.value)
    pool = pack(vol, bs, apron, fmt_out, order[:n_pages], normalize=normalize)
    k, b = mapping(bs, apron)
    page_k = np.full(max(n_pages, 1), k, np.float32)
    page_b = np.full((max(n_pages, 1), 3), b, np.float32)
    raw = (not normalize) and vol.dtype == np.uint8 and fmt_out != L.FMT_U8
    return Bricked(dims=tuple(int(d) for d in vol.shape[::-1]), bs=bs, apron=apron, grid=tuple(grid), fmt=fmt_out,
                   pool=pool, vaddr=vaddr, page_k=page_k, page_b=page_b, cheb=cheb, mean=bmean, empty=empty,
                   order=order[:nn.value].copy(), n_nonempty=nn.value, bmin=bmin, bmax=bmax, raw_counts=raw)


def residency_batches(vol: np.ndarray, b: Bricked, pages_per_batch: int, n_batches: int | None = None):
    """Split the non-empty bricks (in the page-table order) into residency batches of at most
    pages_per_batch pages, like LosslessBlockPagingState::next (lossless_block_paging.cc:137-174):
    each batch gets its own pool (pages 0..n-1) and a page table mapping only its bricks.
    Returns (fmt, pad, pool, vaddr, page_k, page_b) tuples."""
    out = []
    order = b.order
    nb = (len(order) + pages_per_batch - 1) // pages_per_batch
    for i in range(nb if n_batches is None else min(nb, n_batches)):
        bricks = order[i * pages_per_batch:(i + 1) * pages_per_batch]
        pool = pack(vol, b.bs, b.apron, b.fmt, bricks, normalize=not b.raw_counts)
        vaddr = np.full(b.vaddr.shape, -1, np.int32)
        vaddr.reshape(-1)[bricks] = np.arange(len(bricks), dtype=np.int32)
        out.append((b.fmt, b.pad, pool, vaddr, b.page_k[:1].repeat(len(bricks)),
                    np.repeat(b.page_b[:1], len(bricks), axis=0)))
    return out


def single_brick(vol: np.ndarray, fmt_out: int = L.FMT_F32) -> Bricked:
    """One brick covering a cubic volume, no apron (block_size = raw dim): k = 1, b = 0."""
    n = vol.shape[0]
    assert vol.shape == (n, n, n), "single_brick needs a cubic volume"
    return make_bricked(vol, n, 0, fmt_out, empty_fn=None, resident=1)


def volume_empty_fn(tf: np.ndarray):
    """Volume-mode emptiness: the TF alpha is zero over the brick's [min, max]."""
    def fn(bmin, bmax):
        out = np.zeros(bmin.shape, bool)
        for idx in np.ndindex(bmin.shape):
            out[idx] = tf_alpha_zero_on(tf, float(bmin[idx]), float(bmax[idx]))
        return out
    return fn


def iso_empty_fn(iso: float):
    """Isosurface emptiness: the iso value lies outside the brick's [min, max]."""
    def fn(bmin, bmax):
        return (bmin > iso) | (bmax < iso)
    return fn


# ------------------------------------------------------------- scenes ------
README_TF = np.array([
    [0, 0, 0, 0], [1, 0, 0, 1], [1, 0.5, 0, 1], [1, 1, 0, 1], [0, 1, 0, 1],
    [0, 1, 1, 1], [0, 0, 1, 1], [1, 0, 1, 1], [0, 0, 0, 0]], np.float32)
"""9-point RGBA transfer function of the README Aneurism config (README.md:30-44)."""


@dataclass
class SceneGeometry:
    """BasicRenderer::init geometry (include/hydrant/basic_renderer.hpp:57-84), binary32."""
    f_dim: np.ndarray      # raw / block_size
    center: np.ndarray     # exhibit centre = f_dim / 2
    du: np.float32         # 1 / block_size
    step: np.float32       # du / 4 / sample_rate


def geometry(raw_dims, block_size: int, sample_rate: float = 1.0) -> SceneGeometry:
    raw = np.array(raw_dims, np.float32)
    blk = np.float32(block_size)
    f_dim = raw / blk
    center = f_dim / np.float32(2.0)
    du = np.float32(1.0) / blk
    step = (du / np.float32(4.0)) / np.float32(sample_rate)
    return SceneGeometry(f_dim=f_dim, center=center, du=np.float32(du), step=np.float32(step))
