"""ctypes binding of libhydrant_hip.so (include/hydrant_hip.h, hydrant_dataset.h, hydrant_renderer.h).

The library is built in-tree by ``hydrant_amd.build.build_native`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there is
no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhydrant_hip.so"

HYD_OK, HYD_ERR_ARG, HYD_ERR_HIP, HYD_ERR_STATE, HYD_ERR_RANGE = 0, 1, 2, 3, 4
FMT_U8, FMT_F16, FMT_F32 = 0, 1, 2
MODE_DEFAULT, MODE_PARTITION, MODE_PAGING = 0, 1, 2
SCHED_REFERENCE, SCHED_LATTICE = 0, 1
BLOCKS_VOLUME, BLOCKS_SOLID, BLOCKS_PAGING = 0, 1, 2
ISO_COLOR, ISO_NORMAL, ISO_POSITION = 0, 1, 2

f3 = C.c_float * 3
f16 = C.c_float * 16
i3 = C.c_int32 * 3


class View(C.Structure):
    _fields_ = [("trans", f16), ("ctg_fovy_2", C.c_float), ("ray_o", f3)]


class Camera(C.Structure):
    _fields_ = [("position", f3), ("target", f3), ("up", f3), ("fovy", C.c_float)]


class VolumeParams(C.Structure):
    _fields_ = [
        ("bbox_min", f3), ("bbox_max", f3), ("step", C.c_float), ("du", C.c_float),
        ("max_steps", C.c_int32), ("mode", C.c_int32), ("rank", C.c_float), ("density", C.c_float),
        ("schedule", C.c_int32), ("ert", C.c_float),
    ]


class VolumePixel(C.Structure):
    _fields_ = [("o", f3), ("d", f3), ("nsteps", C.c_int32), ("step_size", C.c_int32),
                ("v", C.c_float * 4), ("theta", f3), ("phi", C.c_float)]


class FetchPixel(C.Structure):
    _fields_ = [("theta", f3), ("phi", C.c_float), ("val", C.c_float * 4)]


class TraceRec(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("ix", "iy", "iz", "pgid", "tix", "tiy", "tiz", "tjx", "tjy", "tjz")]


class BlocksParams(C.Structure):
    _fields_ = [("bbox_min", f3), ("bbox_max", f3), ("step", C.c_float), ("max_steps", C.c_int32),
                ("mode", C.c_int32), ("density", C.c_float)]


class IsoParams(C.Structure):
    _fields_ = [
        ("bbox_min", f3), ("bbox_max", f3), ("step", C.c_float), ("du", C.c_float),
        ("max_steps", C.c_int32), ("mode", C.c_int32), ("to_world", f16), ("eye_pos", f3),
        ("light_pos", f3), ("surface_color", f3), ("isovalue", C.c_float), ("corrected", C.c_int32),
        ("bisect_steps", C.c_int32),
    ]


class IsoPixel(C.Structure):
    _fields_ = [("v", C.c_float * 4), ("depth", C.c_float), ("hit", C.c_int32)]


P = C.c_void_p
I32 = C.c_int32
PI32 = C.POINTER(C.c_int32)
PF = C.POINTER(C.c_float)


class VolumeDesc(C.Structure):
    _fields_ = [("grid", i3), ("bs", I32), ("apron", I32), ("bbox_max", f3), ("step", C.c_float),
                ("value_scale", C.c_float), ("vaddr", PI32), ("chebyshev", PI32)]


class BatchView(C.Structure):
    _fields_ = [("view", View), ("width", I32), ("height", I32), ("volume", I32), ("reserved", I32),
                ("out_offset", C.c_int64)]


class DiffParams(C.Structure):
    _fields_ = [("dims", i3), ("n_steps", I32), ("ckpt_every", I32), ("sigma", C.c_float)]


class Dataset(C.Structure):
    """hyd_dataset (include/hydrant_renderer.h)."""
    _fields_ = [("dims", C.c_int32 * 3), ("fmt", C.c_int32), ("voxels", C.c_void_p), ("block_size", C.c_int32),
                ("apron", C.c_int32)]


class SegmentBatch(C.Structure):
    """hyd_segment_batch (include/hydrant_hip.h)."""
    _fields_ = [("fmt", C.c_int32), ("pad_edge", C.c_int32), ("n_pages", C.c_int32), ("pool_dev", C.c_void_p),
                ("vaddr", C.POINTER(C.c_int32)), ("page_k", C.POINTER(C.c_float)), ("page_b", C.POINTER(C.c_float))]


class PackageInfo(C.Structure):
    """hyd_package_info (include/hydrant_renderer.h)."""
    _fields_ = [("raw", C.c_int32 * 3), ("dim", C.c_int32 * 3), ("block_size", C.c_int32), ("padding", C.c_int32),
                ("fmt", C.c_int32), ("has_mean", C.c_int32)]


class BatchConfig(C.Structure):
    _fields_ = [("tile_w", I32), ("unroll", I32), ("waves_per_cu", I32), ("group", I32)]

    def as_dict(self):
        g = max(1, self.group)
        return {"tile_w": self.tile_w * (2 if g == 4 else 1), "tile_h": 16 if g == 4 else 8, "unroll": self.unroll,
                "waves_per_cu": self.waves_per_cu, "waves_per_group": g}

# name -> (restype, argtypes); every symbol the public headers declare.
SIGNATURES = {
    # hydrant_hip.h
    "hyd_last_error": (C.c_char_p, []),
    "hyd_version": (I32, []),
    "hyd_create": (I32, [I32, C.POINTER(P)]),
    "hyd_destroy": (I32, [P]),
    "hyd_set_stream": (I32, [P, P]),
    "hyd_synchronize": (I32, [P]),
    "hyd_orbit_camera": (I32, [PF, PF, C.c_float, C.POINTER(Camera)]),
    "hyd_make_view": (I32, [C.POINTER(Camera), PF, PF, C.POINTER(View)]),
    "hyd_exhibit_iet": (I32, [PF, PF, PF]),
    "hyd_mat4_inverse": (I32, [PF, PF]),
    "hyd_upload_pool": (I32, [P, I32, I32, I32, P]),
    "hyd_bind_pool_device": (I32, [P, I32, I32, I32, P]),
    "hyd_write_pool_pages": (I32, [P, I32, I32, P]),
    "hyd_set_page_table": (I32, [P, PI32, PI32, I32, PF, PF, I32]),
    "hyd_set_empty_grid": (I32, [P, PI32, PI32]),
    "hyd_set_mean_grid": (I32, [P, PI32, PF]),
    "hyd_set_tf": (I32, [P, I32, PF]),
    "hyd_render_volume": (I32, [P, C.POINTER(VolumeParams), C.POINTER(View), I32, I32, P, P, P]),
    "hyd_resume_volume": (I32, [P, C.POINTER(VolumeParams), I32, I32, P, P, P]),
    "hyd_pixel_pass": (I32, [P, P, I32, I32, C.POINTER(C.c_uint8), P]),
    "hyd_fetch_pass": (I32, [P, P, I32, I32, P]),
    "hyd_composite": (I32, [P, I32, P, PI32, I32, I32, P, P]),
    "hyd_composite_noalloc": (I32, [P, I32, P, PI32, I32, I32, P, P]),
    "hyd_srgb8_thresholds": (I32, [PF]),
    "hyd_iso_fetch_pass": (I32, [P, P, I32, I32, P]),
    "hyd_composite_iso": (I32, [P, I32, P, PI32, I32, I32, P, P]),
    "hyd_segments_begin": (I32, [P, I32, I32, I32, P]),
    "hyd_segments_pass": (I32, [P, C.POINTER(VolumeParams), C.POINTER(View), I32, I32, I32, P, P, P]),
    "hyd_segment_set_create": (I32, [P, I32, C.POINTER(SegmentBatch), C.POINTER(P)]),
    "hyd_segment_set_destroy": (I32, [P]),
    "hyd_segments_pass_set": (I32, [P, P, C.POINTER(VolumeParams), C.POINTER(View), I32, I32, I32, P, P, P]),
    "hyd_segments_composite": (I32, [P, I32, I32, I32, P, P, P]),
    "hyd_trace_volume": (I32, [P, C.POINTER(VolumeParams), C.POINTER(View), I32, I32, I32, I32,
                               C.POINTER(TraceRec), I32, PI32]),
    "hyd_render_blocks": (I32, [P, C.POINTER(BlocksParams), C.POINTER(View), I32, I32, P]),
    "hyd_render_iso": (I32, [P, C.POINTER(IsoParams), C.POINTER(View), I32, I32, P, P]),
    "hyd_diff_ckpt_floats": (I32, [C.POINTER(DiffParams), PI32]),
    "hyd_diff_forward": (I32, [P, C.POINTER(DiffParams), C.POINTER(View), I32, I32, I32, P, P, I32, P, P, P]),
    "hyd_diff_backward": (I32, [P, C.POINTER(DiffParams), C.POINTER(View), I32, I32, I32, P, P, I32, P, P, P, P]),
    "hyd_set_volumes": (I32, [P, I32, C.POINTER(VolumeDesc)]),
    "hyd_batch_create": (I32, [P, C.POINTER(BatchView), I32, C.POINTER(P)]),
    "hyd_batch_destroy": (I32, [P]),
    "hyd_batch_output_pixels": (I32, [P, C.POINTER(C.c_int64)]),
    "hyd_batch_render": (I32, [P, P, C.POINTER(BatchConfig), C.c_float, C.c_float, P, P]),
    "hyd_batch_capture": (I32, [P, P, C.POINTER(BatchConfig), C.c_float, C.c_float, P]),
    "hyd_batch_replay": (I32, [P, P]),
    "hyd_batch_autotune": (I32, [P, P, C.c_float, C.c_float, P, I32, C.POINTER(BatchConfig),
                                 C.POINTER(C.c_float)]),
    "hyd_batch_autotune_schedule": (I32, [P, P, C.c_float, C.c_float, P, I32, C.POINTER(BatchConfig),
                                          C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    # hydrant_dataset.h
    "hyd_gen_blobs_u8": (I32, [I32, I32, I32, I32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_double, P]),
    "hyd_gen_perlin_u8": (I32, [I32, I32, C.c_float, C.c_float, C.c_float, C.c_uint64, P]),
    "hyd_gen_sdf_spheres_f16": (I32, [I32, I32, C.c_float, C.c_float, C.c_float, C.c_uint64, P]),
    "hyd_f32_to_f16": (C.c_uint16, [C.c_float]),
    "hyd_f16_to_f32": (C.c_float, [C.c_uint16]),
    "hyd_pack_bricks": (I32, [P, I32, PI32, I32, I32, I32, I32, PI32, I32, P]),
    "hyd_brick_stats": (I32, [P, I32, PI32, I32, I32, PF, PF, PF]),
    "hyd_tf_alpha_zero_on": (I32, [PF, I32, C.c_float, C.c_float]),
    "hyd_chebyshev": (I32, [PI32, C.POINTER(C.c_uint8), PI32]),
    "hyd_seeded_page_table": (I32, [PI32, C.POINTER(C.c_uint8), I32, C.c_uint64, PI32, PI32, PI32]),
    # hydrant_residency.h
    "hyd_cull_bricks": (I32, [C.POINTER(View), PI32, PI32, PI32, PI32, C.c_int64, PI32, C.POINTER(C.c_int64)]),
    "hyd_residency_create": (I32, [PI32, I32, I32, PI32, C.POINTER(P)]),
    "hyd_residency_destroy": (I32, [P]),
    "hyd_residency_update": (I32, [P, C.POINTER(View), PI32, PI32, C.POINTER(C.c_int64)]),
    "hyd_residency_commit": (I32, [P, I32, PI32]),
    "hyd_residency_vaddr": (I32, [P, PI32]),
    "hyd_residency_stats": (I32, [P, C.POINTER(C.c_int64), PI32]),
    # hydrant_renderer.h
    "hyd_renderer_candidates": (I32, [C.c_char_p, I32, C.POINTER(I32)]),
    "hyd_renderer_create": (I32, [P, C.POINTER(Dataset), C.c_char_p, C.POINTER(P)]),
    "hyd_package_read_info": (I32, [C.c_char_p, C.POINTER(PackageInfo)]),
    "hyd_package_read_thumbnails": (I32, [C.c_char_p, PI32, PF, C.c_int64]),
    "hyd_renderer_create_package": (I32, [P, C.c_char_p, C.c_char_p, C.POINTER(P)]),
    "hyd_renderer_update": (I32, [P, C.c_char_p]),
    "hyd_renderer_resolution": (I32, [P, C.POINTER(I32), C.POINTER(I32)]),
    "hyd_renderer_offline_render": (I32, [P, C.c_char_p, P]),
    "hyd_renderer_destroy": (I32, [P]),
    "hyd_renderer_parse_config": (I32, [C.c_char_p, C.c_char_p, I32, C.POINTER(I32), C.POINTER(I32)]),
    "hyd_renderer_parse_camera": (I32, [C.c_char_p, C.POINTER(Camera)]),
}

_lib = None


class HydError(RuntimeError):
    """A libhydrant_hip call returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libhydrant_hip.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the HIP path has no fallback)")
        handle = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(fn: str, code: int) -> None:
    if code != HYD_OK:
        msg = lib().hyd_last_error().decode(errors="replace")
        raise HydError(fn, code, msg)


def call(fn: str, *args) -> None:
    check(fn, getattr(lib(), fn)(*args))


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array's data as POINTER(ctype)."""
    return arr.ctypes.data_as(C.POINTER(ctype))
