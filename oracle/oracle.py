"""ctypes wrapper of the oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  The oracle is a
C++ restatement of the reference's multi-thread CPU backend (oracle.hpp);
precision 32 follows the reference's binary32 arithmetic operation by
operation, precision 64 is the FP64 ground truth.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

d3 = C.c_double * 3


class OView(C.Structure):
    _fields_ = [("trans", C.c_double * 16), ("ctg_fovy_2", C.c_double), ("ray_o", d3)]


class OVolumeArgs(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("bbox_min", d3), ("bbox_max", d3), ("step", C.c_double),
        ("du", C.c_double), ("max_steps", C.c_int), ("mode", C.c_int), ("rank", C.c_double),
        ("density", C.c_double), ("schedule", C.c_int), ("ert", C.c_double), ("tf_n", C.c_int),
        ("tf", C.POINTER(C.c_float)), ("cheb_dim", C.c_int * 3), ("cheb", C.POINTER(C.c_int)),
        ("vaddr_dim", C.c_int * 3), ("vaddr", C.POINTER(C.c_int)), ("lowest_blkcnt", C.c_int),
        ("page_k", C.POINTER(C.c_float)), ("page_b", C.POINTER(C.c_float)), ("pool", C.POINTER(C.c_float)),
        ("pad", C.c_int), ("map_bs", C.c_int),
    ]


class OBlocksArgs(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("bbox_min", d3), ("bbox_max", d3), ("step", C.c_double),
        ("max_steps", C.c_int), ("mode", C.c_int), ("density", C.c_double), ("dim", C.c_int * 3),
        ("cheb", C.POINTER(C.c_int)), ("mean", C.POINTER(C.c_float)),
        ("vaddr", C.POINTER(C.c_int)), ("vdim", C.c_int * 3), ("lowest", C.c_int),
    ]


class OIsoArgs(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("bbox_min", d3), ("bbox_max", d3), ("step", C.c_double),
        ("du", C.c_double), ("max_steps", C.c_int), ("mode", C.c_int), ("to_world", C.c_double * 16),
        ("eye_pos", d3), ("light_pos", d3), ("surface_color", d3), ("isovalue", C.c_double),
        ("cheb_dim", C.c_int * 3), ("cheb", C.POINTER(C.c_int)), ("vaddr_dim", C.c_int * 3),
        ("vaddr", C.POINTER(C.c_int)), ("page_k", C.POINTER(C.c_float)), ("page_b", C.POINTER(C.c_float)),
        ("pool", C.POINTER(C.c_float)), ("pad", C.c_int), ("corrected", C.c_int), ("bisect_steps", C.c_int),
    ]


class OTrace(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ix", "iy", "iz", "pgid", "tix", "tiy", "tiz", "tjx", "tjy", "tjz")]


_lib = None


def build() -> Path:
    res = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        PD = C.POINTER(C.c_double)
        L.orc_orbit_position.argtypes = [C.c_int, PD, PD]
        L.orc_make_view.argtypes = [C.c_int, PD, PD, PD, C.c_double, PD, PD, C.POINTER(OView)]
        L.orc_inverse.argtypes = [C.c_int, PD, PD]
        L.orc_exhibit_iet.argtypes = [C.c_int, PD, PD, PD]
        L.orc_gen_ray.argtypes = [C.c_int, C.POINTER(OView), C.c_int, C.c_int, C.c_int, C.c_int, PD]
        L.orc_intersect.argtypes = [C.c_int, PD, PD, PD, PD, PD]
        L.orc_brick_sample.argtypes = [C.c_int, C.POINTER(C.c_float), C.c_int, PD, C.POINTER(C.c_int)]
        L.orc_brick_sample.restype = C.c_double
        L.orc_tf_sample.argtypes = [C.c_int, C.POINTER(C.c_float), C.c_int, C.c_double, PD]
        L.orc_volume_render.argtypes = [C.c_int, C.c_int, C.POINTER(OView), C.POINTER(OVolumeArgs), PD, PD,
                                        C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.orc_volume_resume.argtypes = [C.c_int, C.c_int, C.POINTER(OVolumeArgs), PD, C.POINTER(C.c_longlong)]
        L.orc_volume_render_pixels.argtypes = [C.c_int, C.c_int, C.POINTER(OView), C.POINTER(OVolumeArgs), C.c_int,
                                               C.POINTER(C.c_int), PD, C.POINTER(C.c_longlong)]
        L.orc_volume_trace.argtypes = [C.POINTER(OView), C.POINTER(OVolumeArgs), C.c_int, C.c_int,
                                       C.POINTER(OTrace), C.c_int]
        L.orc_pixel_pass.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_ubyte),
                                     C.POINTER(C.c_ubyte)]
        L.orc_composite.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                    C.POINTER(C.c_float), C.POINTER(C.c_ubyte)]
        L.orc_blocks_render.argtypes = [C.c_int, C.c_int, C.POINTER(OView), C.POINTER(OBlocksArgs), PD]
        L.orc_iso_render.argtypes = [C.c_int, C.c_int, C.POINTER(OView), C.POINTER(OIsoArgs), PD,
                                     C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def orbit_position(arm, prec=32) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_orbit_position(prec, _pd(np.asarray(arm, np.float64)), _pd(out))
    return out


def make_view(pos, target, up, fovy, ex_center, ex_size, prec=32) -> OView:
    v = OView()
    a = [np.asarray(x, np.float64) for x in (pos, target, up, ex_center, ex_size)]
    lib().orc_make_view(prec, _pd(a[0]), _pd(a[1]), _pd(a[2]), float(fovy), _pd(a[3]), _pd(a[4]), C.byref(v))
    return v


def view_from_dict(d: dict) -> OView:
    v = OView()
    v.trans[:] = [float(x) for x in d["trans"]]
    v.ctg_fovy_2 = float(d["ctg_fovy_2"])
    v.ray_o[:] = [float(x) for x in d["ray_o"]]
    return v


def view_to_dict(v: OView) -> dict:
    return {"trans": np.array(v.trans[:]), "ctg_fovy_2": v.ctg_fovy_2, "ray_o": np.array(v.ray_o[:])}


def gen_ray(view: OView, x, y, W, H, prec=32) -> np.ndarray:
    out = np.zeros(6)
    lib().orc_gen_ray(prec, C.byref(view), x, y, W, H, _pd(out))
    return out


def intersect(o, d, bmin, bmax, prec=32):
    out = np.zeros(2)
    a = [np.asarray(x, np.float64) for x in (o, d, bmin, bmax)]
    hit = lib().orc_intersect(prec, _pd(a[0]), _pd(a[1]), _pd(a[2]), _pd(a[3]), _pd(out))
    return bool(hit), out[0], out[1]


def brick_sample(data: np.ndarray, u, prec=32):
    data = np.ascontiguousarray(data, np.float32)
    taps = np.zeros(6, np.int32)
    r = lib().orc_brick_sample(prec, _pf(data), data.shape[0], _pd(np.asarray(u, np.float64)), _pi(taps))
    return r, taps


def tf_sample(tf: np.ndarray, s: float, prec=32) -> np.ndarray:
    tf = np.ascontiguousarray(tf, np.float32)
    out = np.zeros(4)
    lib().orc_tf_sample(prec, _pf(tf), tf.size // 4, float(s), _pd(out))
    return out


def inverse(m: np.ndarray, prec=32) -> np.ndarray:
    out = np.zeros(16)
    lib().orc_inverse(prec, _pd(np.ascontiguousarray(m, np.float64)), _pd(out))
    return out


def exhibit_iet(center, size, prec=32) -> np.ndarray:
    out = np.zeros(16)
    lib().orc_exhibit_iet(prec, _pd(np.asarray(center, np.float64)), _pd(np.asarray(size, np.float64)), _pd(out))
    return out


class VolumeScene:
    """Arguments of one volume render held alive for the oracle."""

    def __init__(self, W, H, bbox_min, bbox_max, step, du, tf, cheb, vaddr, page_k, page_b, pool_values, pad,
                 max_steps=4_000_000, mode=0, rank=0.0, density=1.0, schedule=0, ert=0.0, lowest_blkcnt=0,
                 map_bs=0):
        self.keep = [np.ascontiguousarray(tf, np.float32).reshape(-1), np.ascontiguousarray(cheb, np.int32),
                     np.ascontiguousarray(vaddr, np.int32), np.ascontiguousarray(page_k, np.float32),
                     np.ascontiguousarray(page_b, np.float32).reshape(-1),
                     np.ascontiguousarray(pool_values, np.float32)]
        tfa, ch, va, pk, pb, pv = self.keep
        a = OVolumeArgs()
        a.W, a.H = W, H
        a.bbox_min[:] = [float(x) for x in bbox_min]
        a.bbox_max[:] = [float(x) for x in bbox_max]
        a.step, a.du = float(step), float(du)
        a.max_steps, a.mode, a.rank, a.density = int(max_steps), int(mode), float(rank), float(density)
        a.schedule, a.ert = int(schedule), float(ert)
        a.tf_n, a.tf = tfa.size // 4, _pf(tfa)
        a.cheb_dim[:] = list(ch.shape[::-1])
        a.cheb = _pi(ch)
        a.vaddr_dim[:] = list(va.shape[::-1])
        a.vaddr = _pi(va)
        a.lowest_blkcnt = int(lowest_blkcnt)
        a.page_k, a.page_b, a.pool, a.pad = _pf(pk), _pf(pb), _pf(pv), int(pad)
        a.map_bs = int(map_bs)
        self.args = a

    def render(self, view: OView, prec=32, nthreads=0, film=True, counts=False):
        a = self.args
        P = a.W * a.H
        film_a = np.zeros((a.H, a.W, 16)) if film else None
        rgba = np.zeros((a.H, a.W, 4))
        cnt = np.zeros((a.H, a.W), np.int64) if counts else None
        tot = C.c_longlong(0)
        lib().orc_volume_render(prec, nthreads, C.byref(view), C.byref(a), _pd(film_a) if film else None, _pd(rgba),
                                cnt.ctypes.data_as(C.POINTER(C.c_longlong)) if counts else None, C.byref(tot))
        return {"film": film_a, "rgba": rgba, "counts": cnt, "samples": tot.value}

    def render_pixels(self, view: OView, xy: np.ndarray, prec=64, nthreads=0):
        """RGBA of the listed pixels (xy: (n, 2) int array of x, y)."""
        xy = np.ascontiguousarray(xy, np.int32)
        out = np.zeros((xy.shape[0], 4))
        tot = C.c_longlong(0)
        lib().orc_volume_render_pixels(prec, nthreads, C.byref(view), C.byref(self.args), xy.shape[0], _pi(xy),
                                       _pd(out), C.byref(tot))
        return out, tot.value

    def resume(self, film: np.ndarray, prec=32, nthreads=0):
        film = np.ascontiguousarray(film, np.float64).copy()
        tot = C.c_longlong(0)
        lib().orc_volume_resume(prec, nthreads, C.byref(self.args), _pd(film), C.byref(tot))
        return {"film": film, "rgba": film[..., 8:12].copy(), "samples": tot.value}

    def trace(self, view: OView, x, y, max_recs=8192) -> np.ndarray:
        recs = (OTrace * max_recs)()
        n = lib().orc_volume_trace(C.byref(view), C.byref(self.args), x, y, recs, max_recs)
        return np.frombuffer(recs, dtype=np.int32).reshape(max_recs, 10)[:n].copy()


class ODiffArgs(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("K", C.c_int), ("sigma", C.c_double), ("dims", C.c_int * 3),
                ("vol", C.POINTER(C.c_float)), ("tf", C.POINTER(C.c_float)), ("tf_n", C.c_int)]


class DiffScene:
    """Differentiable-render oracle: forward RGBA and exact adjoint on listed pixels."""

    def __init__(self, W, H, volume: np.ndarray, tf: np.ndarray, K=512, sigma=8.0):
        self.vol = np.ascontiguousarray(volume, np.float32)
        self.tf = np.ascontiguousarray(tf, np.float32).reshape(-1)
        a = ODiffArgs()
        a.W, a.H, a.K, a.sigma = W, H, int(K), float(sigma)
        a.dims[:] = list(self.vol.shape[::-1])
        a.vol, a.tf, a.tf_n = _pf(self.vol), _pf(self.tf), self.tf.size // 4
        self.args = a
        L = lib()
        if not getattr(L, "_diff_typed", False):
            L.orc_diff_render.argtypes = [C.c_int, C.c_int, C.POINTER(OView), C.POINTER(ODiffArgs), C.c_int,
                                          C.POINTER(C.c_int), C.POINTER(C.c_double)]
            L.orc_diff_grad.argtypes = [C.c_int, C.POINTER(OView), C.POINTER(ODiffArgs), C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
            L._diff_typed = True

    def render(self, view: OView, xy: np.ndarray, prec=64, nthreads=0) -> np.ndarray:
        xy = np.ascontiguousarray(xy, np.int32)
        out = np.zeros((xy.shape[0], 4))
        lib().orc_diff_render(prec, nthreads, C.byref(view), C.byref(self.args), xy.shape[0], _pi(xy), _pd(out))
        return out

    def grad(self, view: OView, xy: np.ndarray, grad_rgba: np.ndarray, prec=64):
        """(dL/dvolume, dL/dTF) for the listed pixels given dL/dRGBA (n x 4)."""
        xy = np.ascontiguousarray(xy, np.int32)
        g = np.ascontiguousarray(grad_rgba, np.float64)
        gv = np.zeros(self.vol.size)
        gt = np.zeros(self.tf.size)
        lib().orc_diff_grad(prec, C.byref(view), C.byref(self.args), xy.shape[0], _pi(xy), _pd(g), _pd(gv), _pd(gt))
        return gv.reshape(self.vol.shape), gt.reshape(-1, 4)


def pixel_pass(rgba: np.ndarray, clear=(0, 0, 0)) -> np.ndarray:
    rgba = np.ascontiguousarray(rgba, np.float32)
    H, W = rgba.shape[:2]
    out = np.zeros((H, W, 3), np.uint8)
    cc = (C.c_ubyte * 3)(*clear)
    lib().orc_pixel_pass(W, H, _pf(rgba), cc, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return out


def composite(partials: np.ndarray, z_order) -> tuple[np.ndarray, np.ndarray]:
    partials = np.ascontiguousarray(partials, np.float32)
    N, H, W = partials.shape[:3]
    z = np.ascontiguousarray(z_order, np.int32)
    rgba = np.zeros((H, W, 4), np.float32)
    rgb = np.zeros((H, W, 3), np.uint8)
    lib().orc_composite(N, W, H, _pf(partials), _pi(z), _pf(rgba), rgb.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return rgba, rgb


def srgb8_check(thresholds: np.ndarray, nthreads=0) -> tuple[int, int]:
    """Mismatches of a 256-entry sRGB byte threshold table against the oracle's bytes
    over every float in [0, 2] plus special values; (count, bits of the first)."""
    L = lib()
    if not getattr(L, "_srgb_typed", False):
        L.orc_srgb8_check.argtypes = [C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_uint)]
        L.orc_srgb8_check.restype = C.c_longlong
        L._srgb_typed = True
    t = np.ascontiguousarray(thresholds, np.float32)
    fb = C.c_uint(0)
    n = L.orc_srgb8_check(_pf(t), int(nthreads), C.byref(fb))
    return int(n), int(fb.value)


def iso_fetch(px: np.ndarray):
    """IsosurfaceShaderKernel::fetch of (H, W, 6) isosurface pixels: (bytes (H, W, 3), depth (H, W))."""
    px = np.ascontiguousarray(px, np.float64)
    H, W = px.shape[:2]
    val = np.zeros((H, W, 3), np.uint8)
    depth = np.zeros((H, W), np.float32)
    L = lib()
    L.orc_iso_fetch.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_ubyte), C.POINTER(C.c_float)]
    L.orc_iso_fetch(H * W, _pd(px), val.ctypes.data_as(C.POINTER(C.c_ubyte)), _pf(depth))
    return val, depth


def composite_iso(val: np.ndarray, depth: np.ndarray, z_order):
    """Depth-min sort-last composite of N isosurface partials: val (N, H, W, 3) bytes,
    depth (N, H, W); returns (rgb (H, W, 3), depth (H, W))."""
    val = np.ascontiguousarray(val, np.uint8)
    depth = np.ascontiguousarray(depth, np.float32)
    N, H, W = depth.shape
    z = np.ascontiguousarray(z_order, np.int32)
    rgb = np.zeros((H, W, 3), np.uint8)
    dout = np.zeros((H, W), np.float32)
    L = lib()
    L.orc_composite_iso.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_ubyte), C.POINTER(C.c_float),
                                    C.POINTER(C.c_int), C.POINTER(C.c_ubyte), C.POINTER(C.c_float)]
    L.orc_composite_iso(N, W, H, val.ctypes.data_as(C.POINTER(C.c_ubyte)), _pf(depth), _pi(z),
                        rgb.ctypes.data_as(C.POINTER(C.c_ubyte)), _pf(dout))
    return rgb, dout


def blocks_render(view: OView, W, H, bbox_min, bbox_max, step, cheb, mean, mode=0, density=1e-2,
                  max_steps=4_000_000, prec=32, nthreads=0, vaddr=None, lowest=0) -> np.ndarray:
    """Blocks.Volume (mode 0) / Blocks.Solid (1), or the Paging view (2) over page table `vaddr` (z, y, x)."""
    ch = np.ascontiguousarray(cheb, np.int32)
    mn = np.ascontiguousarray(mean if mean is not None else np.zeros(ch.shape), np.float32)
    va = np.ascontiguousarray(vaddr if vaddr is not None else np.full(ch.shape, -1), np.int32)
    a = OBlocksArgs()
    a.W, a.H = W, H
    a.bbox_min[:] = [float(x) for x in bbox_min]
    a.bbox_max[:] = [float(x) for x in bbox_max]
    a.step, a.max_steps, a.mode, a.density = float(step), int(max_steps), int(mode), float(density)
    a.dim[:] = list(ch.shape[::-1])
    a.cheb, a.mean = _pi(ch), _pf(mn)
    a.vaddr, a.lowest = _pi(va), int(lowest)
    a.vdim[:] = list(va.shape[::-1])
    rgba = np.zeros((H, W, 4))
    lib().orc_blocks_render(prec, nthreads, C.byref(view), C.byref(a), _pd(rgba))
    return rgba


class IsoScene:
    def __init__(self, W, H, bbox_min, bbox_max, step, du, to_world, eye_pos, light_pos, surface_color, isovalue,
                 cheb, vaddr, page_k, page_b, pool_values, pad, mode=0, corrected=0, bisect_steps=8,
                 max_steps=4_000_000):
        self.keep = [np.ascontiguousarray(cheb, np.int32), np.ascontiguousarray(vaddr, np.int32),
                     np.ascontiguousarray(page_k, np.float32), np.ascontiguousarray(page_b, np.float32).reshape(-1),
                     np.ascontiguousarray(pool_values, np.float32)]
        ch, va, pk, pb, pv = self.keep
        a = OIsoArgs()
        a.W, a.H = W, H
        a.bbox_min[:] = [float(x) for x in bbox_min]
        a.bbox_max[:] = [float(x) for x in bbox_max]
        a.step, a.du, a.max_steps, a.mode = float(step), float(du), int(max_steps), int(mode)
        a.to_world[:] = [float(x) for x in np.asarray(to_world).reshape(-1)]
        a.eye_pos[:] = [float(x) for x in eye_pos]
        a.light_pos[:] = [float(x) for x in light_pos]
        a.surface_color[:] = [float(x) for x in surface_color]
        a.isovalue = float(isovalue)
        a.cheb_dim[:] = list(ch.shape[::-1])
        a.cheb = _pi(ch)
        a.vaddr_dim[:] = list(va.shape[::-1])
        a.vaddr = _pi(va)
        a.page_k, a.page_b, a.pool, a.pad = _pf(pk), _pf(pb), _pf(pv), int(pad)
        a.corrected, a.bisect_steps = int(corrected), int(bisect_steps)
        self.args = a

    def render(self, view: OView, prec=32, nthreads=0):
        a = self.args
        out = np.zeros((a.H, a.W, 6))
        tot = C.c_longlong(0)
        lib().orc_iso_render(prec, nthreads, C.byref(view), C.byref(a), _pd(out), C.byref(tot))
        return {"px": out, "samples": tot.value}
