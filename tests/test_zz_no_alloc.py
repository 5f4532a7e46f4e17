"""hyd_composite_noalloc (the z order staged in the library's fixed device array, no
allocation) against hyd_composite, through the C ABI: two calls in a row, the second with
fewer partials and a smaller image, then a call beyond the array's size is refused."""
import ctypes as C

import numpy as np
import pytest

from hydrant_amd import _lib as L

pytestmark = pytest.mark.gpu


def _partials(rng, n, H, W):
    p = np.zeros((n, H, W, 8), np.float32)
    for i in range(n):
        a = rng.random((H, W)).astype(np.float32) * 0.6
        rgb = rng.random((H, W, 3)).astype(np.float32) * a[..., None]
        p[i, ..., 0:3] = rgb
        p[i, ..., 3] = 1 - a
        p[i, ..., 4:7] = rgb
        p[i, ..., 7] = a
    return p


def _run(fn, h, parts_t, z, W, H):
    import torch

    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.uint8, device="cuda")
    za = np.ascontiguousarray(z, np.int32)
    L.call(fn, h, len(z), C.c_void_p(parts_t.data_ptr()), L.ptr(za, C.c_int32), W, H, C.c_void_p(rgba.data_ptr()),
           C.c_void_p(rgb.data_ptr()))
    return rgba.cpu().numpy(), rgb.cpu().numpy()


def test_composite_noalloc_matches_composite(gpu_ctx):
    import torch

    rng = np.random.default_rng(21)
    h = gpu_ctx.h
    for n, H, W in [(7, 48, 40), (3, 20, 33)]:
        parts = torch.from_numpy(_partials(rng, n, H, W)).cuda()
        z = rng.permutation(n)
        print(f"noalloc composite: {n} partials of {W}x{H}", flush=True)
        want = _run("hyd_composite", h, parts, z, W, H)
        got = _run("hyd_composite_noalloc", h, parts, z, W, H)
        assert np.array_equal(got[0].view(np.int32), want[0].view(np.int32))
        assert np.array_equal(got[1], want[1])
    big = 4097
    za = np.arange(big, dtype=np.int32)
    code = L.lib().hyd_composite_noalloc(h, big, C.c_void_p(parts.data_ptr()), L.ptr(za, C.c_int32), W, H,
                                         None, None)
    assert code == L.HYD_ERR_RANGE and b"at most 4096" in L.lib().hyd_last_error()
