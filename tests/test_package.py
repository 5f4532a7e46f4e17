"""CPU tests of the dataset package reader (hyd_package_read_info / _thumbnails,
include/hydrant_renderer.h): the committed package's sizes and thumbnails against
brute force over its voxels, write/read round trips for every voxel format, and
the reader's refusals of malformed packages.  The reference reads the same fields
of package_meta.json (core/renderer.hpp:53-59) and the level-0 "chebyshev" and
"mean" thumbnails (volume.cc:47-51, blocks.cc:35-44); varch's own binary layout of
those files is not available here, so the layout of the files is this library's
and only the meaning of the values is the reference's."""
import json
import shutil
from pathlib import Path

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import dataset as D

PKG = Path(__file__).resolve().parent / "golden" / "package_blobs_u8_44x40x30_bs8"


def _brute_stats(vol_norm: np.ndarray, bs: int, apron: int):
    """Apron-inclusive max (edge-clamped) and interior mean per brick, grid (z, y, x)."""
    nz, ny, nx = vol_norm.shape
    g = [(nx + bs - 1) // bs, (ny + bs - 1) // bs, (nz + bs - 1) // bs]
    bmax = np.zeros((g[2], g[1], g[0]), np.float32)
    bmean = np.zeros_like(bmax)
    for bz in range(g[2]):
        for by in range(g[1]):
            for bx in range(g[0]):
                zi = np.clip(np.arange(bz * bs - apron, bz * bs + bs + apron), 0, nz - 1)
                yi = np.clip(np.arange(by * bs - apron, by * bs + bs + apron), 0, ny - 1)
                xi = np.clip(np.arange(bx * bs - apron, bx * bs + bs + apron), 0, nx - 1)
                bmax[bz, by, bx] = vol_norm[np.ix_(zi, yi, xi)].max()
                inner = vol_norm[bz * bs:(bz + 1) * bs, by * bs:(by + 1) * bs, bx * bs:(bx + 1) * bs]
                bmean[bz, by, bx] = np.float32(inner.astype(np.float64).mean())
    return bmax, bmean


def _brute_chebyshev(empty: np.ndarray) -> np.ndarray:
    """0 for a brick with data, else the L-inf distance to the nearest such brick."""
    full = np.argwhere(~empty)
    out = np.zeros(empty.shape, np.int32)
    for idx in np.argwhere(empty):
        out[tuple(idx)] = np.abs(full - idx).max(axis=1).min()
    return out


def _voxels(root: Path) -> np.ndarray:
    meta = json.loads((root / "package_meta.json").read_text())
    lvl = meta["sample_levels"][0]
    dt = {"u8": np.uint8, "f16": "<f2", "f32": "<f4"}[lvl["voxels"]["format"]]
    r = lvl["raw"]
    return np.fromfile(root / lvl["voxels"]["path"], dt).reshape(r["z"], r["y"], r["x"])


def test_committed_package_info():
    info = D.read_package_info(PKG)
    assert list(info.raw) == [44, 40, 30]
    assert list(info.dim) == [6, 5, 4]
    assert (info.block_size, info.padding, info.fmt, info.has_mean) == (8, 1, L.FMT_U8, 1)


def test_committed_package_thumbnails_match_voxels():
    """The thumbnails the reader returns equal brute force over the package's voxels:
    empty = apron-inclusive max <= 0, chebyshev = L-inf distance, mean = interior mean."""
    vol = _voxels(PKG)
    cheb, mean = D.read_package_thumbnails(PKG)
    bmax, bmean = _brute_stats(vol.astype(np.float32) / np.float32(255), 8, 1)
    empty = bmax <= 0
    assert 0 < empty.sum() < empty.size
    want = _brute_chebyshev(empty)
    assert want.max() >= 2          # distances above one brick occur
    assert np.array_equal(cheb, want)
    np.testing.assert_allclose(mean, bmean, rtol=1e-6, atol=0)


@pytest.mark.parametrize("dtype", [np.uint8, np.float16, np.float32])
def test_write_read_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(3)
    vol = rng.uniform(0, 1, (13, 22, 17))
    vol[:, :, :9] = 0
    vol = (vol * 255).astype(np.uint8) if dtype == np.uint8 else vol.astype(dtype)
    meta = D.write_package(tmp_path, vol, block_size=4, padding=1)
    info = D.read_package_info(tmp_path)
    assert list(info.raw) == [17, 22, 13] and list(info.dim) == [5, 6, 4]
    assert info.fmt == {np.uint8: L.FMT_U8, np.float16: L.FMT_F16, np.float32: L.FMT_F32}[dtype]
    assert np.array_equal(_voxels(tmp_path), vol)        # f16: equal as numbers, bits stored
    cheb, mean = D.read_package_thumbnails(tmp_path)
    bmax, bmean = _brute_stats(D.normalised(vol) if dtype != np.float16 else vol.astype(np.float32), 4, 1)
    assert np.array_equal(cheb, _brute_chebyshev(bmax <= 0))
    np.testing.assert_allclose(mean, bmean, rtol=1e-6, atol=1e-7)
    assert meta["sample_levels"][0]["dim"] == {"x": 5, "y": 6, "z": 4}


def test_write_package_empty_override(tmp_path):
    """A caller-given emptiness (say, an iso value's) replaces max <= 0."""
    vol = np.full((8, 8, 16), 200, np.uint8)          # grid (z, y, x) = (2, 2, 4)
    empty = np.broadcast_to(np.array([False, True, True, True]), (2, 2, 4))
    D.write_package(tmp_path, vol, block_size=4, padding=0, empty=empty)
    cheb, _ = D.read_package_thumbnails(tmp_path)
    assert np.array_equal(cheb, np.broadcast_to(np.arange(4, dtype=np.int32), (2, 2, 4)))
    with pytest.raises(ValueError):
        D.write_package(tmp_path, vol, block_size=4, padding=0, empty=empty[:1])


def _copy(tmp_path) -> Path:
    dst = tmp_path / "pkg"
    shutil.copytree(PKG, dst)
    return dst


def _edit_meta(root: Path, fn):
    meta = json.loads((root / "package_meta.json").read_text())
    fn(meta)
    (root / "package_meta.json").write_text(json.dumps(meta))


@pytest.mark.parametrize("case, fragment", [
    ("no_meta", "cannot open"),
    ("dim_mismatch", "does not match"),
    ("bad_format", "format must be"),
    ("no_cheb", "chebyshev"),
    ("short_voxels", "voxels"),
    ("short_cheb", "chebyshev thumbnail"),
    ("short_mean", "mean thumbnail"),
    ("negative_cheb", "negative"),
    ("padding", "block_size / padding"),
    ("no_levels", "sample_levels"),
    ("fractional", "integer"),
])
def test_reader_refuses_malformed_packages(tmp_path, case, fragment):
    root = _copy(tmp_path)
    lvl = lambda m: m["sample_levels"][0]  # noqa: E731
    if case == "no_meta":
        (root / "package_meta.json").unlink()
    elif case == "dim_mismatch":
        _edit_meta(root, lambda m: lvl(m)["dim"].update(x=5))
    elif case == "bad_format":
        _edit_meta(root, lambda m: lvl(m)["voxels"].update(format="u16"))
    elif case == "no_cheb":
        _edit_meta(root, lambda m: lvl(m)["thumbnails"].pop("chebyshev"))
    elif case == "short_voxels":
        (root / "voxels.raw").write_bytes((root / "voxels.raw").read_bytes()[:-1])
    elif case == "short_cheb":
        (root / "thumb_chebyshev.i32").write_bytes((root / "thumb_chebyshev.i32").read_bytes()[:-4])
    elif case == "short_mean":
        (root / "thumb_mean.f32").write_bytes(b"")
    elif case == "negative_cheb":
        c = np.fromfile(root / "thumb_chebyshev.i32", "<i4")
        c[3] = -1
        c.tofile(root / "thumb_chebyshev.i32")
    elif case == "padding":
        _edit_meta(root, lambda m: m.update(padding=9))
    elif case == "no_levels":
        _edit_meta(root, lambda m: m.update(sample_levels=[]))
    elif case == "fractional":
        _edit_meta(root, lambda m: m.update(block_size=8.5))
    with pytest.raises(L.HydError) as e:
        D.read_package_info(root)
    assert e.value.code == L.HYD_ERR_ARG
    assert fragment in str(e.value)


def test_thumbnail_buffer_size_checked():
    import ctypes as C
    cheb = np.empty(10, np.int32)
    with pytest.raises(L.HydError):
        L.call("hyd_package_read_thumbnails", str(PKG).encode(), L.ptr(cheb, C.c_int32), None, cheb.size)
