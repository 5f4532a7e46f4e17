"""GPU tests of rendering from a dataset package on disk (hyd_renderer_create_package,
the reference's RendererFactory(dataset_path), core/renderer.hpp:53-59): the Volume
image against the oracle with the package's chebyshev thumbnail as the empty-space
grid, the thumbnail's effect on the image, and Blocks / Isosurface images equal to the
raw-volume renderer's where both see the same empty bricks and means."""
import copy

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import dataset as D
from hydrant_amd import workloads as W
from hydrant_amd.renderer import Renderer
from oracle import oracle as O
from tests.helpers import oracle_view, oracle_volume
from tests.test_package import PKG, _voxels
from tests.test_renderer_config import README_CONFIG

pytestmark = pytest.mark.gpu


def _cfg(res=96, sample_rate=0.5):
    cfg = copy.deepcopy(README_CONFIG)
    cfg["render"]["resolution"] = [res, res]
    cfg["render"]["params"]["sample_rate"] = sample_rate
    return cfg


def test_volume_from_package_matches_oracle(gpu_ctx):
    """Volume over the committed package: bricks the thumbnail marks (max <= 0) are
    skipped, not those the transfer function makes transparent; bit-exact RGB8."""
    vol = _voxels(PKG)
    cfg = _cfg()
    tf = np.array(cfg["render"]["params"]["transfer_fn"]["values"], np.float32)
    b = D.make_bricked(vol, 8, 1, L.FMT_U8, lambda bmin, bmax: bmax <= 0, resident=1.0, seed=1)
    cheb, _ = D.read_package_thumbnails(PKG)
    assert np.array_equal(b.cheb, cheb)
    sc = W.Scene("package_volume", b, tf, 96, 96, (25.0, 65.0, 2.2), 60.0, D.geometry(vol.shape[::-1], 8, 0.5),
                 density=1e-2, extra={"volume": vol})
    r = Renderer.from_package(gpu_ctx, PKG, cfg)
    assert (r.W, r.H) == (96, 96)
    img = r.offline_render(cfg)
    o = oracle_volume(sc).render(oracle_view(sc), 32, film=False)
    ref = O.pixel_pass(o["rgba"])
    print(f"package Volume: {int(np.count_nonzero(img != ref))} differing bytes, "
          f"lit {int(np.count_nonzero(img.sum(-1)))}", flush=True)
    assert np.array_equal(img, ref)
    assert np.count_nonzero(img.sum(-1)) > 100
    r.close()


def test_volume_thumbnail_marks_are_honoured(gpu_ctx, tmp_path):
    """A thumbnail that marks every brick with data empty renders the clear colour;
    marking one brick changes the image only where rays cross it."""
    vol = _voxels(PKG)
    cfg = _cfg(res=64)
    r = Renderer.from_package(gpu_ctx, PKG, cfg)
    full = r.offline_render(cfg)
    r.close()
    D.write_package(tmp_path / "all", vol, 8, 1, empty=np.ones((4, 5, 6), bool))
    r = Renderer.from_package(gpu_ctx, tmp_path / "all", cfg)
    assert not np.any(r.offline_render(cfg))
    r.close()
    _, _, bmax, _ = D.brick_stats(vol, 8, 1)
    one = bmax <= 0
    one.reshape(-1)[int(np.argmax(bmax))] = True
    D.write_package(tmp_path / "one", vol, 8, 1, empty=one)
    r = Renderer.from_package(gpu_ctx, tmp_path / "one", cfg)
    img = r.offline_render(cfg)
    r.close()
    changed = np.count_nonzero(np.any(img != full, axis=-1))
    assert 0 < changed < np.count_nonzero(full.sum(-1))


@pytest.mark.parametrize("mode", ["Volume", "Solid"])
def test_blocks_from_package_equals_raw(gpu_ctx, mode):
    """Blocks shades with the package's mean thumbnail; the raw-volume renderer derives
    the same means and emptiness from the same voxels."""
    cfg = {"resolution": [80, 64], "renderer": "Blocks", "params": {"mode": mode, "density": 5e-2}}
    cam = {"orbit": {"arm": [25, 65, 2.2]}}
    a = Renderer.from_package(gpu_ctx, PKG, cfg)
    b = Renderer(gpu_ctx, _voxels(PKG), cfg, block_size=8, apron=1)
    ia, ib = a.offline_render(cam), b.offline_render(cam)
    assert np.array_equal(ia, ib)
    assert np.count_nonzero(ia.sum(-1)) > 50
    a.close()
    b.close()


def test_isosurface_from_package_equals_raw(gpu_ctx, tmp_path):
    """With a thumbnail that marks the bricks the iso value cannot cross, the package
    renderer's image equals the raw-volume renderer's."""
    vol = _voxels(PKG)
    iso = 0.4
    _, bmin, bmax, _ = D.brick_stats(vol, 8, 1)
    D.write_package(tmp_path, vol, 8, 1, empty=(bmin > iso) | (bmax < iso))
    cfg = {"resolution": [96, 96], "renderer": "Isosurface", "params": {"mode": "Normal", "isovalue": iso}}
    cam = {"orbit": {"arm": [25, 65, 2.2]}}
    a = Renderer.from_package(gpu_ctx, tmp_path, cfg)
    b = Renderer(gpu_ctx, vol, cfg, block_size=8, apron=1)
    ia, ib = a.offline_render(cam), b.offline_render(cam)
    print(f"package Isosurface: lit {int(np.count_nonzero(ia.sum(-1)))}, "
          f"differing bytes {int(np.count_nonzero(ia != ib))}", flush=True)
    assert np.array_equal(ia, ib)
    assert np.count_nonzero(ia.sum(-1)) > 50
    a.close()
    b.close()


def test_package_errors_need_no_render(gpu_ctx, tmp_path):
    with pytest.raises(L.HydError, match="cannot open"):
        Renderer.from_package(gpu_ctx, tmp_path / "missing", _cfg())
    with pytest.raises(L.HydError, match="unknown renderer|renderer"):
        Renderer.from_package(gpu_ctx, PKG, {"renderer": "Nope"})
