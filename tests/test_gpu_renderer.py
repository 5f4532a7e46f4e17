"""GPU tests of the JSON renderer interface (include/hydrant_renderer.h): the
front end's RGB8 images against the oracle (Volume, bit-exact) and against the
same kernels driven directly (Blocks, Isosurface), residency batches, update()."""
import copy

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import dataset as D
from hydrant_amd import workloads as W
from hydrant_amd.renderer import Renderer
from oracle import oracle as O
from tests.helpers import oracle_view, oracle_volume
from tests.test_renderer_config import README_CONFIG

pytestmark = pytest.mark.gpu


def _volume_case(n=64, res=96, bs=16, sample_rate=0.5):
    vol = D.gen_blobs(n, 16, W.SEED_BLOBS, sig=(8.0, 40.0), sigma_scale=n / 256.0, amp=(0.3, 1.0))
    cfg = copy.deepcopy(README_CONFIG)
    cfg["render"]["resolution"] = [res, res]
    cfg["render"]["params"]["sample_rate"] = sample_rate
    tf = np.array(cfg["render"]["params"]["transfer_fn"]["values"], np.float32)
    b = D.make_bricked(vol, bs, 1, L.FMT_U8, D.volume_empty_fn(tf), resident=1.0, seed=1)
    sc = W.Scene("json_volume", b, tf, res, res, (25.0, 65.0, 2.2), 60.0, D.geometry((n, n, n), bs, sample_rate),
                 density=1e-2, extra={"volume": vol})
    return vol, cfg, sc


def test_volume_readme_config_matches_oracle(gpu_ctx):
    vol, cfg, sc = _volume_case()
    r = Renderer(gpu_ctx, vol, cfg, block_size=16, apron=1)
    img = r.offline_render(cfg)
    o = oracle_volume(sc).render(oracle_view(sc), 32, film=False)
    ref = O.pixel_pass(o["rgba"])
    diff = int(np.count_nonzero(img != ref))
    print(f"Volume via JSON: {diff} differing bytes, lit pixels {int(np.count_nonzero(img.sum(-1)))}", flush=True)
    assert np.array_equal(img, ref)
    assert np.count_nonzero(img.sum(-1)) > 100
    r.close()


def test_volume_schema_config_equals_readme_form(gpu_ctx):
    """Schema spelling (mode "Default", flat TF values) renders the same image."""
    vol, cfg, _ = _volume_case(n=48, res=64)
    r1 = Renderer(gpu_ctx, vol, cfg, block_size=16, apron=1)
    a = r1.offline_render(cfg["camera"])
    flat = copy.deepcopy(cfg)
    p = flat["render"]["params"]
    p["mode"] = "Default"
    p["transfer_fn"]["values"] = [x for q in p["transfer_fn"]["values"] for x in q]
    r2 = Renderer(gpu_ctx, vol, flat["render"], block_size=16, apron=1)
    b = r2.offline_render(flat)
    assert np.array_equal(a, b)
    r1.close()
    r2.close()


def _nearest_first(b, view_o):
    """The front end's residency order: non-empty bricks by float32 squared distance of
    the brick centre to the camera (object space), ties by brick index."""
    gz, gy, gx = b.empty.shape
    ids, d2 = [], []
    for i in range(b.empty.size):
        if b.empty.reshape(-1)[i]:
            continue
        x, y, z = i % gx, (i // gx) % gy, i // (gx * gy)
        c = [np.float32(x) + np.float32(0.5) - np.float32(view_o[0]),
             np.float32(y) + np.float32(0.5) - np.float32(view_o[1]),
             np.float32(z) + np.float32(0.5) - np.float32(view_o[2])]
        ids.append(i)
        d2.append(np.float32(np.float32(c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]))
    order = sorted(range(len(ids)), key=lambda k: (d2[k], k))
    return [ids[k] for k in order]


def test_volume_residency_batches_match_oracle_resume(gpu_ctx):
    """mem_limit_mb below the dataset: bricks become resident nearest first in batches
    (emit on the first, resume on the next, volume.cc:95-134); the oracle replays the
    same batches."""
    vol, cfg, sc = _volume_case(n=64, res=64, bs=16)
    b = sc.bricked
    pad = 16 + 2
    per = 20
    cfg["render"]["params"]["mem_limit_mb"] = (per + 0.5) * pad ** 3 / 1048576.0
    r = Renderer(gpu_ctx, vol, cfg, block_size=16, apron=1)
    img = r.offline_render(cfg)
    order = _nearest_first(b, list(sc.view().ray_o))
    assert len(order) > 2 * per, "the case needs at least three batches"
    view = oracle_view(sc)
    film = None
    for k in range(0, len(order), per):
        vaddr = np.full(b.vaddr.shape, -1, np.int32)
        sel = np.array(order[k:k + per])
        vaddr.reshape(-1)[sel] = b.vaddr.reshape(-1)[sel]
        osc = oracle_volume(sc, vaddr=vaddr)
        out = osc.render(view, 32, film=True) if film is None else osc.resume(film, 32)
        film = out["film"]
    ref = O.pixel_pass(out["rgba"])
    print(f"batches {(len(order) + per - 1) // per}: differing bytes {int(np.count_nonzero(img != ref))}", flush=True)
    assert np.array_equal(img, ref)
    r.close()


@pytest.mark.parametrize("mode", ["Volume", "Solid"])
def test_blocks_json_matches_kernel(gpu_ctx, mode):
    sc = W.blocks(n=96, W=80, H=64, bs=16)
    cfg = {"resolution": [80, 64], "renderer": "Blocks", "params": {"mode": mode, "density": 1e-2}}
    r = Renderer(gpu_ctx, _blob(96), cfg, block_size=16, apron=0)   # the volume W.blocks builds
    img = r.offline_render({"orbit": {"arm": list(sc.arm)}})
    b = sc.bricked
    gpu_ctx.set_empty_grid(b.cheb)
    gpu_ctx.set_mean_grid(b.mean)
    m = L.BLOCKS_VOLUME if mode == "Volume" else L.BLOCKS_SOLID
    rgba = gpu_ctx.render_blocks(W.blocks_params(sc, m), sc.view(), sc.W, sc.H)
    ref = gpu_ctx.pixel_pass(rgba).cpu().numpy()
    o = O.pixel_pass(O.blocks_render(oracle_view(sc), sc.W, sc.H, sc.bbox_min, sc.bbox_max, sc.geom.step, b.cheb,
                                     b.mean, mode=m, density=sc.density).astype(np.float32))
    assert np.array_equal(img, ref)
    assert int(np.max(np.abs(img.astype(int) - o.astype(int)))) <= 1
    assert np.count_nonzero(img.sum(-1)) > 50
    r.close()


def test_paging_renderer_offline_is_the_clear_colour(gpu_ctx):
    """The reference's Paging renderer renders nothing offline (paging.cc:51-56)."""
    cfg = {"resolution": [40, 24], "renderer": "Paging",
           "params": {"clear_color": [0.2, 0.4, 1.0], "mem_limit_mb": 512}}
    r = Renderer(gpu_ctx, _blob(48), cfg, block_size=16, apron=1)
    img = r.offline_render({"orbit": {"arm": [25, 65, 2.2]}})
    assert img.shape == (24, 40, 3)
    assert np.all(img == np.array([51, 102, 255], np.uint8))
    with pytest.raises(L.HydError, match="mem_limit_mb"):
        r.update({"mem_limit_mb": 0})
    r.close()


def _blob(n):
    return D.gen_blobs(n, 16, W.SEED_BLOBS, sig=(8.0, 40.0), sigma_scale=n / 256.0, amp=(0.3, 1.0))


@pytest.mark.parametrize("mode", ["Color", "Normal"])
def test_isosurface_json_matches_kernel(gpu_ctx, mode):
    sc = W.isosurface(n=128, W=96, H=96, bs=32, corrected=1)
    sdf = sc.extra["volume"]
    if sdf.dtype == np.uint16:   # half bits
        sdf = sdf.view(np.float16)
    cfg = {"resolution": [96, 96], "renderer": "Isosurface",
           "params": {"mode": mode, "isovalue": 0.0, "corrected": 1, "bisect_steps": 8}}
    r = Renderer(gpu_ctx, sdf, cfg, block_size=32, apron=1)
    img = r.offline_render({"orbit": {"arm": list(sc.arm)}})
    gpu_ctx.load(sc.bricked)
    p = W.iso_params(sc, L.ISO_COLOR if mode == "Color" else L.ISO_NORMAL)
    px = gpu_ctx.render_iso(p, sc.view(), sc.W, sc.H)["px"].cpu().numpy()
    ref = O.pixel_pass(np.ascontiguousarray(px[..., :4]))
    print(f"iso {mode}: hits {int((px[..., 5].view(np.int32) == 1).sum())}, differing bytes "
          f"{int(np.count_nonzero(img != ref))}", flush=True)
    assert np.array_equal(img, ref)
    assert np.count_nonzero(img.sum(-1)) > 100
    r.close()


def test_update_and_errors(gpu_ctx):
    vol, cfg, _ = _volume_case(n=48, res=48)
    r = Renderer(gpu_ctx, vol, cfg, block_size=16, apron=1)
    a = r.offline_render(cfg)
    r.update({"transfer_fn": {"values": [0, 0, 0, 0, 0, 1, 0, 1]}, "density": 0.05})
    b = r.offline_render(cfg)
    assert not np.array_equal(a, b)
    assert np.all(b[..., 0] == 0) and np.any(b[..., 1] > 0)   # green ramp only
    with pytest.raises(L.HydError, match="Cpu"):
        r.update({"device": "Cpu"})
    with pytest.raises(L.HydError, match="unknown value"):
        r.update({"mode": "Fancy"})
    with pytest.raises(L.HydError, match="orbit"):
        r.offline_render({"perspective": {"fovy": 30}})
    r.close()
