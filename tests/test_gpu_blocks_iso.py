"""GPU parity of the Blocks and Isosurface kernels against the oracle."""
import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import workloads as W
from oracle import oracle as O
from tests.helpers import max_abs, oracle_view

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [L.BLOCKS_VOLUME, L.BLOCKS_SOLID])
def test_blocks(gpu_ctx, mode):
    sc = W.blocks(n=128, W=96, H=96, bs=16)
    b = sc.bricked
    gpu_ctx.set_empty_grid(b.cheb)
    gpu_ctx.set_mean_grid(b.mean)
    p = W.blocks_params(sc, mode)
    g = gpu_ctx.render_blocks(p, sc.view(), sc.W, sc.H).cpu().numpy()
    o = O.blocks_render(oracle_view(sc), sc.W, sc.H, sc.bbox_min, sc.bbox_max, sc.geom.step, b.cheb, b.mean,
                        mode=mode, density=sc.density)
    err = max_abs(g, o.astype(np.float32))
    print(f"blocks mode {mode}: max-abs {err:.3g}, empty bricks {int(b.empty.sum())}/{b.empty.size}", flush=True)
    assert err <= 1e-4
    assert np.any(g[..., 3] > 0)


def test_paging_view(gpu_ctx):
    """The Paging renderer's residency view over a seeded page table with faults
    and two page levels, against the oracle."""
    sc = W.blocks(n=128, W=96, H=96, bs=16)
    b = sc.bricked
    rng = np.random.default_rng(11)
    vaddr = rng.integers(0, 64, size=b.cheb.shape).astype(np.int32)
    vaddr[rng.random(b.cheb.shape) < 0.05] = -1
    lowest = 32
    gpu_ctx.set_empty_grid(b.cheb)
    gpu_ctx.set_page_table(vaddr, np.ones(64, np.float32), np.zeros((64, 3), np.float32), lowest)
    p = W.blocks_params(sc, L.BLOCKS_PAGING)
    g = gpu_ctx.render_blocks(p, sc.view(), sc.W, sc.H).cpu().numpy()
    o = O.blocks_render(oracle_view(sc), sc.W, sc.H, sc.bbox_min, sc.bbox_max, sc.geom.step, b.cheb, None,
                        mode=2, vaddr=vaddr, lowest=lowest)
    err = max_abs(g, o.astype(np.float32))
    print(f"paging view: max-abs {err:.3g}, faults {int((vaddr < 0).sum())}", flush=True)
    assert err <= 1e-6
    assert np.any(g[..., 1] > 0) and np.any(g[..., 0] > 0)   # both page levels seen


def _iso_oracle(sc, p):
    b = sc.bricked
    return O.IsoScene(sc.W, sc.H, sc.bbox_min, sc.bbox_max, p.step, p.du, np.array(p.to_world[:]),
                      p.eye_pos[:], p.light_pos[:], p.surface_color[:], p.isovalue, b.cheb, b.vaddr, b.page_k,
                      b.page_b, b.pool_values(), b.pad, mode=p.mode, corrected=p.corrected,
                      bisect_steps=p.bisect_steps, max_steps=p.max_steps)


@pytest.mark.parametrize("corrected", [0, 1])
@pytest.mark.parametrize("mode", [L.ISO_COLOR, L.ISO_NORMAL, L.ISO_POSITION])
def test_isosurface(gpu_ctx, corrected, mode):
    sc = W.isosurface(n=128, W=96, H=96, bs=32, corrected=corrected)
    b = sc.bricked
    gpu_ctx.load(b)
    p = W.iso_params(sc, mode)
    out = gpu_ctx.render_iso(p, sc.view(), sc.W, sc.H, count=True)
    gpu_ctx.synchronize()
    px = out["px"].cpu().numpy()
    g_hit = px[..., 5].view(np.int32)
    g_depth = px[..., 4]
    r = _iso_oracle(sc, p).render(oracle_view(sc), 32)
    o = r["px"]
    o_hit = o[..., 5].astype(np.int32)
    assert np.array_equal(g_hit, o_hit), int((g_hit != o_hit).sum())
    hit = g_hit == 1
    depth_vox = np.max(np.abs(g_depth[hit] - o[..., 4][hit].astype(np.float32))) * b.bs if hit.any() else 0.0
    err = max_abs(px[..., :4], o[..., :4].astype(np.float32))
    print(f"iso corrected={corrected} mode={mode}: hits {int(hit.sum())}, depth err {depth_vox:.3g} voxel, "
          f"colour max-abs {err:.3g}, samples {int(out['samples'].item())}/{r['samples']}", flush=True)
    assert depth_vox <= 1e-3
    assert err <= 1e-4
    assert int(out["samples"].item()) == r["samples"]
    if corrected:
        assert hit.sum() > 0


def test_isosurface_refuses_smaller_pool(gpu_ctx):
    """A pool uploaded after the page table with fewer pages than the table names is refused
    on the host before any launch (the march would read past the pool's end)."""
    sc = W.isosurface(n=64, W=16, H=16, bs=32)
    b = sc.bricked
    gpu_ctx.load(b)
    p = W.iso_params(sc, L.ISO_COLOR)
    assert b.n_pages > 1
    gpu_ctx.upload_pool(b.fmt, b.pad, b.pool[:1])
    with pytest.raises(L.HydError) as e:
        gpu_ctx.render_iso(p, sc.view(), sc.W, sc.H)
    assert "pool holds 1" in str(e.value)


def test_isosurface_fetch_and_depth_composite(gpu_ctx):
    """Isosurface sort-last: three partial renders (each with a third of the bricks
    resident; rays stop at a fault), the fetch pass (isosurface_shader.cu:42-49) and the
    depth-min composite (isosurface.cc:272-289) on the GPU, against the oracle's restatements
    of both on the GPU partials (bitwise) and on the oracle's own partials."""
    import torch

    sc = W.isosurface(n=128, W=96, H=80, bs=32, corrected=1)
    b = sc.bricked
    gpu_ctx.load(b)
    p = W.iso_params(sc, L.ISO_COLOR)
    rng = np.random.default_rng(12)
    part = rng.integers(0, 3, size=b.vaddr.shape)
    g_fetch, o_val, o_depth = [], [], []
    for i in range(3):
        va = np.where(part == i, b.vaddr, -1).astype(np.int32)
        gpu_ctx.set_page_table(va, b.page_k, b.page_b)
        px = gpu_ctx.render_iso(p, sc.view(), sc.W, sc.H)["px"]
        g_fetch.append(gpu_ctx.iso_fetch_pass(px))
        osc = O.IsoScene(sc.W, sc.H, sc.bbox_min, sc.bbox_max, p.step, p.du, np.array(p.to_world[:]),
                         p.eye_pos[:], p.light_pos[:], p.surface_color[:], p.isovalue, b.cheb, va, b.page_k,
                         b.page_b, b.pool_values(), b.pad, mode=p.mode, corrected=p.corrected,
                         bisect_steps=p.bisect_steps, max_steps=p.max_steps)
        v, d = O.iso_fetch(osc.render(oracle_view(sc), 32)["px"])
        o_val.append(v)
        o_depth.append(d)
    parts = torch.stack(g_fetch).contiguous()
    z = [1, 2, 0]
    rgb, depth = gpu_ctx.composite_iso(parts, z)
    pn = parts.cpu().numpy()
    g_val, g_depth = pn[..., 0:3], pn[..., 4:8].copy().view(np.float32)[..., 0]
    want_rgb, want_depth = O.composite_iso(g_val, g_depth, z)
    assert np.array_equal(rgb.cpu().numpy(), want_rgb)
    assert np.array_equal(depth.cpu().numpy(), want_depth)
    o_rgb, o_d = O.composite_iso(np.stack(o_val), np.stack(o_depth), z)
    hit = np.isfinite(o_d)
    assert np.array_equal(np.isfinite(depth.cpu().numpy()), hit) and hit.sum() > 200
    assert np.max(np.abs(depth.cpu().numpy()[hit] - o_d[hit])) * b.bs <= 1e-3
    assert np.max(np.abs(rgb.cpu().numpy().astype(int) - o_rgb.astype(int))) <= 1
    print(f"iso composite: {int(hit.sum())} hit pixels, bytes equal to the oracle's on "
          f"{float(np.mean(rgb.cpu().numpy() == o_rgb)):.4f} of channels", flush=True)
