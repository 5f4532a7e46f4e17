"""Rendering through a pool and page table built by the residency manager: bricks are
committed nearest first into pool slots (hyd_write_pool_pages) and the page table comes
from hyd_residency_vaddr.  With every brick resident the image equals the fully loaded
render bit for bit; with a pool smaller than the view needs, the GPU render of that
partial residency matches the oracle's (the 1e-4 volume bar, equal sample counts)."""
import dataclasses

import numpy as np
import pytest

from hydrant_amd import workloads as W
from hydrant_amd.residency import Residency
from tests.helpers import max_abs, oracle_view, oracle_volume

pytestmark = pytest.mark.gpu


def _scene(arm):
    sc = W.paged_volume(n=256, W=128, H=128, bs=32, resident=1.0)
    return dataclasses.replace(sc, arm=arm)


def _resident(ctx, sc, max_pages):
    """Commit the view's missing bricks into a fresh pool of max_pages slots."""
    b = sc.bricked
    ctx.upload_pool(b.fmt, b.pad, np.zeros((max_pages,) + b.pool.shape[1:], b.pool.dtype))
    res = Residency(b.cheb.shape, max_pages)
    missing = res.update(sc.view(), b.cheb)
    pool = np.zeros((max_pages,) + b.pool.shape[1:], b.pool.dtype)
    for i in missing:
        s = res.commit(int(i))
        assert s >= 0
        page = b.pool[b.vaddr.ravel()[i]][None]
        ctx.write_pool_pages(s, page)
        pool[s] = page[0]
    vaddr = res.vaddr()
    page_k = np.full(max_pages, b.page_k[0], np.float32)
    page_b = np.tile(b.page_b[0], (max_pages, 1)).astype(np.float32)
    ctx.set_page_table(vaddr, page_k, page_b)
    ctx.set_empty_grid(b.cheb)
    ctx.set_mean_grid(b.mean)
    ctx.set_tf(sc.tf)
    res.close()
    return dataclasses.replace(b, pool=pool, vaddr=vaddr, page_k=page_k, page_b=page_b), len(missing)


def test_full_residency_renders_the_loaded_image(gpu_ctx):
    sc = _scene((25.0, 65.0, 3.0))
    b = sc.bricked
    gpu_ctx.load(b, sc.tf)
    ref = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=False, count=True)
    ref_rgba, ref_n = ref["rgba"].cpu().numpy(), int(ref["samples"].item())
    rb, n_missing = _resident(gpu_ctx, sc, int(np.sum(b.cheb == 0)))
    assert n_missing == int(np.sum(b.cheb == 0))   # the whole volume is in view
    out = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=False, count=True)
    gpu_ctx.synchronize()
    assert int(out["samples"].item()) == ref_n
    assert np.array_equal(out["rgba"].cpu().numpy(), ref_rgba)


def test_partial_residency_matches_the_oracle(gpu_ctx):
    sc = _scene((25.0, 65.0, 2.2))
    rb, n_missing = _resident(gpu_ctx, sc, 300)
    assert n_missing == 300
    out = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=False, count=True)
    gpu_ctx.synchronize()
    osc = oracle_volume(dataclasses.replace(sc, bricked=rb))
    r = osc.render(oracle_view(sc), 32, film=False)
    err = max_abs(out["rgba"].cpu().numpy(), r["rgba"].astype(np.float32))
    print(f"partial residency (300 of {int(np.sum(sc.bricked.cheb == 0))} bricks): max-abs {err:.3g}, "
          f"samples gpu {int(out['samples'].item())} oracle {r['samples']}", flush=True)
    assert err <= 1e-4
    assert int(out["samples"].item()) == r["samples"]


def test_lowest_level_fallback_matches_the_oracle(gpu_ctx):
    """Non-resident bricks fall back to their lowest-level page (a coarse level-1 block
    with the reference's per-brick mapping) instead of stopping the ray: the GPU render of
    that pool and page table matches the oracle's, and composites more samples than the
    same residency without the fallback."""
    from hydrant_amd import dataset as D

    sc = _scene((25.0, 65.0, 2.2))
    b = sc.bricked
    low = D.lowest_level(sc.extra["volume"], b.bs, b.apron, b.fmt, level=1)
    n0, mp = low.pages.shape[0], 200
    pool = np.concatenate([low.pages, np.zeros((mp,) + b.pool.shape[1:], b.pool.dtype)])
    gpu_ctx.upload_pool(b.fmt, b.pad, pool)
    basic = np.arange(n0, dtype=np.int32).reshape(b.cheb.shape)   # brick i falls back to page i
    res = Residency(b.cheb.shape, mp, first_slot=n0, basic_vaddr=basic)
    for i in res.update(sc.view(), b.cheb):
        s = res.commit(int(i))
        assert s >= n0
        pool[s] = b.pool[b.vaddr.ravel()[i]]
        gpu_ctx.write_pool_pages(s, pool[s:s + 1])
    vaddr = res.vaddr()
    res.close()
    page_k = np.concatenate([low.page_k, np.full(mp, b.page_k[0], np.float32)])
    page_b = np.concatenate([low.page_b, np.tile(b.page_b[0], (mp, 1)).astype(np.float32)])
    gpu_ctx.set_page_table(vaddr, page_k, page_b)
    gpu_ctx.set_empty_grid(b.cheb)
    gpu_ctx.set_mean_grid(b.mean)
    gpu_ctx.set_tf(sc.tf)
    out = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=False, count=True)
    gpu_ctx.synchronize()
    rb = dataclasses.replace(b, pool=pool, vaddr=vaddr, page_k=page_k, page_b=page_b)
    r = oracle_volume(dataclasses.replace(sc, bricked=rb)).render(oracle_view(sc), 32, film=False)
    no_lod = np.where(vaddr >= n0, vaddr, -1).astype(np.int32)
    r0 = oracle_volume(dataclasses.replace(sc, bricked=dataclasses.replace(rb, vaddr=no_lod))).render(
        oracle_view(sc), 32, film=False)
    err = max_abs(out["rgba"].cpu().numpy(), r["rgba"].astype(np.float32))
    n = int(out["samples"].item())
    print(f"lowest-level fallback: max-abs {err:.3g}, samples gpu {n} oracle {r['samples']} "
          f"(without fallback {r0['samples']})", flush=True)
    assert err <= 1e-4
    assert n == r["samples"]
    assert n > r0["samples"]
