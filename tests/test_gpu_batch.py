"""GPU tests of the batched persistent renderer (hyd_batch_*).

Bar: per-view error against the FP64 oracle (lattice schedule, exact u8 source
values) <= 1e-3 max-abs for FP16 and u8 bricks; FP32 bricks agree with the FP32
oracle to 1e-4; graph replays equal the eager launch bitwise; the launch shape
(autotuner knobs) never changes the output.
"""
import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import workloads as W
from hydrant_amd.hip import batch_config
from tests.helpers import error_pixels, exact_pool_values, oracle_view, oracle_volume

pytestmark = pytest.mark.gpu


def _small(fmt, n_views=6, normalize=False):
    return W.batched(sizes=(64, 96, 128), n_views=n_views, res_choices=(64, 96, 128, 160), fmt=fmt, bs=32,
                     apron=1, seed=77, normalize=normalize)


def _setup(ctx, wl):
    ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, wl.pool)
    ctx.set_tf(wl.tf)
    ctx.set_volumes(wl.volume_descs())
    return ctx.batch(wl.hip_views())


def _view_errors(wl, out, batch, prec, exact, views=None):
    errs = []
    for i in (range(len(wl.views)) if views is None else views):
        sc = wl.scene(i)
        img = batch.view_image(out, i).cpu().numpy()
        xy = error_pixels(sc.W, sc.H)
        pv = exact_pool_values(sc) if exact else None
        osc = oracle_volume(sc, pool_values=pv, exact_mapping=exact)
        ref, _ = osc.render_pixels(oracle_view(sc), xy, prec=prec)
        got = img[xy[:, 1], xy[:, 0]]
        errs.append(float(np.max(np.abs(got.astype(np.float64) - ref))))
    return errs


@pytest.mark.parametrize("fmt", [L.FMT_F16, L.FMT_U8, L.FMT_F32])
def test_batch_vs_fp64_oracle(gpu_ctx, fmt):
    wl = _small(fmt)
    b = _setup(gpu_ctx, wl)
    out, cnt = b.render(batch_config(8, 1, 8), wl.density, wl.ert, count=True)
    gpu_ctx.synchronize()
    errs = _view_errors(wl, out, b, 64, exact=True)
    print(f"fmt {fmt}: per-view max-abs vs FP64 {['%.2e' % e for e in errs]}, samples {int(cnt.item())}", flush=True)
    assert max(errs) <= 1e-3
    if fmt == L.FMT_F32:
        e32 = _view_errors(wl, out, b, 32, exact=False)
        assert max(e32) <= 1e-4, e32
    b.close()


def test_non_square_and_ragged_views(gpu_ctx):
    """Views whose width and height differ and are not multiples of the tile edge (the
    reference scales a pixel's uv by the image height, so x spans more than [-1, 1] on
    wide views): every launch shape renders the same image, within 1e-3 of FP64."""
    import torch

    wl = _small(L.FMT_F16, n_views=4)
    for v, (w, h) in zip(wl.views, [(96, 64), (70, 130), (130, 70), (33, 47)]):
        v["W"], v["H"] = w, h
    b = _setup(gpu_ctx, wl)
    ref = None
    for cfg in (batch_config(8, 4, 8), batch_config(32, 2, 8), batch_config(16, 1, 12), batch_config(8, 4, 8, 4)):
        out, _ = b.render(cfg, wl.density, wl.ert)
        torch.cuda.synchronize()
        if ref is None:
            ref = out.clone()
            errs = _view_errors(wl, out, b, 64, exact=True)
            print(f"non-square views: per-view max-abs vs FP64 {['%.2e' % e for e in errs]}", flush=True)
            assert max(errs) <= 1e-3
        assert torch.equal(out, ref)
    b.close()


def test_batch_normalised_fp16_storage_error(gpu_ctx):
    """FP16 bricks holding float(v)/255 (instead of the exact count v) carry up to 2.4e-4
    rounding per voxel; the image error against FP64 is reported and bounded loosely."""
    wl = _small(L.FMT_F16, normalize=True)
    b = _setup(gpu_ctx, wl)
    out, _ = b.render(batch_config(8, 1, 8), wl.density, wl.ert)
    gpu_ctx.synchronize()
    errs = _view_errors(wl, out, b, 64, exact=True)
    print(f"normalised fp16: per-view max-abs vs FP64 {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-2
    b.close()


def test_launch_shape_does_not_change_output(gpu_ctx):
    import torch

    wl = _small(L.FMT_F16, n_views=4)
    b = _setup(gpu_ctx, wl)
    ref, c0 = b.render(batch_config(8, 1, 8), wl.density, wl.ert, count=True)
    for tw, un, wpc, g in [(16, 2, 4, 1), (32, 1, 16, 1), (32, 2, 2, 1), (8, 4, 32, 1), (16, 4, 8, 1),
                           (8, 4, 8, 4), (8, 2, 12, 4), (8, 1, 3, 4)]:
        out, c1 = b.render(batch_config(tw, un, wpc, g), wl.density, wl.ert, count=True)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (tw, un, wpc, g)
        assert int(c1.item()) == int(c0.item())
    best, ms = b.autotune(wl.density, wl.ert, b.new_output(), reps=1)
    assert best.tile_w in (8, 16, 32) and best.unroll in (1, 2, 4) and ms > 0
    b.close()


def test_graph_replay_matches_eager(gpu_ctx):
    """Capture the persistent launch in a HIP graph; every replay (with eager work in
    between) equals the eager result bitwise."""
    import torch

    wl = _small(L.FMT_F16, n_views=5)
    b = _setup(gpu_ctx, wl)
    cfg = batch_config(8, 4, 8, 4)
    ref, _ = b.render(cfg, wl.density, wl.ert)
    out = b.new_output()
    b.capture(cfg, wl.density, wl.ert, out)
    for rep in range(3):
        out.fill_(-1.0)
        junk = torch.randn(1024, device="cuda").sum()  # eager work between replays
        b.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
        del junk
    b.close()


def test_replay_refused_after_state_change(gpu_ctx):
    """A captured graph holds the device pointers of the pool, transfer function and
    volume set: after any of them is replaced, replay is refused until the batch is
    captured again, and the new capture replays the eager image."""
    import torch

    wl = _small(L.FMT_F16, n_views=3)
    b = _setup(gpu_ctx, wl)
    cfg = batch_config(8, 4, 8)
    out = b.new_output()
    b.capture(cfg, wl.density, wl.ert, out)
    b.replay()
    for change in (lambda: gpu_ctx.set_tf(wl.tf),
                   lambda: gpu_ctx.set_volumes(wl.volume_descs()),
                   lambda: gpu_ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, wl.pool)):
        change()
        with pytest.raises(L.HydError) as e:
            b.replay()
        assert e.value.code == L.HYD_ERR_STATE
        b.capture(cfg, wl.density, wl.ert, out)
        out.fill_(-1.0)
        b.replay()
        ref, _ = b.render(cfg, wl.density, wl.ert)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
    b.close()


def test_batch_refuses_views_beyond_the_volume_set(gpu_ctx):
    """A batch whose views name a volume that a later volume set no longer holds is
    refused before any launch."""
    wl = _small(L.FMT_F16, n_views=6)
    b = _setup(gpu_ctx, wl)
    assert max(v["volume"] for v in wl.views) > 0
    gpu_ctx.set_volumes(wl.volume_descs()[:1])
    with pytest.raises(L.HydError) as e:
        b.render(batch_config(8, 4, 8), wl.density, wl.ert)
    assert e.value.code == L.HYD_ERR_STATE
    b.close()


def test_batch_refuses_a_pool_smaller_than_the_volume_set(gpu_ctx):
    """A pool uploaded after the volume set with fewer pages than its cells name is
    refused before any launch."""
    wl = _small(L.FMT_F16, n_views=2)
    b = _setup(gpu_ctx, wl)
    gpu_ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, np.ascontiguousarray(wl.pool[:1]))
    with pytest.raises(L.HydError) as e:
        b.render(batch_config(8, 4, 8), wl.density, wl.ert)
    assert e.value.code == L.HYD_ERR_STATE
    b.close()


def test_batch_rejects_bad_config(gpu_ctx):
    import ctypes as C

    wl = _small(L.FMT_F16, n_views=2)
    b = _setup(gpu_ctx, wl)
    out = b.new_output()
    for bad in (batch_config(12, 1, 8), batch_config(16, 1, 8, 4), batch_config(8, 1, 8, 2)):
        code = L.lib().hyd_batch_render(gpu_ctx.h, b.h, C.byref(bad), 0.01, 0.9995, C.c_void_p(out.data_ptr()),
                                        None)
        assert code == L.HYD_ERR_ARG
    b.close()


def test_pool_change_refreshes_oriented_layouts():
    """The batched renderer reads the pool through y- and z-fastest layouts derived at
    its first launch after an upload or bind: a new pool (uploaded, then bound from
    device memory) must render exactly as a fresh context does."""
    import dataclasses

    import torch
    from hydrant_amd.hip import Context

    wl = _small(L.FMT_F16, n_views=4)
    pad = wl.bs + 2 * wl.apron
    ctx = Context(0)
    b = _setup(ctx, wl)
    cfg = batch_config(8, 4, 8)
    first = b.render(cfg, wl.density, wl.ert)[0].clone()
    # the same pages with the voxel order reversed within each page
    bricks = np.ascontiguousarray(wl.pool).reshape(-1, pad ** 3)
    other = np.ascontiguousarray(bricks[:, ::-1]).reshape(wl.pool.shape)
    ctx.upload_pool(wl.fmt, pad, other)
    second = b.render(cfg, wl.density, wl.ert)[0].clone()
    ctx2 = Context(0)
    b2 = _setup(ctx2, dataclasses.replace(wl, pool=other))
    ref = b2.render(cfg, wl.density, wl.ert)[0]
    torch.cuda.synchronize()
    assert not torch.equal(second, first)
    assert torch.equal(second, ref)
    # the original pool again, now bound from device memory
    dev = torch.from_numpy(np.ascontiguousarray(wl.pool).view(np.uint8).reshape(-1).copy()).cuda()
    ctx.bind_pool_device(wl.fmt, pad, bricks.shape[0], dev.data_ptr())
    third = b.render(cfg, wl.density, wl.ert)[0]
    torch.cuda.synchronize()
    assert torch.equal(third, first)
    b.close()
    b2.close()
    ctx2.close()
    ctx.close()


@pytest.mark.parametrize("fmt", [L.FMT_F16, L.FMT_U8])
@pytest.mark.parametrize("switch", ["HYD_BATCH_NO_QUADS", "HYD_BATCH_NO_PAIRS", "HYD_BATCH_SINGLE_LAYOUT"])
def test_layout_fallbacks_render_the_same(monkeypatch, switch, fmt):
    """The quad layouts (the FP16 default), the row-pair layouts (the u8 default, and FP16's
    with HYD_BATCH_NO_QUADS=1 or when the device cannot hold the quads), the y-/z-fastest
    layouts (HYD_BATCH_NO_PAIRS=1, or when the device cannot hold the pairs) and the
    x-fastest layout alone (HYD_BATCH_SINGLE_LAYOUT=1, when it cannot hold either) hold the
    same voxels: same image, bitwise."""
    import torch
    from hydrant_amd.hip import Context

    wl = _small(fmt, n_views=4)
    cfg = batch_config(8, 4, 8)
    ctx = Context(0)
    b = _setup(ctx, wl)
    ref = b.render(cfg, wl.density, wl.ert)[0].clone()
    monkeypatch.setenv(switch, "1")
    ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, wl.pool)   # the switch is read when the pool changes
    out = b.render(cfg, wl.density, wl.ert)[0]
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    b.close()
    ctx.close()


def test_batch_rejects_pools_beyond_32_bit_offsets():
    """The batched kernel addresses voxels by 32-bit byte offsets: a pool of more than
    4 GiB is refused by the host before any kernel runs (the pool is only declared that
    large; no launch reads it)."""
    import torch
    from hydrant_amd.hip import Context

    wl = _small(L.FMT_F16, n_views=2)
    pad = wl.bs + 2 * wl.apron
    ctx = Context(0)
    b = _setup(ctx, wl)
    dev = torch.from_numpy(np.ascontiguousarray(wl.pool).view(np.uint8).reshape(-1).copy()).cuda()
    pages = (1 << 32) // (pad ** 3 * 2) + 1
    ctx.bind_pool_device(wl.fmt, pad, pages, dev.data_ptr())
    with pytest.raises(L.HydError) as e:
        b.render(batch_config(8, 4, 8), wl.density, wl.ert)
    assert e.value.code == L.HYD_ERR_RANGE
    b.close()
    ctx.close()


@pytest.mark.parametrize("fmt", [L.FMT_F16, L.FMT_U8, L.FMT_F32])
def test_batch_64_bricks_vs_fp64_oracle(monkeypatch, fmt):
    """64^3 bricks with a one-voxel apron take the kernels compiled for a padded edge of
    66 (the 96-view benchmark's bricks), through each pool layout: quads (the FP16
    default), row pairs, the single-voxel oriented layouts and the x-fastest pool alone.  Error
    against the FP64 oracle within 1e-3, and the layouts give the same image."""
    import torch
    from hydrant_amd.hip import Context

    wl = W.batched(sizes=(128, 192), n_views=4, res_choices=(96, 128), fmt=fmt, bs=64, apron=1, seed=91)
    ctx = Context(0)
    b = _setup(ctx, wl)
    ref = None
    for switch in (None, "HYD_BATCH_NO_QUADS", "HYD_BATCH_NO_PAIRS", "HYD_BATCH_SINGLE_LAYOUT"):
        if switch is not None:
            monkeypatch.setenv(switch, "1")
            ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, wl.pool)   # the switch is read when the pool changes
        out, cnt = b.render(batch_config(8, 4, 16), wl.density, wl.ert, count=True)
        torch.cuda.synchronize()
        errs = _view_errors(wl, out, b, 64, exact=True)
        print(f"pad 66 fmt {fmt} ({switch}): per-view max-abs vs FP64 {['%.2e' % e for e in errs]}, "
              f"samples {int(cnt.item())}", flush=True)
        assert max(errs) <= 1e-3
        if ref is None:
            ref = out.clone()
        assert torch.equal(out, ref)
    b.close()
    ctx.close()


@pytest.mark.parametrize("fmt", [L.FMT_F16, L.FMT_U8])
def test_large_bricks_unstaged_orientation(monkeypatch, fmt):
    """128^3 bricks with a one-voxel apron (padded edge 130): for FP16 the two source
    planes a workgroup of the orientation passes needs exceed 64 KB of LDS, so the quad,
    row-pair and single-voxel layouts are written without LDS staging (u8 planes still
    fit).  Every layout renders the x-fastest pool's image bit for bit, within 1e-3 of
    the FP64 oracle."""
    import torch
    from hydrant_amd.hip import Context

    wl = W.batched(sizes=(256,), n_views=3, res_choices=(64, 96), fmt=fmt, bs=128, apron=1, seed=93)
    if fmt == L.FMT_F16:
        assert 2 * (wl.bs + 2 * wl.apron) ** 2 * 2 > 65536   # two FP16 planes exceed the staging limit
    ctx = Context(0)
    b = _setup(ctx, wl)
    ref = None
    for switch in ("HYD_BATCH_SINGLE_LAYOUT", "HYD_BATCH_NO_PAIRS", "HYD_BATCH_NO_QUADS", None):
        for s in ("HYD_BATCH_SINGLE_LAYOUT", "HYD_BATCH_NO_PAIRS", "HYD_BATCH_NO_QUADS"):
            monkeypatch.delenv(s, raising=False)
        if switch is not None:
            monkeypatch.setenv(switch, "1")
        ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, wl.pool)   # the switches are read when the pool changes
        out = b.render(batch_config(8, 4, 8), wl.density, wl.ert)[0]
        torch.cuda.synchronize()
        if ref is None:
            ref = out.clone()
            errs = _view_errors(wl, out, b, 64, exact=True)
            print(f"pad 130 fmt {fmt}: per-view max-abs vs FP64 {['%.2e' % e for e in errs]}", flush=True)
            assert max(errs) <= 1e-3
        else:
            assert torch.equal(out, ref), f"layout switch {switch} changed the image"
    b.close()
    ctx.close()


@pytest.mark.parametrize("fmt", [L.FMT_F16, L.FMT_U8])
@pytest.mark.parametrize("switch", [None, "HYD_BATCH_NO_QUADS", "HYD_BATCH_NO_PAIRS", "HYD_BATCH_SINGLE_LAYOUT"])
def test_pool_page_writes_refresh_oriented_layouts(monkeypatch, switch, fmt):
    """Pages overwritten in place (hyd_write_pool_pages) after the oriented copies were
    made render as a fresh context with the modified pool does: eagerly, and from a graph
    captured before the write (the copies are updated in place)."""
    import dataclasses

    import torch
    from hydrant_amd.hip import Context

    if switch is not None:
        monkeypatch.setenv(switch, "1")
    wl = _small(fmt, n_views=4)
    pad = wl.bs + 2 * wl.apron
    cfg = batch_config(8, 4, 8)
    ctx = Context(0)
    b = _setup(ctx, wl)
    out = b.new_output()
    b.capture(cfg, wl.density, wl.ert, out)
    first = b.render(cfg, wl.density, wl.ert)[0].clone()
    new_pool = np.ascontiguousarray(wl.pool).copy()
    flat = new_pool.reshape(new_pool.shape[0], -1)
    for p in (1, 3, 4, 5):
        flat[p] = flat[p][::-1].copy()
    ctx.write_pool_pages(1, new_pool[1:2])
    ctx.write_pool_pages(3, new_pool[3:6])
    eager = b.render(cfg, wl.density, wl.ert)[0].clone()
    out.fill_(-1.0)
    b.replay()
    ctx2 = Context(0)
    b2 = _setup(ctx2, dataclasses.replace(wl, pool=new_pool))
    ref = b2.render(cfg, wl.density, wl.ert)[0]
    torch.cuda.synchronize()
    assert not torch.equal(eager, first)
    assert torch.equal(eager, ref)
    assert torch.equal(out, ref)
    b.close()
    b2.close()
    ctx.close()
    ctx2.close()


def test_autotune_schedule_within_budget(gpu_ctx):
    """The step-schedule autotuner times only termination thresholds whose truncation
    (1 - ert) fits half the error budget, and the chosen schedule's image is within the
    budget of the FP64 oracle run without termination."""
    wl = _small(L.FMT_F16, n_views=4)
    b = _setup(gpu_ctx, wl)
    out = b.new_output()
    for budget in (1e-3, 2e-2):
        cfg, ert, ms = b.autotune_schedule(wl.density, budget, out, reps=1)
        assert ert == 0.0 or 1.0 - ert <= budget / 2 + 1e-7, (budget, ert)
        assert ms > 0.0
        out, _ = b.render(cfg, wl.density, ert, out=out)
        gpu_ctx.synchronize()
        wl.ert = 0.0
        errs = _view_errors(wl, out, b, 64, exact=True)
        print(f"budget {budget}: ert {ert}, config {cfg.as_dict()}, max-abs vs FP64 (no termination) "
              f"{max(errs):.3g}", flush=True)
        assert max(errs) <= budget
    b.close()
