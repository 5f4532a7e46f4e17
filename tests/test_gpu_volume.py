"""GPU parity of the volume march (libhydrant_hip via the C ABI) against the oracle.

Bar: page/brick indexing bit-exact (per-sample traces equal), identical
composited-sample counts, and images within 1e-4 max-abs per channel in FP32;
the reference schedule reproduces the reference CPU backend's binary32
operation order, so its films and images must be bit-equal to the oracle's.
"""
import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import workloads as W
from tests.helpers import film_fields, max_abs, oracle_film_fields, oracle_view, oracle_volume

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _compare(ctx, sc, trace_pixels=(), check_film=True):
    ctx.load(sc.bricked, sc.tf)
    view = sc.view()
    p = sc.volume_params()
    out = ctx.render_volume(p, view, sc.W, sc.H, film=True, count=True)
    ctx.synchronize()
    g = film_fields(out["film"].cpu().numpy())
    g_rgba = out["rgba"].cpu().numpy()
    g_samples = int(out["samples"].item())
    osc = oracle_volume(sc)
    ov = oracle_view(sc)
    r = osc.render(ov, 32, film=True)
    o = oracle_film_fields(r["film"])
    err = max_abs(g_rgba, r["rgba"].astype(np.float32))
    print(f"{sc.name} {sc.W}x{sc.H}: max-abs {err:.3g}, samples gpu {g_samples} oracle {r['samples']}", flush=True)
    exact = sc.schedule == L.SCHED_REFERENCE
    assert err <= (0.0 if exact else TOL)
    assert g_samples == r["samples"]
    # the image-only kernel (no film, no theta/phi) renders the film frame's image
    ro = ctx.render_volume(p, view, sc.W, sc.H, film=False, count=True)
    ctx.synchronize()
    assert np.array_equal(ro["rgba"].cpu().numpy(), g_rgba)
    assert int(ro["samples"].item()) == g_samples
    if check_film:
        for k in ("v", "theta", "phi", "o"):
            assert max_abs(g[k], o[k]) <= (0.0 if exact else TOL), k
        if sc.schedule == L.SCHED_LATTICE:   # nsteps slot holds the float ray extent
            assert np.array_equal(g["nsteps"].view(np.float32), r["film"][..., 6].astype(np.float32))
        else:
            assert np.array_equal(g["nsteps"], o["nsteps"])
        assert np.array_equal(g["step_size"], o["step_size"])
    for (x, y) in trace_pixels:
        gt = ctx.trace_volume(p, view, sc.W, sc.H, x, y)
        ot = osc.trace(ov, x, y)
        assert gt.shape == ot.shape, (x, y, gt.shape, ot.shape)
        assert np.array_equal(gt, ot), (x, y)
    return g, g_rgba, r


def test_plain_volume_small(gpu_ctx):
    sc = W.plain_volume(n=64, W=96, H=80)
    _compare(gpu_ctx, sc, trace_pixels=[(48, 40), (10, 70), (90, 5)])


def test_plain_volume_f16_and_u8_pool(gpu_ctx):
    for fmt in (L.FMT_F16, L.FMT_U8):
        sc = W.plain_volume(n=48, W=64, H=64, fmt=fmt)
        _compare(gpu_ctx, sc, trace_pixels=[(32, 32)])


def test_plain_volume_density_one(gpu_ctx):
    # density 1 is the reference Volume shader literally (escalation triggers early)
    sc = W.plain_volume(n=64, W=64, H=64, density=1.0)
    _compare(gpu_ctx, sc)


@pytest.mark.parametrize("fmt", [L.FMT_F32, L.FMT_U8])
def test_ragged_frame_and_escalation(gpu_ctx, fmt):
    """A frame whose width and height are not multiples of the tile (partial tiles
    in the last column and row of the rotated tile grid; u8 frames run several
    lanes per ray), at density 1, so the step escalates inside groups."""
    sc = W.plain_volume(n=40, W=67, H=53, fmt=fmt, density=1.0)
    _compare(gpu_ctx, sc, trace_pixels=[(33, 26), (66, 52)])


@pytest.mark.slow
def test_plain_volume_full(gpu_ctx):
    sc = W.plain_volume()   # 256^3 FP32 brick, 512x512
    _compare(gpu_ctx, sc, trace_pixels=[(256, 256), (200, 300)], check_film=True)


def test_paged_small(gpu_ctx):
    # 256^3 Perlin in 64^3 + 1 apron bricks, 40 % resident: page faults stop rays
    sc = W.paged_volume(n=256, W=128, H=128)
    g, _, _ = _compare(gpu_ctx, sc, trace_pixels=[(64, 64), (30, 90), (100, 20)])


def test_lattice_schedule(gpu_ctx):
    sc = W.paged_volume(n=128, W=96, H=96, bs=32, resident=1.0, schedule=L.SCHED_LATTICE)
    sc.ert = 0.99
    _compare(gpu_ctx, sc)


def test_resume_equals_single_pass(gpu_ctx):
    """Partial residency, then make everything resident and resume: bit-equal to one pass
    (volume.cc:95-134 emit/march batching; shader.hpp:146-157)."""
    import torch

    sc = W.paged_volume(n=128, W=96, H=96, bs=32, resident=1.0)
    b = sc.bricked
    full_vaddr = b.vaddr.copy()
    part = full_vaddr.copy()
    rng = np.random.default_rng(5)
    drop = rng.random(part.shape) < 0.5
    part[drop] = -1
    ctx = gpu_ctx
    ctx.load(b, sc.tf)
    p = sc.volume_params()
    view = sc.view()
    ref = ctx.render_volume(p, view, sc.W, sc.H, film=True)
    ctx.set_page_table(part, b.page_k, b.page_b)
    first = ctx.render_volume(p, view, sc.W, sc.H, film=True)
    ctx.set_page_table(full_vaddr, b.page_k, b.page_b)
    resumed = ctx.resume_volume(p, first["film"])
    ctx.synchronize()
    # bitwise (finished rays hold negative nsteps, a NaN pattern when viewed as float)
    assert torch.equal(resumed["film"].view(torch.int32), ref["film"].view(torch.int32))
    # and the oracle's resume agrees
    osc_part = oracle_volume(sc, vaddr=part)
    r1 = osc_part.render(oracle_view(sc), 32, film=True)
    osc_full = oracle_volume(sc)
    r2 = osc_full.resume(r1["film"], 32)
    assert max_abs(resumed["rgba"].cpu().numpy(), r2["rgba"]) <= TOL


def test_sort_last_segments_equal_single_march(gpu_ctx):
    """In-GPU sort-last composite of per-brick partials over 3 residency batches equals the
    single lattice march (all bricks resident, no early termination) to rounding, takes the
    same samples, and matches the oracle's single march."""
    from hydrant_amd import dataset as D

    sc = W.paged_volume(n=256, W=128, H=128, bs=64, resident=1.0, schedule=L.SCHED_LATTICE)
    sc.ert = 0.0
    b = sc.bricked
    vol = sc.extra["volume"]
    ctx = gpu_ctx
    ctx.load(b, sc.tf)
    p, view = sc.volume_params(), sc.view()
    ref = ctx.render_volume(p, view, sc.W, sc.H, film=False, count=True)
    per = max(1, int(round(0.4 * b.n_nonempty)))
    batches = D.residency_batches(vol, b, per)
    assert len(batches) == 3
    rgba, cnt = ctx.render_sort_last(p, view, sc.W, sc.H, batches, b.grid, count=True)
    ctx.synchronize()
    err = max_abs(rgba.cpu().numpy(), ref["rgba"].cpu().numpy())
    print(f"sort-last vs single march: max-abs {err:.3g}, samples {int(cnt.item())} / {int(ref['samples'].item())}",
          flush=True)
    assert err <= 1e-5
    assert int(cnt.item()) == int(ref["samples"].item())
    o = oracle_volume(sc).render(oracle_view(sc), 32, film=False)
    assert max_abs(rgba.cpu().numpy(), o["rgba"]) <= TOL
    # the three batches kept in device memory and rendered by one launch: the same
    # slots, so the same image bit for bit and the same samples
    import torch

    dev = [(fmt, pad, torch.from_numpy(pool).cuda(), vaddr, pk, pb) for fmt, pad, pool, vaddr, pk, pb in batches]
    sset = ctx.segment_set(dev)
    rgba_set, cnt_set = ctx.render_sort_last_set(p, view, sc.W, sc.H, sset, count=True)
    ctx.synchronize()
    assert torch.equal(rgba_set, rgba)
    assert int(cnt_set.item()) == int(cnt.item())
    sset.close()
    # only two of the three batches ever resident: rays stop at the first missing brick
    part, _ = ctx.render_sort_last(p, view, sc.W, sc.H, batches[:2], b.grid)
    vaddr2 = np.full(b.vaddr.shape, -1, np.int32)
    vaddr2.reshape(-1)[b.order[:2 * per]] = 0   # any page: only residency matters for the stop
    full_pool = b.pool_values()
    vaddr_o = b.vaddr.copy()
    vaddr_o[vaddr2 < 0] = -1
    o2 = oracle_volume(sc, vaddr=vaddr_o).render(oracle_view(sc), 32, film=False)
    assert max_abs(part.cpu().numpy(), o2["rgba"]) <= TOL


def test_pixel_pass_matches_oracle(gpu_ctx):
    from oracle import oracle as O

    sc = W.plain_volume(n=32, W=40, H=30)
    gpu_ctx.load(sc.bricked, sc.tf)
    out = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=False)
    rgb = gpu_ctx.pixel_pass(out["rgba"], clear=(7, 8, 9)).cpu().numpy()
    ref = O.pixel_pass(out["rgba"].cpu().numpy(), clear=(7, 8, 9))
    assert np.array_equal(rgb, ref)


def test_fetch_and_composite(gpu_ctx):
    """Sort-last composite of partial images (volume.cc:267-292) against the oracle."""
    import torch
    from oracle import oracle as O

    rng = np.random.default_rng(3)
    N, H, Wd = 4, 24, 32
    parts = []
    for i in range(N):
        a = rng.random((H, Wd)).astype(np.float32) * 0.5
        rgb = rng.random((H, Wd, 3)).astype(np.float32) * a[..., None]
        p = np.zeros((H, Wd, 8), np.float32)
        p[..., 0:3] = rgb
        p[..., 3] = 1 - a
        p[..., 4:7] = rgb
        p[..., 7] = a
        parts.append(p)
    parts = np.stack(parts)
    z = [2, 0, 3, 1]
    g_rgba, g_rgb = gpu_ctx.composite(torch.from_numpy(parts).cuda(), z)
    o_rgba, o_rgb = O.composite(parts, z)
    assert max_abs(g_rgba.cpu().numpy(), o_rgba) == 0.0
    # the sRGB bytes come from host-computed thresholds: equal to the host's bytes
    assert np.array_equal(g_rgb.cpu().numpy(), o_rgb)


@pytest.mark.parametrize("mode", [L.MODE_PARTITION, L.MODE_PAGING])
def test_recolor_modes(gpu_ctx, mode):
    """Partition and Paging recolor (volume_shader.cu:54-66) on a paged scene, bit for bit
    against the oracle (film, image, per-sample traces, sample counts)."""
    sc = W.paged_volume(n=256, W=96, H=96, resident=0.5)
    sc.mode = mode
    sc.rank = 0.3
    b = sc.bricked
    b.lowest_blkcnt = b.n_pages // 2
    _, g_rgba, r = _compare(gpu_ctx, sc, trace_pixels=[(48, 48), (20, 70)])
    assert max_abs(g_rgba, r["rgba"].astype(np.float32)) == 0.0
    lit = g_rgba[..., 3] > 0
    assert lit.sum() > 100
    if mode == L.MODE_PARTITION:
        # every sample's colour is proportional to (1 - rank, 0, rank)
        assert np.all(g_rgba[..., 1] == 0)
        np.testing.assert_allclose(g_rgba[..., 2][lit] * 0.7, g_rgba[..., 0][lit] * 0.3, rtol=1e-5, atol=1e-7)
    else:
        assert np.all(g_rgba[..., 2] == 0)
        assert (g_rgba[..., 0][lit] > 0).any() and (g_rgba[..., 1][lit] > 0).any()   # both page levels


def test_fetch_pass_matches_oracle_film(gpu_ctx):
    """hyd_fetch_pass (VolumeShaderKernel::fetch, volume_shader.cu:23-30) copies theta, phi
    and v of every film pixel: equal to the GPU film bitwise and to the oracle's film."""
    import torch

    sc = W.paged_volume(n=128, W=64, H=48, bs=32, resident=0.6)
    gpu_ctx.load(sc.bricked, sc.tf)
    out = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=True)
    f = gpu_ctx.fetch_pass(out["film"])
    film = out["film"]
    assert torch.equal(f[..., 0:4].view(torch.int32), film[..., 12:16].view(torch.int32))
    assert torch.equal(f[..., 4:8].view(torch.int32), film[..., 8:12].view(torch.int32))
    o = oracle_film_fields(oracle_volume(sc).render(oracle_view(sc), 32, film=True)["film"])
    fn = f.cpu().numpy()
    assert max_abs(fn[..., 0:3], o["theta"]) == 0.0
    assert max_abs(fn[..., 3], o["phi"]) == 0.0
    assert max_abs(fn[..., 4:8], o["v"]) == 0.0
    assert (fn[..., 7] > 0).sum() > 100


def test_error_paths(gpu_ctx):
    import ctypes as C

    h = C.c_void_p()
    L.call("hyd_create", 0, C.byref(h))
    try:
        sc = W.plain_volume(n=16, W=8, H=8)
        p = sc.volume_params()
        v = sc.view()
        code = L.lib().hyd_render_volume(h, C.byref(p), C.byref(v), 8, 8, None, None, None)
        assert code == L.HYD_ERR_STATE
        assert b"pool" in L.lib().hyd_last_error()
        big = np.zeros((300, 4), np.float32)
        code = L.lib().hyd_set_tf(h, 300, L.ptr(big, C.c_float))
        assert code == L.HYD_ERR_RANGE
    finally:
        L.call("hyd_destroy", h)


def test_offsets64_variant_equals_default(gpu_ctx):
    """The 64-bit-offset march (pools of 4 GiB and more) renders the same film, bit for bit,
    as the 32-bit one; a child process forces it on a small pool."""
    import os
    import subprocess
    import sys
    import textwrap

    code = textwrap.dedent("""
        import sys, torch
        sys.path.insert(0, '.')
        from hydrant_amd import workloads as W
        from hydrant_amd.hip import Context
        sc = W.paged_volume(n=256, W=64, H=64, resident=0.5, density=0.05)
        ctx = Context(0)
        ctx.load(sc.bricked, sc.tf)
        out = ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=True, count=True)
        torch.cuda.synchronize()
        torch.save({'film': out['film'].cpu(), 'n': int(out['samples'].item())}, sys.argv[1])
    """)
    outs = []
    for flag in ("0", "1"):
        path = f"/tmp/hyd_off64_{os.getpid()}_{flag}.pt"
        env = dict(os.environ, HYD_VOLUME_OFFSETS64=flag)
        subprocess.run([sys.executable, "-c", code, path], check=True, env=env, timeout=300)
        outs.append(torch_load(path))
        os.remove(path)
    import torch

    assert outs[0]["n"] == outs[1]["n"] > 0
    assert torch.equal(outs[0]["film"].view(torch.int32), outs[1]["film"].view(torch.int32))


def torch_load(path):
    import torch

    return torch.load(path, weights_only=True)
