#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

A step is one render pass of the workload over inputs already resident in
HBM.  Workloads:
  plain    one 512x512 view of a 256^3 Gaussian-blob volume stored as one FP32
           brick, reference step schedule (escalation), density 1e-2 (the plain
           single-view FP32 form).
  batched  the batched FP16-brick form (hydrant_amd/hip.py Batch over
           k_batch.hip): many views of several volumes in one persistent launch.
  diff     the differentiable form (hydrant_amd/diff.py): 32 views at 256^2 of a
           128^3 FP32 volume, 512 samples per ray, transmittance checkpoints every
           64; forward and forward + adjoint times, and a 200-step Adam fit.
  paged    one 512x512 view of a 1024^3 u8 Perlin volume in 64^3 + 1 bricks, 40%
           resident, reference schedule (ms/frame).
  iso      one 1024x1024 isosurface view of a 512^3 FP16 sphere SDF (ms/frame).
Unit of work: one composited sample (a march step that does a TF lookup and a
front-to-back composite); the count comes from the kernel's counting variant
and equals the oracle's.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"

METRIC = "Msamples/s + ms/frame at 512², plain fp32 vs batched fp16; fwd+adjoint ms"
HBM_PEAK_GBS = 8000.0


def cpu_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE in KiB per dispatch, FETCH
    scaled by the plain-copy calibration factor), or None."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())[kernel]
    except (OSError, KeyError, ValueError):
        return None
    return int(rec["fetch_kib"] * 1024 * rec["fetch_calibration"] + rec["write_kib"] * 1024)


def time_kernel(torch, fn, steps: int, warmup: int):
    """Warm up, then time exactly `steps` calls: wall clock bracketed by synchronize,
    and per-launch HIP event durations on torch's current stream (the library's stream)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        fn()
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall, ev


def bench_plain(torch, args):
    from hydrant_amd import workloads as W
    from hydrant_amd.hip import Context

    sc = W.plain_volume(n=256, W=512, H=512)
    ctx = Context(0)
    ctx.load(sc.bricked, sc.tf)
    p = sc.volume_params()
    view = sc.view()
    rgba = torch.empty((sc.H, sc.W, 4), dtype=torch.float32, device="cuda")
    cnt = ctx.render_volume(p, view, sc.W, sc.H, film=False, count=True, rgba=rgba)["samples"]
    torch.cuda.synchronize()
    n_samples = int(cnt.item())

    def step():
        ctx.render_volume(p, view, sc.W, sc.H, film=False, count=False, rgba=rgba)

    wall, ev = time_kernel(torch, step, args.steps, args.warmup)
    ms_step = wall * 1e3 / args.steps
    k_ms = float(np.mean(ev))
    bytes_per_launch = n_samples * 8 * 4 + sc.W * sc.H * 16
    res = {
        "workload": "plain_volume_fp32",
        "n_samples": n_samples,
        "ms_per_step": ms_step,
        "kernel_ms": k_ms,
        "kernel_ms_median": float(np.median(ev)),
        "msamples_per_s": n_samples / (ms_step * 1e-3) / 1e6,
        "alg_bytes": bytes_per_launch,
        "achieved_gbs": bytes_per_launch / (k_ms * 1e-3) / 1e9,
        "scene": sc,
        "rgba": rgba,
    }
    ctx.close()
    return res


def bench_paged(torch, args):
    """Paged Volume: 1024^3 u8 Perlin fBm in 64^3 + 1 bricks, 40% of the bricks resident
    through a seeded page table (a ray stops at a non-resident brick), empty bricks
    skipped by the chebyshev grid, 512^2, reference schedule, u8 pool."""
    from hydrant_amd import workloads as W
    from hydrant_amd.hip import Context

    t0 = time.perf_counter()
    sc = W.paged_volume()
    gen_s = time.perf_counter() - t0
    ctx = Context(0)
    ctx.load(sc.bricked, sc.tf)
    p = sc.volume_params()
    view = sc.view()
    rgba = torch.empty((sc.H, sc.W, 4), dtype=torch.float32, device="cuda")
    n_samples = int(ctx.render_volume(p, view, sc.W, sc.H, film=False, count=True, rgba=rgba)["samples"].item())

    def step():
        ctx.render_volume(p, view, sc.W, sc.H, film=False, count=False, rgba=rgba)

    wall, ev = time_kernel(torch, step, args.steps, args.warmup)
    ms = wall * 1e3 / args.steps
    k_ms = float(np.mean(ev))
    alg = n_samples * 8 * 1 + sc.W * sc.H * 16
    ctx.close()
    return {"ms_per_frame": round(ms, 4), "kernel_ms": round(k_ms, 4), "samples_per_frame": n_samples,
            "msamples_per_s": round(n_samples / (ms * 1e-3) / 1e6, 3),
            "achieved_gbs": round(alg / (k_ms * 1e-3) / 1e9, 1),
            "roofline_frac": round(alg / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "alg_bytes_rule": "samples x 8 taps x 1 B + pixels x 16 B",
            "resident_pages": int((sc.bricked.vaddr >= 0).sum()), "bricks": int(sc.bricked.vaddr.size),
            "volume": 1024, "res": sc.W, "gen_s": round(gen_s, 1), "scene": sc, "rgba": rgba}


def bench_paged_sort_last(torch, args):
    """Paged Volume as configured: 1024^3 u8 Perlin in 64^3 + 1 bricks, rendered over
    residency batches of 40% of the bricks (1638 pages each) until every non-empty brick
    has been resident, each batch writing its bricks' per-ray partials, then the in-GPU
    sort-last composite (lattice schedule, no early termination).  Every batch's pool is
    in HBM before the timed region, the batches held as one segment set: a step clears the
    partial slots, renders every batch's segments in one launch and composites.  The
    sequential form (bind each batch's pool, set its page table, one pass per batch) is
    timed beside it."""
    from hydrant_amd import _lib as L
    from hydrant_amd import dataset as D
    from hydrant_amd import workloads as W
    from hydrant_amd.hip import Context

    t0 = time.perf_counter()
    sc = W.paged_volume(resident=1.0, schedule=L.SCHED_LATTICE)
    sc.ert = 0.0
    b = sc.bricked
    per = int(round(0.4 * b.vaddr.size))
    host_batches = D.residency_batches(sc.extra["volume"], b, per)
    gen_s = time.perf_counter() - t0
    ctx = Context(0)
    ctx.set_tf(sc.tf)
    ctx.set_empty_grid(b.cheb)
    dev_batches = [(fmt, pad, torch.from_numpy(pool).cuda(), vaddr, pk, pb)
                   for fmt, pad, pool, vaddr, pk, pb in host_batches]
    p, view = sc.volume_params(), sc.view()
    max_slots = int(sum(b.grid))
    slots = torch.empty((max_slots, sc.H, sc.W, 4), dtype=torch.float32, device="cuda")
    nslots = torch.zeros((sc.H, sc.W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((sc.H, sc.W, 4), dtype=torch.float32, device="cuda")
    sset = ctx.segment_set(dev_batches)
    _, cnt = ctx.render_sort_last_set(p, view, sc.W, sc.H, sset, slots, nslots, rgba, count=True)
    torch.cuda.synchronize()
    n_samples = int(cnt.item())

    def step_seq():
        ctx.render_sort_last_resident(p, view, sc.W, sc.H, dev_batches, b.grid, slots, nslots, rgba)

    def step():
        ctx.render_sort_last_set(p, view, sc.W, sc.H, sset, slots, nslots, rgba)

    wall_seq, _ = time_kernel(torch, step_seq, args.steps, args.warmup)
    wall, ev = time_kernel(torch, step, args.steps, args.warmup)
    ms = wall * 1e3 / args.steps
    alg = n_samples * 8 * 1 + sc.W * sc.H * 16
    res = {"ms_per_frame": round(ms, 4), "samples_per_frame": n_samples,
           "sequential_batches_ms_per_frame": round(wall_seq * 1e3 / args.steps, 4),
           "msamples_per_s": round(n_samples / (ms * 1e-3) / 1e6, 3),
           "achieved_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "roofline_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "alg_bytes_rule": "samples x 8 taps x 1 B + pixels x 16 B, over the whole frame (segments, clear, fold)",
           "batches": len(dev_batches), "pages_per_batch": per, "bricks": int(b.vaddr.size),
           "nonempty_bricks": int(b.n_nonempty), "max_slots": max_slots, "volume": 1024, "res": sc.W,
           "gen_s": round(gen_s, 1), "scene": sc, "rgba": rgba}
    sset.close()
    ctx.close()
    return res


def sort_last_check(res):
    """The sort-last frame against the FP32 oracle's single lattice march over every
    brick, on every 7th pixel."""
    from tests.helpers import oracle_view, oracle_volume

    sc = res["scene"]
    idx = np.arange(0, sc.W * sc.H, 7)
    xy = np.stack([idx % sc.W, idx // sc.W], 1).astype(np.int32)
    ref, _ = oracle_volume(sc).render_pixels(oracle_view(sc), xy, prec=32, nthreads=cpu_threads())
    got = res["rgba"].cpu().numpy()[xy[:, 1], xy[:, 0]].astype(np.float64)
    return {"sort_last_max_abs_vs_oracle": float(np.max(np.abs(got - ref))), "sort_last_pixels_checked": int(len(idx))}


def bench_iso(torch, args):
    """Isosurface: 512^3 FP16 SDF of 64 seeded spheres (smooth union), 64^3 + 1 bricks,
    1024^2, iso 0, bracket on the sign change then 8 bisections, Phong colour."""
    from hydrant_amd import _lib as L
    from hydrant_amd import workloads as W
    from hydrant_amd.hip import Context

    sc = W.isosurface()
    ctx = Context(0)
    ctx.load(sc.bricked)
    p = W.iso_params(sc, L.ISO_COLOR)
    view = sc.view()
    px = torch.empty((sc.H, sc.W, 6), dtype=torch.float32, device="cuda")
    n_samples = int(ctx.render_iso(p, view, sc.W, sc.H, count=True, px=px)["samples"].item())

    def step():
        ctx.render_iso(p, view, sc.W, sc.H, count=False, px=px)

    wall, ev = time_kernel(torch, step, args.steps, args.warmup)
    ms = wall * 1e3 / args.steps
    k_ms = float(np.mean(ev))
    alg = n_samples * 8 * 2 + sc.W * sc.H * 24
    hits = int((px[..., 5].view(torch.int32) == 1).sum().item())
    ctx.close()
    return {"ms_per_frame": round(ms, 4), "kernel_ms": round(k_ms, 4), "samples_per_frame": n_samples,
            "msamples_per_s": round(n_samples / (ms * 1e-3) / 1e6, 3),
            "achieved_gbs": round(alg / (k_ms * 1e-3) / 1e9, 1),
            "roofline_frac": round(alg / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "alg_bytes_rule": "samples x 8 taps x 2 B + pixels x 24 B",
            "hits": hits, "volume": 512, "res": sc.W, "scene": sc, "params": p, "px": px}


def paged_iso_checks(paged, iso):
    """The paged image against the FP32 oracle on every 7th pixel (reference schedule, the
    kernel reproduces the CPU backend's arithmetic); the isosurface hit mask (XOR count)
    and hit depth against the FP32 oracle's full frame."""
    from oracle import oracle as O
    from tests.helpers import oracle_view, oracle_volume

    out = {}
    if paged is not None:
        sc = paged["scene"]
        xs, ys = np.meshgrid(np.arange(0, sc.W, 7), np.arange(0, sc.H, 7))
        xy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
        ref, _ = oracle_volume(sc).render_pixels(oracle_view(sc), xy, prec=32, nthreads=cpu_threads())
        got = paged["rgba"].cpu().numpy()[xy[:, 1], xy[:, 0]].astype(np.float64)
        out["paged_max_abs_vs_oracle"] = float(np.max(np.abs(got - ref)))
        out["paged_pixels_checked"] = int(xy.shape[0])
    if iso is not None:
        sc, p = iso["scene"], iso["params"]
        b = sc.bricked
        osc = O.IsoScene(sc.W, sc.H, sc.bbox_min, sc.bbox_max, p.step, p.du, np.array(p.to_world[:]), p.eye_pos[:],
                         p.light_pos[:], p.surface_color[:], p.isovalue, b.cheb, b.vaddr, b.page_k, b.page_b,
                         b.pool_values(), b.pad, mode=p.mode, corrected=p.corrected, bisect_steps=p.bisect_steps,
                         max_steps=p.max_steps)
        r = osc.render(oracle_view(sc), 32, nthreads=cpu_threads())
        px = iso["px"].cpu().numpy()
        g_hit, o_hit = px[..., 5].view(np.int32), r["px"][..., 5].astype(np.int32)
        hit = g_hit == 1
        out["iso_hit_xor"] = int(np.count_nonzero(g_hit != o_hit))
        out["iso_max_depth_diff_voxel"] = float(np.max(np.abs(px[..., 4][hit] - r["px"][..., 4][hit])) * b.bs) \
            if hit.any() else 0.0
        out["iso_samples_equal"] = bool(r["samples"] == iso["samples_per_frame"])
    return out


MAX_ERROR = 1e-3   # max-abs per channel against the FP64 oracle


def bench_batched(torch, args, fmt=None):
    """96 views (resolutions in {256..1024}^2) of three blob volumes (256^3, 512^3, 768^3)
    as FP16 (or u8) bricks, one persistent launch replayed from a HIP graph; launch shape
    and early-termination threshold from the autotuner (error budget 1e-3)."""
    from hydrant_amd import _lib as L
    from hydrant_amd import workloads as W
    from hydrant_amd.hip import Context

    fmt = L.FMT_F16 if fmt is None else fmt
    t0 = time.perf_counter()
    wl = W.batched(fmt=fmt)
    t_gen = time.perf_counter() - t0
    ctx = Context(0)
    ctx.upload_pool(wl.fmt, wl.bs + 2 * wl.apron, wl.pool)
    ctx.set_tf(wl.tf)
    ctx.set_volumes(wl.volume_descs())
    b = ctx.batch(wl.hip_views())
    out = b.new_output()
    if args.config:
        from hydrant_amd.hip import batch_config

        parts = [int(x) for x in args.config.split(",")]
        best, tune_ms = batch_config(*parts), None
    else:
        best, wl.ert, tune_ms = b.autotune_schedule(wl.density, MAX_ERROR, out, reps=2)
    _, cnt = b.render(best, wl.density, wl.ert, out=out, count=True)
    torch.cuda.synchronize()
    n_samples = int(cnt.item())
    b.capture(best, wl.density, wl.ert, out)

    def step():
        b.replay()

    wall, ev = time_kernel(torch, step, args.steps, args.warmup)
    ms_step = wall * 1e3 / args.steps
    k_ms = float(np.mean(ev))
    n_px = b.n_pixels
    vb = {L.FMT_U8: 1, L.FMT_F16: 2, L.FMT_F32: 4}[fmt]
    alg = n_samples * 8 * vb + n_px * 16
    res = {"n_samples": n_samples, "ms_per_step": ms_step, "kernel_ms": k_ms,
           "msamples_per_s": n_samples / (ms_step * 1e-3) / 1e6, "alg_bytes": alg, "bytes_per_voxel": vb,
           "achieved_gbs": alg / (k_ms * 1e-3) / 1e9, "config": best.as_dict(), "autotune_best_ms": tune_ms,
           "ert": wl.ert, "n_pixels": n_px, "gen_s": t_gen, "wl": wl, "batch": b, "out": out, "ctx": ctx}
    return res


def bench_diff(torch, args):
    """Differentiable form: forward (with checkpoints) and forward + adjoint over 32 views at
    256^2 of a 128^3 FP32 volume, K = 512; then an Adam fit (lr 1e-2) of a constant volume
    and a grey TF to the hidden pair's renders."""
    from hydrant_amd import workloads as W
    from hydrant_amd.diff import DiffRenderer
    from hydrant_amd.hip import Context

    wl = W.differentiable()
    ctx = Context(0)
    r = DiffRenderer(ctx, wl.views(), wl.W, wl.H, (wl.n,) * 3, wl.n_steps, wl.ckpt_every, wl.sigma)
    vol = torch.from_numpy(wl.hidden_volume).cuda()
    tf = torch.from_numpy(wl.hidden_tf).cuda()
    target, ck, cnt = r.forward_raw(vol, tf, count=True)
    torch.cuda.synchronize()
    n_samples = int(cnt.item())
    g = torch.randn(target.shape, generator=torch.Generator(device="cuda").manual_seed(7), device="cuda")

    def fwd():
        r.forward_raw(vol, tf)

    def fwd_adj():
        _, c, _ = r.forward_raw(vol, tf)
        r.backward_raw(vol, tf, c, g)

    wall_f, ev_f = time_kernel(torch, fwd, args.steps, args.warmup)
    wall_b, ev_b = time_kernel(torch, fwd_adj, args.steps, args.warmup)
    f_ms, fb_ms = wall_f * 1e3 / args.steps, wall_b * 1e3 / args.steps
    res = {"views": len(wl.arms), "res": wl.W, "volume": wl.n, "n_steps": wl.n_steps, "ckpt_every": wl.ckpt_every,
           "samples_per_forward": n_samples, "fwd_ms": round(f_ms, 4), "fwd_kernel_ms": round(float(np.mean(ev_f)), 4),
           "fwd_adjoint_ms": round(fb_ms, 4), "fwd_adjoint_event_ms": round(float(np.mean(ev_b)), 4),
           "ratio_fwd_adjoint_vs_fwd": round(fb_ms / f_ms, 3),
           "fwd_msamples_per_s": round(n_samples / (f_ms * 1e-3) / 1e6, 3)}
    if args.diff_fit_steps > 0:
        v = torch.from_numpy(wl.init_volume).cuda().requires_grad_(True)
        t = torch.from_numpy(wl.init_tf).cuda().requires_grad_(True)
        opt = torch.optim.Adam([v, t], lr=1e-2)
        losses = []
        t0 = time.perf_counter()
        for _ in range(args.diff_fit_steps):
            opt.zero_grad()
            loss = torch.mean((r.render(v, t) - target) ** 2)
            loss.backward()
            opt.step()
            losses.append(loss.detach())
        torch.cuda.synchronize()
        ls = [float(x.item()) for x in losses]
        res["adam_fit"] = {"steps": args.diff_fit_steps, "lr": 1e-2, "loss_first": ls[0], "loss_last": ls[-1],
                           "loss_min": min(ls), "s": round(time.perf_counter() - t0, 3)}
    ctx.close()
    return res


def batched_errors(res, stride: int = 16):
    """Per-view max-abs error of a batched output against the FP64 oracle over every view:
    every stride-th pixel (offset by the view index), exact u8 source values, lattice
    schedule with early termination off, so the error includes the truncation of the
    autotuned termination threshold."""
    from tests.helpers import exact_pool_values, oracle_view, oracle_volume

    wl, b, out = res["wl"], res["batch"], res["out"]
    pools = {}
    errs = []
    t0 = time.perf_counter()
    for i in range(len(wl.views)):
        sc = wl.scene(i)
        sc.ert = 0.0
        vi = wl.views[i]["volume"]
        if vi not in pools:
            pools[vi] = exact_pool_values(sc)
        idx = np.arange(i % stride, sc.W * sc.H, stride)
        xy = np.stack([idx % sc.W, idx // sc.W], 1).astype(np.int32)
        osc = oracle_volume(sc, pool_values=pools[vi], exact_mapping=True)
        ref, _ = osc.render_pixels(oracle_view(sc), xy, prec=64, nthreads=cpu_threads())
        img = b.view_image(out, i).cpu().numpy()
        got = img[xy[:, 1], xy[:, 0]].astype(np.float64)
        errs.append({"view": i, "res": sc.W, "volume": vi, "pixels": int(xy.shape[0]),
                     "max_abs": float(np.max(np.abs(got - ref)))})
    summary = {"views": len(errs), "pixels": int(sum(e["pixels"] for e in errs)),
               "max_abs": max(e["max_abs"] for e in errs), "stride": stride, "oracle_ert": 0.0,
               "cpu_s": round(time.perf_counter() - t0, 1)}
    return errs, summary


def cpu_baseline_batched(res, budget_s: float = 12.0):
    """FP32 oracle (lattice schedule) on whole views of the batched workload until the budget
    is spent; FP16 storage is replaced by the exact u8 values, which costs the CPU the same."""
    from tests.helpers import exact_pool_values, oracle_view, oracle_volume

    wl = res["wl"]
    nthreads = cpu_threads()
    samples, t, views = 0, 0.0, 0
    for i in range(len(wl.views)):
        sc = wl.scene(i)
        osc = oracle_volume(sc, pool_values=exact_pool_values(sc), exact_mapping=True)
        t0 = time.perf_counter()
        r = osc.render(oracle_view(sc), 32, nthreads=nthreads, film=False)
        t += time.perf_counter() - t0
        samples += r["samples"]
        views += 1
        if t > budget_s:
            break
    return {"value": round(samples / t / 1e6, 3), "unit": "Msamples/s", "cores": nthreads, "kind": "port",
            "sample": f"first {views} of {len(wl.views)} batched views rendered whole by the FP32 oracle "
                      f"(lattice schedule, {samples} samples in {t:.1f} s, {cpu_model()})"}


def cpu_baseline_plain(sc, budget_s: float = 12.0):
    """Oracle FP32 multi-thread path (the reference CPU backend restated) on the same
    workload; whole frames, 1 warm-up, then as many frames as fit the budget."""
    from oracle import oracle as O
    from tests.helpers import oracle_view, oracle_volume

    nthreads = cpu_threads()
    osc = oracle_volume(sc)
    ov = oracle_view(sc)
    r = osc.render(ov, 32, nthreads=nthreads, film=False)
    frames, t0 = 0, time.perf_counter()
    while True:
        r = osc.render(ov, 32, nthreads=nthreads, film=False)
        frames += 1
        if time.perf_counter() - t0 > budget_s or frames >= 20:
            break
    dt = (time.perf_counter() - t0) / frames
    return {"value": r["samples"] / dt / 1e6, "unit": "Msamples/s", "cores": nthreads, "kind": "port",
            "sample": f"{frames} full 512x512 frames of the plain FP32 workload after 1 warm-up "
                      f"({dt * 1e3:.1f} ms/frame, {cpu_model()})",
            "ms_per_frame": dt * 1e3, "rgba": r["rgba"]}


def fp64_errors(plain, paged, sortlast, iso):
    """Each frame against the FP64 oracle of the same schedule: plain (full frame) and paged
    (every 7th pixel) under the reference schedule, whose escalation thresholds and integer
    step budget can move a whole sample between FP32 and FP64 on a few rays (bounded by one
    sample's opacity, density x max TF alpha = 1e-2); sort-last (every 7th pixel) under the
    lattice schedule; the isosurface hit mask and depth (full frame)."""
    from oracle import oracle as O
    from tests.helpers import oracle_view, oracle_volume

    out = {}
    nt = cpu_threads()
    if plain is not None:
        sc = plain["scene"]
        ref = oracle_volume(sc).render(oracle_view(sc), 64, nthreads=nt, film=False)["rgba"]
        d = np.abs(plain["rgba"].cpu().numpy().astype(np.float64) - ref).max(axis=-1)
        out["plain"] = {"max_abs": float(d.max()), "p999": float(np.percentile(d, 99.9)),
                        "frac_within_1e-4": float(np.mean(d <= 1e-4)), "pixels": int(d.size)}
    for name, res, sched in (("paged", paged, None), ("sort_last", sortlast, None)):
        if res is None:
            continue
        sc = res["scene"]
        idx = np.arange(0, sc.W * sc.H, 7)
        xy = np.stack([idx % sc.W, idx // sc.W], 1).astype(np.int32)
        ref, _ = oracle_volume(sc).render_pixels(oracle_view(sc), xy, prec=64, nthreads=nt)
        got = res["rgba"].cpu().numpy()[xy[:, 1], xy[:, 0]].astype(np.float64)
        d = np.abs(got - ref).max(axis=-1)
        out[name] = {"max_abs": float(d.max()), "p999": float(np.percentile(d, 99.9)),
                     "frac_within_1e-4": float(np.mean(d <= 1e-4)), "pixels": int(d.size)}
    if iso is not None:
        sc, p = iso["scene"], iso["params"]
        b = sc.bricked
        osc = O.IsoScene(sc.W, sc.H, sc.bbox_min, sc.bbox_max, p.step, p.du, np.array(p.to_world[:]), p.eye_pos[:],
                         p.light_pos[:], p.surface_color[:], p.isovalue, b.cheb, b.vaddr, b.page_k, b.page_b,
                         b.pool_values(), b.pad, mode=p.mode, corrected=p.corrected, bisect_steps=p.bisect_steps,
                         max_steps=p.max_steps)
        r = osc.render(oracle_view(sc), 64, nthreads=nt)["px"]
        px = iso["px"].cpu().numpy()
        g_hit, o_hit = px[..., 5].view(np.int32) == 1, r[..., 5] == 1
        both = g_hit & o_hit
        dd = np.abs(px[..., 4][both] - r[..., 4][both]) * b.bs if both.any() else np.zeros(1)
        # a grazing hit in one precision can be a miss in the other, and the ray then
        # hits a farther surface: a few pixels differ by whole surfaces
        out["isosurface"] = {"hit_xor": int(np.count_nonzero(g_hit != o_hit)), "hits": int(o_hit.sum()),
                             "max_depth_diff_voxel": float(dd.max()),
                             "p999_depth_diff_voxel": float(np.percentile(dd, 99.9)),
                             "depth_diff_over_1_voxel": int(np.count_nonzero(dd > 1.0))}
    return out


def batched_info(res, plain):
    from hydrant_amd import _lib as L

    name = {L.FMT_U8: "u8", L.FMT_F16: "f16", L.FMT_F32: "f32"}[res["wl"].fmt]
    return {"msamples_per_s": round(res["msamples_per_s"], 3), "ms_per_batch": round(res["ms_per_step"], 4),
            "samples_per_batch": res["n_samples"], "views": len(res["wl"].views), "pixels": res["n_pixels"],
            "volumes": [256, 512, 768], "brick_format": name, "launch_config": res["config"],
            "ert": res["ert"], "error_budget": MAX_ERROR,
            "kernel": f"batch_kernel<{name}> (lattice schedule, HIP graph replay)",
            "kernel_ms": round(res["kernel_ms"], 4),
            "achieved_gbs": round(res["achieved_gbs"], 1),
            "alg_bytes_rule": f"samples x 8 taps x {res['bytes_per_voxel']} B + pixels x 16 B",
            "ratio_vs_plain_fp32": round(res["msamples_per_s"] / plain["msamples_per_s"], 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batched", choices=["batched", "plain"],
                    help="headline workload (the other one is reported alongside)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and the FP64 error checks")
    ap.add_argument("--err-stride", type=int, default=16,
                    help="batched views: every n-th pixel of every view is checked against the FP64 oracle")
    ap.add_argument("--no-paged-iso", action="store_true", help="skip the paged, sort-last and isosurface legs")
    ap.add_argument("--no-u8", action="store_true", help="skip the u8 batched leg")
    ap.add_argument("--only", choices=["plain", "batched", "batched_u8", "diff", "paged", "sortlast", "iso"],
                    help="profiling aid: run only this workload's timed launches and print its line")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-form leg")
    ap.add_argument("--diff-fit-steps", type=int, default=200, help="Adam steps of the differentiable fit")
    ap.add_argument("--config", help="batched launch shape tile_w,unroll,waves_per_cu[,group] (skips the autotuner)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="profiling aid: copy a MIB-sized buffer 3 times (scales FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sort-last rendering is not provided by this library", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    from hydrant_amd import _lib as L

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4}), flush=True)
        return
    if args.only == "diff":
        print(json.dumps(bench_diff(torch, args)), flush=True)
        return
    if args.only:
        fns = {"plain": bench_plain, "batched": bench_batched, "paged": bench_paged, "iso": bench_iso,
               "sortlast": bench_paged_sort_last,
               "batched_u8": lambda t, a: bench_batched(t, a, fmt=L.FMT_U8)}
        res = fns[args.only](torch, args)
        print(json.dumps({k: v for k, v in res.items() if isinstance(v, (int, float, str, dict))}), flush=True)
        return

    plain = bench_plain(torch, args)
    print(f"# plain fp32: {plain['msamples_per_s']:.1f} Msamples/s, {plain['ms_per_step']:.3f} ms/frame",
          file=sys.stderr, flush=True)
    batched = bench_batched(torch, args)
    print(f"# batched fp16: {batched['msamples_per_s']:.1f} Msamples/s, {batched['ms_per_step']:.3f} ms/batch, "
          f"config {batched['config']}, ert {batched['ert']}", file=sys.stderr, flush=True)
    batched_u8 = None
    if not args.no_u8:
        batched_u8 = bench_batched(torch, args, fmt=L.FMT_U8)
        print(f"# batched u8: {batched_u8['msamples_per_s']:.1f} Msamples/s, {batched_u8['ms_per_step']:.3f} "
              f"ms/batch, config {batched_u8['config']}, ert {batched_u8['ert']}", file=sys.stderr, flush=True)

    diff = None if args.no_diff else bench_diff(torch, args)
    if diff is not None:
        print(f"# diff: fwd {diff['fwd_ms']:.3f} ms, fwd+adjoint {diff['fwd_adjoint_ms']:.3f} ms "
              f"(x{diff['ratio_fwd_adjoint_vs_fwd']:.2f})", file=sys.stderr, flush=True)

    paged = iso = sortlast = None
    if not args.no_paged_iso:
        paged = bench_paged(torch, args)
        print(f"# paged u8: {paged['ms_per_frame']:.3f} ms/frame, {paged['msamples_per_s']:.1f} Msamples/s",
              file=sys.stderr, flush=True)
        sortlast = bench_paged_sort_last(torch, args)
        print(f"# paged u8 sort-last ({sortlast['batches']} residency batches): {sortlast['ms_per_frame']:.3f} "
              f"ms/frame, {sortlast['msamples_per_s']:.1f} Msamples/s (one pass per batch: "
              f"{sortlast['sequential_batches_ms_per_frame']:.3f} ms)", file=sys.stderr, flush=True)
        iso = bench_iso(torch, args)
        print(f"# isosurface fp16: {iso['ms_per_frame']:.3f} ms/frame, {iso['msamples_per_s']:.1f} Msamples/s",
              file=sys.stderr, flush=True)

    head = batched if args.workload == "batched" else plain
    plain_info = {"msamples_per_s": round(plain["msamples_per_s"], 3), "ms_per_frame": round(plain["ms_per_step"], 4),
                  "samples_per_frame": plain["n_samples"], "kernel": "volume_kernel<f32> (reference schedule)",
                  "kernel_ms": round(plain["kernel_ms"], 4), "achieved_gbs": round(plain["achieved_gbs"], 1),
                  "roofline_frac": round(plain["achieved_gbs"] / HBM_PEAK_GBS, 4),
                  "alg_bytes_rule": "samples x 8 taps x 4 B + pixels x 16 B"}
    if head is batched:
        workload = ("batched_fp16: 96 views (seeded resolutions in {256,384,512,768,1024}^2, seeded orbit "
                    "angles) of three Gaussian-blob volumes 256^3/512^3/768^3 as FP16 64^3+1 bricks, one "
                    f"persistent launch replayed from a HIP graph, lattice schedule (early termination at "
                    f"{batched['ert']}, autotuned within a 1e-3 error budget), README TF, density 1e-2, "
                    "step 0.5 voxel")
        bytes_per_voxel, kern = 2, "batch_kernel<f16>"
    else:
        workload = ("plain_volume_fp32: 256^3 u8 Gaussian-blob volume as one FP32 brick, 512x512, orbit "
                    "[25,65,2.2], README 9-point TF, step 0.5 voxel, density 1e-2, reference schedule")
        bytes_per_voxel, kern = 4, "volume_kernel<f32>"
    out = {
        "metric": METRIC,
        "value": round(head["msamples_per_s"], 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "samples_per_step": head["n_samples"]},
        "plain_fp32": plain_info,
        "batched_fp16": batched_info(batched, plain),
        "batched_u8": None if batched_u8 is None else batched_info(batched_u8, plain),
        "diff": diff,
        "paged_u8": None if paged is None else {k: v for k, v in paged.items() if k not in ("scene", "rgba")},
        "paged_sort_last_u8": None if sortlast is None else {k: v for k, v in sortlast.items()
                                                             if k not in ("scene", "rgba")},
        "isosurface_f16": None if iso is None else {k: v for k, v in iso.items()
                                                    if k not in ("scene", "params", "px")},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(kern),
                     "kernel": kern, "kernel_ms": round(head["kernel_ms"], 4),
                     "alg_bytes_per_launch": head["alg_bytes"],
                     "alg_bytes_rule": f"samples x 8 taps x {bytes_per_voxel} B + pixels x 16 B"},
    }
    if not args.no_cpu:
        if sortlast is not None:
            out["paged_sort_last_u8"].update(sort_last_check(sortlast))
        if paged is not None or iso is not None:
            chk = paged_iso_checks(paged, iso)
            if paged is not None:
                out["paged_u8"].update({k: v for k, v in chk.items() if k.startswith("paged_")})
            if iso is not None:
                out["isosurface_f16"].update({k: v for k, v in chk.items() if k.startswith("iso_")})
        out["fp64_error"] = fp64_errors(plain, paged, sortlast, iso)
        print("# fp64 errors of the single-view frames done", file=sys.stderr, flush=True)
        for key, res in (("batched_fp16", batched), ("batched_u8", batched_u8)):
            if res is None:
                continue
            errs, summary = batched_errors(res, args.err_stride)
            out[key]["max_abs_vs_fp64"] = summary["max_abs"]
            out[key]["fp64_check"] = summary
            out[key]["per_view_error"] = [round(e["max_abs"], 8) for e in errs]
            print(f"# {key}: max-abs {summary['max_abs']:.3g} vs FP64 over {summary['views']} views, "
                  f"{summary['pixels']} pixels", file=sys.stderr, flush=True)
        cb = cpu_baseline_plain(plain["scene"])
        plain_cpu = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        plain_cpu["value"] = round(cb["value"], 3)
        if head is batched:
            out["cpu_baseline"] = cpu_baseline_batched(batched)
            out["cpu_baseline_plain"] = plain_cpu
        else:
            out["cpu_baseline"] = plain_cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
