#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

A step is one render pass of the workload over inputs already resident in
HBM.  Workloads:
  plain    one 512x512 view of a 256^3 Gaussian-blob volume stored as one FP32
           brick, reference step schedule (escalation), density 1e-2 (the plain
           single-view FP32 form).
  batched  the batched FP16-brick form (see hydrant_amd/batch.py): many views of
           several volumes in one persistent launch.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sort-last design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    res = bench_plain(torch, args)
    sc = res["scene"]
    out = {
        "metric": METRIC,
        "value": round(res["msamples_per_s"], 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain_volume_fp32: 256^3 u8 Gaussian-blob volume as one FP32 brick, 512x512, "
                               "orbit [25,65,2.2], README 9-point TF, step 0.5 voxel, density 1e-2, "
                               "reference schedule",
                   "resolution": [sc.W, sc.H], "volume": [256, 256, 256], "brick_format": "f32",
                   "samples_per_frame": res["n_samples"]},
        "ms_per_frame": round(res["ms_per_step"], 4),
        "roofline": {"bound": "hbm", "achieved": round(res["achieved_gbs"], 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(res["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "volume_kernel<f32>", "kernel_ms": round(res["kernel_ms"], 4),
                     "alg_bytes_per_launch": res["alg_bytes"]},
    }
    if not args.no_cpu:
        cb = cpu_baseline_plain(sc)
        err = float(np.max(np.abs(res["rgba"].cpu().numpy().astype(np.float64) - cb["rgba"])))
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["value"] = round(cb["value"], 3)
        out["max_abs_vs_cpu"] = err
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
