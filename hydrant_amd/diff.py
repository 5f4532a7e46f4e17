"""Differentiable volume rendering through the C ABI (hyd_diff_forward / hyd_diff_backward).

``render(volume, tf)`` is a ``torch.autograd.Function``: the forward runs the
HIP forward kernel (keeping transmittance checkpoints every ``ckpt_every``
samples), the backward runs the HIP adjoint kernel, which recomputes each
segment from its checkpoint and returns dL/dvolume and dL/dTF.  torch only
provides device memory and the autograd plumbing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class DiffRenderer:
    """Views, image size and model constants of one differentiable render."""

    def __init__(self, ctx, views: list, W: int, H: int, dims, n_steps: int = 512, ckpt_every: int = 64,
                 sigma: float = 8.0):
        self.ctx = ctx
        self.W, self.H = int(W), int(H)
        self.n_views = len(views)
        self.views = (L.View * len(views))(*views)
        p = L.DiffParams()
        p.dims = L.i3(*[int(d) for d in dims])
        p.n_steps, p.ckpt_every, p.sigma = int(n_steps), int(ckpt_every), float(sigma)
        self.params = p
        nck = C.c_int32(0)
        L.call("hyd_diff_ckpt_floats", C.byref(p), C.byref(nck))
        self.nck = nck.value

    def _check_inputs(self, volume, tf):
        n_vox = int(np.prod([int(d) for d in self.params.dims]))
        if volume.numel() != n_vox or not volume.is_contiguous():
            raise ValueError(f"volume must be a contiguous tensor of {n_vox} voxels "
                             f"(dims {list(self.params.dims)}); got {tuple(volume.shape)}")
        if tf.dim() != 2 or tf.shape[1] != 4 or not 1 <= tf.shape[0] <= 256 or not tf.is_contiguous():
            raise ValueError(f"tf must be a contiguous (n, 4) tensor with 1 <= n <= 256; got {tuple(tf.shape)}")

    def forward_raw(self, volume, tf, count: bool = False):
        self._check_inputs(volume, tf)
        t = self.ctx.torch
        self.ctx.sync_stream()
        rgba = t.empty((self.n_views, self.H, self.W, 4), dtype=t.float32, device=self.ctx.dev)
        ck = t.empty((self.n_views, self.H, self.W, self.nck), dtype=t.float32, device=self.ctx.dev)
        cnt = t.zeros(1, dtype=t.int64, device=self.ctx.dev) if count else None
        L.call("hyd_diff_forward", self.ctx.h, C.byref(self.params), self.views, self.n_views, self.W, self.H,
               C.c_void_p(volume.data_ptr()), C.c_void_p(tf.data_ptr()), tf.shape[0],
               C.c_void_p(rgba.data_ptr()), C.c_void_p(ck.data_ptr()), C.c_void_p(cnt.data_ptr() if count else 0))
        return rgba, ck, cnt

    def backward_raw(self, volume, tf, ck, grad_rgba):
        self._check_inputs(volume, tf)
        if tuple(ck.shape) != (self.n_views, self.H, self.W, self.nck) or not ck.is_contiguous():
            raise ValueError(f"checkpoints must be a contiguous {(self.n_views, self.H, self.W, self.nck)} tensor; "
                             f"got {tuple(ck.shape)}")
        if tuple(grad_rgba.shape) != (self.n_views, self.H, self.W, 4):
            raise ValueError(f"grad_rgba must have shape {(self.n_views, self.H, self.W, 4)}; "
                             f"got {tuple(grad_rgba.shape)}")
        import torch

        if grad_rgba.dtype != torch.float32:
            raise ValueError(f"grad_rgba must be float32 (the adjoint reads float4 pixels); got {grad_rgba.dtype}")
        if any(x.device != grad_rgba.device for x in (volume, tf, ck)):
            raise ValueError("volume, tf, checkpoints and grad_rgba must live on one device; got "
                             f"{volume.device}, {tf.device}, {ck.device}, {grad_rgba.device}")
        t = self.ctx.torch
        self.ctx.sync_stream()
        gv = t.empty_like(volume)   # both outputs are overwritten by the adjoint
        gt = t.empty_like(tf)
        grad_rgba = grad_rgba.contiguous()
        L.call("hyd_diff_backward", self.ctx.h, C.byref(self.params), self.views, self.n_views, self.W, self.H,
               C.c_void_p(volume.data_ptr()), C.c_void_p(tf.data_ptr()), tf.shape[0], C.c_void_p(ck.data_ptr()),
               C.c_void_p(grad_rgba.data_ptr()), C.c_void_p(gv.data_ptr()), C.c_void_p(gt.data_ptr()))
        return gv, gt

    def render(self, volume, tf):
        return _RenderFn.apply(volume, tf, self)


def _make_fn():
    import torch

    class RenderFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, volume, tf, renderer: DiffRenderer):
            if volume.dtype != torch.float32 or tf.dtype != torch.float32:
                raise TypeError("volume and tf must be float32")
            if not (volume.is_cuda and tf.is_cuda):
                raise ValueError("volume and tf must be device tensors (the HIP path has no CPU fallback)")
            volume_c = volume.contiguous()
            tf_c = tf.contiguous()
            rgba, ck, _ = renderer.forward_raw(volume_c, tf_c)
            ctx.save_for_backward(volume_c, tf_c, ck)
            ctx.renderer = renderer
            return rgba

        @staticmethod
        def backward(ctx, grad_rgba):
            volume, tf, ck = ctx.saved_tensors
            gv, gt = ctx.renderer.backward_raw(volume, tf, ck, grad_rgba)
            return gv, gt, None

    return RenderFn


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _make_fn()
        return cls.fn.apply(*args)


_RenderFn = _Lazy


def smooth_tf(n: int, seed: int, n_bumps: int = 4) -> np.ndarray:
    """Seeded TF of smooth bumps: RGB = sum of coloured Gaussian bumps in s, alpha likewise,
    clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    s = (np.arange(n) + 0.5) / n
    tf = np.zeros((n, 4), np.float64)
    for _ in range(n_bumps):
        c = rng.uniform(0.15, 0.95)
        w = rng.uniform(0.04, 0.15)
        col = rng.uniform(0.2, 1.0, 3)
        amp = rng.uniform(0.3, 1.0)
        bump = np.exp(-0.5 * ((s - c) / w) ** 2)
        tf[:, :3] += bump[:, None] * col[None, :]
        tf[:, 3] += amp * bump
    return np.clip(tf, 0.0, 1.0).astype(np.float32)
