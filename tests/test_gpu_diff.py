"""GPU tests of the differentiable render (forward + adjoint through the C ABI).

Bars: forward within 1e-4 of the FP64 oracle; adjoint (dL/dTF, dL/dvolume) within
1e-3 relative of the FP64 oracle gradient on a 64x64 crop of 2 views; central
finite differences (h = 1e-4) of the FP64 forward with FP64 reductions agree with
the GPU adjoint on 64 random parameters; a short Adam fit through the autograd Function
lowers the loss.
"""
import numpy as np
import pytest

from hydrant_amd import workloads as W
from hydrant_amd.diff import DiffRenderer
from hydrant_amd.hip import view_to_dict
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _renderer(ctx, wl, n_views=None):
    views = wl.views()[: n_views or len(wl.arms)]
    return DiffRenderer(ctx, views, wl.W, wl.H, (wl.n,) * 3, wl.n_steps, wl.ckpt_every, wl.sigma), views


def _oview(v):
    return O.view_from_dict(view_to_dict(v))


def test_forward_matches_fp64(gpu_ctx):
    import torch

    wl = W.differentiable(n=64, n_views=3, W=96, H=96)
    r, views = _renderer(gpu_ctx, wl)
    vol = torch.from_numpy(wl.hidden_volume).cuda()
    tf = torch.from_numpy(wl.hidden_tf).cuda()
    rgba, _, cnt = r.forward_raw(vol, tf, count=True)
    torch.cuda.synchronize()
    g = rgba.cpu().numpy()
    sc = O.DiffScene(wl.W, wl.H, wl.hidden_volume, wl.hidden_tf, K=wl.n_steps, sigma=wl.sigma)
    ys, xs = np.mgrid[0:wl.H, 0:wl.W]
    xy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    err = 0.0
    for i, v in enumerate(views):
        ref = sc.render(_oview(v), xy, prec=64).reshape(wl.H, wl.W, 4)
        err = max(err, float(np.max(np.abs(g[i] - ref))))
    print(f"diff forward: max-abs vs FP64 {err:.3g}, samples {int(cnt.item())}", flush=True)
    assert err <= 1e-4
    assert g[..., 3].max() > 0.3


def test_adjoint_matches_fp64_gradient(gpu_ctx):
    import torch

    wl = W.differentiable(n=64, n_views=2, W=128, H=128)
    r, views = _renderer(gpu_ctx, wl)
    vol = torch.from_numpy(wl.hidden_volume).cuda()
    tf = torch.from_numpy(wl.hidden_tf).cuda()
    rgba, ck, _ = r.forward_raw(vol, tf)
    rng = np.random.default_rng(11)
    grad = np.zeros((2, wl.H, wl.W, 4), np.float32)
    y0, x0 = 32, 32
    grad[:, y0:y0 + 64, x0:x0 + 64] = rng.normal(size=(2, 64, 64, 4)).astype(np.float32)
    gv, gt = r.backward_raw(vol, tf, ck, torch.from_numpy(grad).cuda())
    torch.cuda.synchronize()
    sc = O.DiffScene(wl.W, wl.H, wl.hidden_volume, wl.hidden_tf, K=wl.n_steps, sigma=wl.sigma)
    ys, xs = np.mgrid[y0:y0 + 64, x0:x0 + 64]
    xy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    ref_v = np.zeros(wl.hidden_volume.shape)
    ref_t = np.zeros((256, 4))
    for i, v in enumerate(views):
        a, b = sc.grad(_oview(v), xy, grad[i, y0:y0 + 64, x0:x0 + 64].reshape(-1, 4), prec=64)
        ref_v += a
        ref_t += b
    ev = float(np.max(np.abs(gv.cpu().numpy() - ref_v)) / np.max(np.abs(ref_v)))
    et = float(np.max(np.abs(gt.cpu().numpy() - ref_t)) / np.max(np.abs(ref_t)))
    # each voxel's own relative error, down to gradients 1e-4 of the largest (an
    # optimiser normalises every parameter's gradient, so small ones must be right too)
    g = gv.cpu().numpy().astype(np.float64)
    big = np.abs(ref_v) >= 1e-4 * np.max(np.abs(ref_v))
    rel = np.abs(g[big] - ref_v[big]) / np.abs(ref_v[big])
    print(f"adjoint vs FP64 gradient: volume rel {ev:.3g}, tf rel {et:.3g}; per-voxel rel p99 "
          f"{np.quantile(rel, 0.99):.3g}, max {rel.max():.3g} over {int(big.sum())} voxels", flush=True)
    assert ev <= 1e-3 and et <= 1e-3
    assert np.quantile(rel, 0.99) <= 1e-3


@pytest.mark.parametrize("n_steps", [None, 98])
def test_adjoint_ragged_sizes(gpu_ctx, n_steps):
    """Image and volume sizes that are not multiples of the 8x8 ray tiles and 8^3
    gradient bricks (partial tiles and bricks), gradient on every pixel, two calls in a
    row (the outputs are overwritten, not accumulated); also a sample count that is not
    a multiple of 4 or 8 (the element-wise paths of the dL/ds stores and loads) or of the
    checkpoint interval."""
    import torch

    wl = W.differentiable(n=44, n_views=2, W=60, H=37, n_steps=n_steps or 512)
    vol_np = wl.hidden_volume[:, :, :41].copy()          # dims (x, y, z) = (41, 44, 44)
    r = DiffRenderer(gpu_ctx, wl.views()[:2], wl.W, wl.H, vol_np.shape[::-1], wl.n_steps, wl.ckpt_every, wl.sigma)
    vol = torch.from_numpy(vol_np).cuda()
    tf = torch.from_numpy(wl.hidden_tf).cuda()
    rng = np.random.default_rng(3)
    grad = rng.normal(size=(2, wl.H, wl.W, 4)).astype(np.float32)
    _, ck, _ = r.forward_raw(vol, tf)
    gv1, gt1 = r.backward_raw(vol, tf, ck, torch.from_numpy(grad).cuda())
    gv, gt = r.backward_raw(vol, tf, ck, torch.from_numpy(grad).cuda())
    torch.cuda.synchronize()
    assert torch.equal(gv, gv1), "voxel gradients are deterministic"
    assert torch.equal(gt, gt1), "TF gradients are deterministic"
    sc = O.DiffScene(wl.W, wl.H, vol_np, wl.hidden_tf, K=wl.n_steps, sigma=wl.sigma)
    ys, xs = np.mgrid[0:wl.H, 0:wl.W]
    xy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    ref_v = np.zeros(vol_np.shape)
    ref_t = np.zeros((256, 4))
    for i, v in enumerate(wl.views()[:2]):
        a, b = sc.grad(_oview(v), xy, grad[i].reshape(-1, 4), prec=64)
        ref_v += a
        ref_t += b
    ev = float(np.max(np.abs(gv.cpu().numpy() - ref_v)) / np.max(np.abs(ref_v)))
    et = float(np.max(np.abs(gt.cpu().numpy() - ref_t)) / np.max(np.abs(ref_t)))
    print(f"ragged adjoint vs FP64 gradient: volume rel {ev:.3g}, tf rel {et:.3g}", flush=True)
    assert ev <= 1e-3 and et <= 1e-3


def test_finite_differences_fp64_reductions(gpu_ctx):
    """GPU adjoint of L = sum(w * render) over a 24x24 crop of 2 views against central
    differences (h = 1e-4) of the FP64 forward with FP64 reductions, on 32 TF entries and 32
    voxels.  The step is taken on the binary32 grid the inputs live on, and the difference
    quotient divides by the step actually taken."""
    import torch

    wl = W.differentiable(n=48, n_views=2, W=64, H=64)
    r, views = _renderer(gpu_ctx, wl)
    vol = torch.from_numpy(wl.hidden_volume).cuda()
    tf = torch.from_numpy(wl.hidden_tf).cuda()
    rng = np.random.default_rng(5)
    y0, x0, c = 20, 20, 24
    wts = np.zeros((2, wl.H, wl.W, 4), np.float32)
    wts[:, y0:y0 + c, x0:x0 + c] = rng.normal(size=(2, c, c, 4)).astype(np.float32)
    _, ck, _ = r.forward_raw(vol, tf)
    gv, gt = r.backward_raw(vol, tf, ck, torch.from_numpy(wts).cuda())
    gv, gt = gv.cpu().numpy(), gt.cpu().numpy()

    sc = O.DiffScene(wl.W, wl.H, wl.hidden_volume, wl.hidden_tf, K=wl.n_steps, sigma=wl.sigma)
    ys, xs = np.mgrid[y0:y0 + c, x0:x0 + c]
    xy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    oviews = [_oview(v) for v in views]
    wcrop = [wts[i, y0:y0 + c, x0:x0 + c].reshape(-1, 4).astype(np.float64) for i in range(2)]

    def loss():
        return sum(float(np.sum(sc.render(ov, xy, prec=64) * w)) for ov, w in zip(oviews, wcrop))

    h = np.float32(1e-4)
    cand = np.argwhere(np.abs(gt) > 1e-3)
    params = [("tf", tuple(int(v) for v in cand[j])) for j in rng.choice(len(cand), min(32, len(cand)), replace=False)]
    flat = np.argsort(np.abs(gv).ravel())[::-1][:1000]
    params += [("vol", tuple(int(v) for v in np.unravel_index(int(j), gv.shape)))
               for j in rng.choice(flat, 32, replace=False)]
    assert len(params) >= 48
    worst = 0.0
    for kind, idx in params:
        arr = sc.tf.reshape(-1, 4) if kind == "tf" else sc.vol
        orig = arr[idx]
        arr[idx] = orig + h
        p = float(arr[idx])
        lp = loss()
        arr[idx] = orig - h
        m = float(arr[idx])
        lm = loss()
        arr[idx] = orig
        fd = (lp - lm) / (p - m)
        ad = float(gt[idx] if kind == "tf" else gv[idx])
        rel = abs(fd - ad) / max(abs(ad), 1e-2)
        worst = max(worst, rel)
        assert rel <= 1e-2, (kind, idx, fd, ad)
    print(f"finite differences (FP64): {len(params)} parameters, worst relative deviation {worst:.3g}", flush=True)


def test_autograd_adam_fit_reduces_loss(gpu_ctx):
    import torch

    wl = W.differentiable(n=48, n_views=6, W=64, H=64)
    r, _ = _renderer(gpu_ctx, wl)
    with torch.no_grad():
        target = r.render(torch.from_numpy(wl.hidden_volume).cuda(), torch.from_numpy(wl.hidden_tf).cuda())
    vol = torch.from_numpy(wl.init_volume).cuda().requires_grad_(True)
    tf = torch.from_numpy(wl.init_tf).cuda().requires_grad_(True)
    opt = torch.optim.Adam([vol, tf], lr=1e-2)
    losses = []
    for _ in range(40):
        opt.zero_grad()
        out = r.render(vol, tf)
        loss = torch.mean((out - target) ** 2)
        loss.backward()
        opt.step()
        losses.append(float(loss.item()))
    print(f"adam: loss {losses[0]:.4g} -> {losses[-1]:.4g}", flush=True)
    assert losses[-1] < 0.25 * losses[0]
    assert vol.grad is not None and torch.isfinite(vol.grad).all()
