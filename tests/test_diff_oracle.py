"""CPU pins of the differentiable-render oracle: its adjoint against FP64 central
finite differences of its own forward, and closed forms of the forward."""
import numpy as np

from oracle import oracle as O


def _setup(seed=0, n=6, W=6, H=5, K=24):
    rng = np.random.default_rng(seed)
    vol = rng.uniform(0.0, 1.0, (n, n, n)).astype(np.float32)
    tf = rng.uniform(0.0, 1.0, (8, 4)).astype(np.float32)
    view = O.make_view(O.orbit_position((30, 40, 2.2), 64), (0, 0, 0), (0, 1, 0), 60.0, (0.5,) * 3, (1,) * 3, 64)
    ys, xs = np.mgrid[0:H, 0:W]
    xy = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    wts = rng.normal(size=(xy.shape[0], 4))
    return vol, tf, view, xy, wts, K


def _loss(vol, tf, view, xy, wts, K, W=6, H=5):
    sc = O.DiffScene(W, H, vol, tf, K=K, sigma=8.0)
    return float(np.sum(sc.render(view, xy, prec=64) * wts))


def test_adjoint_matches_central_differences_fp64():
    vol, tf, view, xy, wts, K = _setup()
    sc = O.DiffScene(6, 5, vol, tf, K=K, sigma=8.0)
    gv, gt = sc.grad(view, xy, wts, prec=64)
    h = 1e-4
    rng = np.random.default_rng(1)
    # TF entries
    for _ in range(12):
        i, ch = int(rng.integers(0, 8)), int(rng.integers(0, 4))
        tp, tm = tf.astype(np.float64).copy(), tf.astype(np.float64).copy()
        tp[i, ch] += h
        tm[i, ch] -= h
        # the oracle reads float32 inputs: evaluate the FD on the float32-representable grid step
        fd = (_loss(vol, tp.astype(np.float32), view, xy, wts, K) - _loss(vol, tm.astype(np.float32), view, xy, wts, K))
        fd /= float(np.float32(tp[i, ch]) - np.float32(tm[i, ch]))
        assert abs(fd - gt[i, ch]) <= 2e-3 * max(1.0, abs(gt[i, ch])), (i, ch, fd, gt[i, ch])
    # voxels with non-zero gradient
    nz = np.argwhere(np.abs(gv) > 1e-3)
    for idx in nz[rng.choice(len(nz), size=min(12, len(nz)), replace=False)]:
        vp, vm = vol.copy(), vol.copy()
        vp[tuple(idx)] += np.float32(h)
        vm[tuple(idx)] -= np.float32(h)
        fd = (_loss(vp, tf, view, xy, wts, K) - _loss(vm, tf, view, xy, wts, K))
        fd /= float(vp[tuple(idx)] - vm[tuple(idx)])
        g = gv[tuple(idx)]
        assert abs(fd - g) <= 2e-3 * max(1.0, abs(g)), (idx, fd, g)


def test_forward_constant_field_closed_form():
    """Constant volume and TF: alpha = 1 - exp(-sigma a dt) per sample, so the ray's opacity is
    1 - exp(-sigma a L) for chord length L, and C = c * A."""
    n = 4
    vol = np.full((n, n, n), 0.5, np.float32)
    tf = np.tile(np.array([0.2, 0.6, 0.9, 0.7], np.float32), (4, 1))
    view = O.make_view((3, 0, 0), (0, 0, 0), (0, 1, 0), 60.0, (0.5,) * 3, (1,) * 3, 64)
    sc = O.DiffScene(2, 2, vol, tf, K=64, sigma=3.0)
    out = sc.render(view, np.array([[1, 1]], np.int32), prec=64)[0]
    c = tf[0].astype(np.float64)      # the stored binary32 TF values
    A = 1 - np.exp(-3.0 * c[3] * 1.0)   # axis ray through the unit box: L = 1
    assert abs(out[3] - A) < 1e-12
    assert np.allclose(out[:3], c[:3] * A, atol=1e-12)


def test_fp32_forward_close_to_fp64():
    vol, tf, view, xy, wts, K = _setup(seed=3)
    sc = O.DiffScene(6, 5, vol, tf, K=K, sigma=8.0)
    assert np.max(np.abs(sc.render(view, xy, prec=32) - sc.render(view, xy, prec=64))) < 1e-5

