"""CPU pins of the oracle's isosurface, Blocks, recolor-mode and empty-skip paths.

Each oracle function is checked against something the mathematics fixes, chosen so
that a plausible mistake (a dropped term, a wrong sign, index or step count) fails:
  * isosurface (isosurface_shader.cu:50-125): an affine SDF, where trilinear
    interpolation is exact, gives the crossing in closed form and bounds the
    bisection error by step / 2^(m+1) after m bisections; an analytic sphere SDF
    gives the hit mask, the depth and the outward normal; the literal rule (its
    prev_value is never updated) hits at the first sample inside the box, at depth
    tnear, with the interpolated point one step before it;
  * Blocks.Volume / Blocks.Solid (blocks_shader.cu:16-50): an axis-aligned row of
    bricks gives 1 - (1 - a)^S per brick, the 0.95 stop and the first non-empty
    brick's colour, including the step lost after each empty-space skip;
  * Partition / Paging recolor (volume_shader.cu:54-66) on a constant field;
  * skip_nblock_steps (volume_shader.cu:5-14) with chebyshev distances 1-3: the
    render with skipping equals the render without it, and the sample count
    follows the skip arithmetic.
"""
import math

import numpy as np
import pytest

from hydrant_amd import dataset as D
from oracle import oracle as O

ARM = (25.0, 65.0, 2.2)
# An orbit camera below and behind the box: its rays enter through the min faces.  A ray
# entering through a max face starts exactly on it, where floor() names the brick past the
# grid; the clamped page lookup then reads the sampler's wrapped value (Wrap addressing),
# which for the tilted plane is a sign change at the first step.
ARM_LOW = (135.0, -35.0, 2.2)


# ---------------------------------------------------------------- helpers --
def _unit_view(prec, arm=ARM):
    pos = O.orbit_position(arm, prec)
    v = O.make_view(pos, (0, 0, 0), (0, 1, 0), 60.0, (0.5, 0.5, 0.5), (1, 1, 1), prec)
    return v, pos


def _to_world(prec=64):
    return O.inverse(O.exhibit_iet((0.5, 0.5, 0.5), (1, 1, 1), prec), prec)


def _iso_scene(sdf, W, H, isovalue, corrected, mode=1, bisect_steps=8, eye=(0, 0, 0)):
    """One n^3 FP32 brick over the unit box (block size n, no apron: k = 1, b = 0)."""
    n = sdf.shape[0]
    one = np.zeros((1, 1, 1), np.int32)
    du = 1.0 / n
    return O.IsoScene(W, H, (0, 0, 0), (1, 1, 1), du / 4, du, _to_world(), eye, (3, 3, 3), (1, 1, 1), isovalue,
                      one, one, np.ones(1), np.zeros((1, 3)), sdf[None].astype(np.float32), n, mode=mode,
                      corrected=corrected, bisect_steps=bisect_steps)


def _rays(view, W, H):
    """FP64 object-space rays of every pixel (pinned against the camera basis elsewhere)."""
    o = np.zeros((H, W, 3))
    d = np.zeros((H, W, 3))
    for y in range(H):
        for x in range(W):
            r = O.gen_ray(view, x, y, W, H, 64)
            o[y, x], d[y, x] = r[:3], r[3:]
    return o, d


def _box_t(o, d):
    """Slab entry/exit of the unit box per ray (FP64)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (0 - o) / d
        t1 = (1 - o) / d
    tn = np.max(np.minimum(t0, t1), axis=-1)
    tf = np.min(np.maximum(t0, t1), axis=-1)
    return tn, tf


def _voxel_centres(n):
    c = (np.arange(n) + 0.5) / n
    return np.meshgrid(c, c, c, indexing="ij")[::-1]     # x, y, z grids of shape (z, y, x)


# --------------------------------------------------- isosurface (A25) -------
N_ISO = 32


def _plane_sdf(g, p0):
    """Affine SDF g . (p - p0) in voxels (block coordinates times n), |g| = 1."""
    x, y, z = _voxel_centres(N_ISO)
    return (g[0] * (x - p0[0]) + g[1] * (y - p0[1]) + g[2] * (z - p0[2])) * N_ISO


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("steps", [8, 4])
def test_iso_plane_bisection_bound(prec, steps):
    """Affine SDF: trilinear is exact away from the faces, so the crossing of each ray is
    t* = g.(p0 - o) / g.d and the corrected rule's depth (the midpoint of the final
    bracket of m bisections of one step) lies within step / 2^(m+1) of it; its normal
    is g oriented toward the side the ray comes from (sprev g)."""
    g = np.array([0.48, 0.6, 0.64])
    p0 = np.array([0.5, 0.45, 0.55])
    W = H = 24
    sc = _iso_scene(_plane_sdf(g, p0), W, H, 0.0, corrected=1, mode=1, bisect_steps=steps)
    view, _ = _unit_view(prec, ARM_LOW)
    r = sc.render(view, prec)
    px = r["px"]
    o, d = _rays(_unit_view(64, ARM_LOW)[0], W, H)
    tstar = ((p0 - o) @ g) / (d @ g)
    pstar = o + d * tstar[..., None]
    tn, tf = _box_t(o, d)
    margin = 1.5 / N_ISO
    inner = np.all((pstar > margin) & (pstar < 1 - margin), axis=-1) & (tstar > tn + 2 * margin) & \
        (tstar < tf - 2 * margin)
    assert inner.sum() > 100
    hit = px[..., 5] == 1
    assert hit[inner].all()
    step = 1.0 / (4 * N_ISO)
    bound = step / 2 ** (steps + 1) + (2e-6 if prec == 32 else 1e-12)
    err = np.abs(px[..., 4] - tstar)[inner]
    assert err.max() <= bound, (err.max(), bound)
    if steps == 4:
        assert err.max() > step / 2 ** 9    # fewer bisections are visible in the error
    sprev = np.sign((o - p0) @ g)[inner]
    assert np.all(sprev == -1)     # this camera sits on the plane's negative side
    nrm = px[..., :3][inner]
    assert np.abs(nrm - sprev[:, None] * g).max() < 2e-6     # voxel values are binary32


def _sphere(centre, r_vox):
    x, y, z = _voxel_centres(N_ISO)
    return np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2) * N_ISO - r_vox


SPHERE_C = np.array([0.52, 0.47, 0.5])
SPHERE_R = 9.0     # voxels


def _sphere_geometry(W, H):
    o, d = _rays(_unit_view(64)[0], W, H)
    oc = SPHERE_C - o
    tc = np.sum(oc * d, axis=-1)
    rho = np.sqrt(np.maximum(np.sum(oc * oc, axis=-1) - tc ** 2, 0))
    return o, d, tc, rho


@pytest.mark.parametrize("prec", [64, 32])
def test_iso_sphere_corrected(prec):
    """Analytic sphere: rays passing nearer the centre than r - 1 voxel hit, rays passing
    farther than r + 1 voxel miss; depth = tc - sqrt(r^2 - rho^2) within the trilinear
    error of |p - c| (1 / (4 r) voxel along the normal, divided by cos of the incidence)
    plus the bisection bound; the normal is the outward radius."""
    W = H = 32
    sc = _iso_scene(_sphere(SPHERE_C, SPHERE_R), W, H, 0.0, corrected=1, mode=1)
    px = sc.render(_unit_view(prec)[0], prec)["px"]
    o, d, tc, rho = _sphere_geometry(W, H)
    r = SPHERE_R / N_ISO
    hit = px[..., 5] == 1
    sure_hit = rho < r - 1.0 / N_ISO
    sure_miss = rho > r + 1.0 / N_ISO
    assert sure_hit.sum() > 50 and sure_miss.sum() > 50
    assert hit[sure_hit].all() and not hit[sure_miss].any()
    steep = rho < 0.8 * r
    cos_i = np.sqrt(1 - (rho[steep] / r) ** 2)
    tstar = tc[steep] - np.sqrt(r ** 2 - rho[steep] ** 2)
    tol = (1 / (4 * SPHERE_R) / cos_i + 1e-3) / N_ISO
    err = np.abs(px[..., 4][steep] - tstar)
    assert np.all(err <= tol), (err / tol).max()
    p = o[steep] + d[steep] * px[..., 4][steep][:, None]
    outward = (p - SPHERE_C) / np.linalg.norm(p - SPHERE_C, axis=-1, keepdims=True)
    assert np.abs(px[..., :3][steep] - outward).max() < 0.02


@pytest.mark.parametrize("prec", [64, 32])
def test_iso_literal_first_sample(prec):
    """Literal rule with iso 0: prev_value stays 0 (isosurface_shader.cu:58), so any
    non-zero first sample is a crossing: every ray that meets the box hits, at depth
    distance(o_clip, origin) = tnear, and the bogus interpolation a = (iso - 0) / (v - 0)
    = 0 puts the shaded point one step before the first sample (Position mode)."""
    W = H = 24
    sdf = _sphere(SPHERE_C, SPHERE_R)
    o, d = _rays(_unit_view(64)[0], W, H)
    tn, tf = _box_t(o, d)
    step = 1.0 / (4 * N_ISO)
    # a ray takes int(chord / step) samples: it needs a chord of one step to sample at all
    inbox = (tf > tn) & ((tf - tn) / step >= 1)
    clear = np.abs((tf - tn) / step - 1) > 1e-3
    r = _iso_scene(sdf, W, H, 0.0, corrected=0, mode=2).render(_unit_view(prec)[0], prec)["px"]
    tol = 1e-12 if prec == 64 else 3e-6
    assert inbox.sum() > 300 and (~inbox).sum() > 10
    assert np.array_equal((r[..., 5] == 1)[clear], inbox[clear])
    assert np.abs(r[..., 4][inbox] - tn[inbox]).max() < tol
    want = o[inbox] + d[inbox] * (tn[inbox] - step)[:, None]
    assert np.abs(r[..., :3][inbox] - want).max() < tol
    assert np.all(np.isinf(r[..., 4][~inbox]))


def test_iso_literal_threshold_is_inner_sphere():
    """Literal rule with iso = -3: sign(v + 3) != sign(3) holds once v <= -3, i.e. at the
    first sample inside the sphere of radius r - 3; the depth is that sample's distance,
    within one step after the analytic entry (plus the trilinear error), and the literal
    normal s(p - 2du) - s(p + 2du) points inward."""
    W = H = 32
    sdf = _sphere(SPHERE_C, SPHERE_R)
    px = _iso_scene(sdf, W, H, -3.0, corrected=0, mode=1).render(_unit_view(64)[0], 64)["px"]
    o, d, tc, rho = _sphere_geometry(W, H)
    r = (SPHERE_R - 3.0) / N_ISO
    hit = px[..., 5] == 1
    assert hit[rho < r - 1.0 / N_ISO].all() and not hit[rho > r + 1.0 / N_ISO].any()
    steep = rho < 0.8 * r
    cos_i = np.sqrt(1 - (rho[steep] / r) ** 2)
    tstar = tc[steep] - np.sqrt(r ** 2 - rho[steep] ** 2)
    slack = (1 / (4 * (SPHERE_R - 3)) / cos_i) / N_ISO
    dt = px[..., 4][steep] - tstar
    step = 1.0 / (4 * N_ISO)
    assert np.all(dt >= -slack) and np.all(dt <= step + slack)
    p = o[steep] + d[steep] * px[..., 4][steep][:, None]
    inward = -(p - SPHERE_C) / np.linalg.norm(p - SPHERE_C, axis=-1, keepdims=True)
    assert np.abs(px[..., :3][steep] - inward).max() < 0.05


# ------------------------------------------------------- blocks (A24) -------
def _row_view():
    """Centre pixel (1, 1) of a 2 x 2 image looks along +x from (-0.5, 1.5, 1.5)."""
    v = O.OView()
    v.trans[:] = [0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, 0, -0.5, 1.5, 1.5, 1]
    v.ctg_fovy_2 = 1.0
    v.ray_o[:] = [-0.5, 1.5, 1.5]
    return v


def _blocks_row(cheb_row, mean_row, mode, density, prec, x_min=0):
    nx = len(cheb_row)
    cheb = np.full((3, 3, nx), 5, np.int32)
    cheb[1, 1, :] = cheb_row
    mean = np.zeros((3, 3, nx), np.float32)
    mean[1, 1, :] = mean_row
    rgba = O.blocks_render(_row_view(), 2, 2, (x_min, 0, 0), (nx, 3, 3), 1.0 / 64, cheb, mean, mode=mode,
                           density=density, prec=prec)
    return rgba[1, 1]


@pytest.mark.parametrize("prec", [32, 64])
def test_blocks_volume_row(prec):
    """Blocks.Volume along an axis ray, step 1/64: 64 samples per non-empty brick, colour
    (rip, 1) * density * mean with rip = (x / nx, 1/3, 1/3); after an empty brick
    (chebyshev 1) the skip lands on the next brick's first sample and the caller's step
    passes it, so that brick takes 63."""
    nx = 4
    mean = [0.8, 0.0, 0.5, 0.6]
    density = 0.01
    got = _blocks_row([0, 1, 0, 0], mean, 0, density, prec)
    rips = [(x / nx, 1 / 3, 1 / 3) for x in range(nx)]
    A, rgb = 0.0, np.zeros(3)
    for S, m, rip in zip([64, 0, 63, 64], mean, rips):
        a = density * m
        An = 1 - (1 - A) * (1 - a) ** S
        rgb += np.asarray(rip) * (An - A)
        A = An
    tol = 1e-12 if prec == 64 else 2e-6
    np.testing.assert_allclose(got, [*rgb, A], atol=tol)


def test_blocks_volume_stops_at_095():
    """Opacity 0.1 per sample: alpha passes 0.95 at sample ceil(log 0.05 / log 0.9) = 29,
    which ends the ray inside the first brick (later bricks add nothing)."""
    got = _blocks_row([0, 0, 0], [1.0, 1.0, 1.0], 0, 0.1, 64)
    S = math.ceil(math.log(0.05) / math.log(0.9))
    A = 1 - 0.9 ** S
    np.testing.assert_allclose(got, [0.0, A / 3, A / 3, A], atol=1e-12)


@pytest.mark.parametrize("prec", [32, 64])
def test_blocks_solid_first_brick(prec):
    """Blocks.Solid: the first brick with chebyshev 0 gives (rip, 1)."""
    np.testing.assert_allclose(_blocks_row([2, 1, 0, 0], [0] * 4, 1, 1.0, prec), [2 / 4, 1 / 3, 1 / 3, 1.0],
                               atol=1e-7)
    np.testing.assert_allclose(_blocks_row([0, 1, 0], [0] * 3, 1, 1.0, prec), [0.0, 1 / 3, 1 / 3, 1.0], atol=1e-7)
    np.testing.assert_allclose(_blocks_row([1, 1, 1], [0] * 3, 1, 1.0, prec), [0, 0, 0, 0], atol=0)
    # a box starting at x = 1: rip = clamp(ip, min, max - 1) / (max - min) = (1 / 3, 1/3, 1/3)
    np.testing.assert_allclose(_blocks_row([0, 0, 0, 0], [0] * 4, 1, 1.0, prec, x_min=1), [1 / 3, 1 / 3, 1 / 3, 1],
                               atol=1e-7)


# ------------------------------------------------ recolor modes (A15) -------
def _const_axis_scene(mode, rank, lowest, density, prec):
    n = 16
    pool = np.full((1, n, n, n), np.float32(0.4), np.float32)
    tf = np.tile(np.asarray([0.3, 0.6, 0.2, 0.5], np.float32), (4, 1))
    du = 1.0 / n
    step = du / 4 / 0.5
    v = O.make_view((3, 0, 0), (0, 0, 0), (0, 1, 0), 60.0, (0.5, 0.5, 0.5), (1, 1, 1), prec)
    sc = O.VolumeScene(2, 2, (0, 0, 0), (1, 1, 1), step, du, tf, np.zeros((1, 1, 1)), np.zeros((1, 1, 1)),
                       np.ones(1), np.zeros((1, 3)), pool, n, mode=mode, rank=rank, density=density,
                       lowest_blkcnt=lowest)
    r = sc.render(v, prec, film=False, counts=True)
    return r["rgba"][1, 1], int(r["counts"][1, 1])


@pytest.mark.parametrize("prec", [32, 64])
def test_partition_recolor(prec):
    """Partition: rgb = mix((1,0,0), (0,0,1), rank) * |tf.rgb| before premultiplying, so
    v.rgb = ((1 - rank) |c|, 0, rank |c|) * v.w and v.w = 1 - (1 - a)^S."""
    c = np.float32([0.3, 0.6, 0.2]).astype(np.float64)    # the TF table is binary32
    rank = 0.3
    got, S = _const_axis_scene(1, rank, 0, 0.02, prec)
    a = 0.5 * 0.02
    A = 1 - (1 - a) ** S
    want = np.array([(1 - rank) * np.linalg.norm(c) * A, 0.0, rank * np.linalg.norm(c) * A, A])
    np.testing.assert_allclose(got, want, atol=1e-12 if prec == 64 else 3e-6)
    base, S0 = _const_axis_scene(0, rank, 0, 0.02, prec)
    assert S0 == S and abs(base[3] - A) < 3e-6


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("lowest,colour", [(0, (0, 1, 0)), (1, (1, 0, 0))])
def test_paging_recolor(prec, lowest, colour):
    """Paging: page 0 is green when 0 >= lowest_blkcnt, else red."""
    got, S = _const_axis_scene(2, 0.0, lowest, 0.02, prec)
    a = 0.5 * 0.02
    A = 1 - (1 - a) ** S
    np.testing.assert_allclose(got, [*(np.asarray(colour) * A), A], atol=1e-12 if prec == 64 else 3e-6)


# --------------------------------------------------- empty skip (A8) --------
def _cheb_brute(nonempty):
    """L-inf brick distance to the nearest non-empty brick (0 for non-empty)."""
    idx = np.argwhere(nonempty)
    out = np.zeros(nonempty.shape, np.int32)
    for p in np.ndindex(nonempty.shape):
        out[p] = int(np.min(np.max(np.abs(idx - np.array(p)), axis=1)))
    return out


def _shell_volume(nonempty, bs, value=200, seed=0):
    """u8 volume whose non-empty bricks hold values in their interior and zero in a one-voxel
    shell; empty bricks are zero.  Every sample within half a voxel of a face shared with
    an empty brick then reads eight zero taps."""
    gz, gy, gx = nonempty.shape
    vol = np.zeros((gz * bs, gy * bs, gx * bs), np.uint8)
    rng = np.random.default_rng(seed)
    for z, y, x in np.argwhere(nonempty):
        blk = rng.integers(value // 2, value, size=(bs - 2,) * 3)
        vol[z * bs + 1:(z + 1) * bs - 1, y * bs + 1:(y + 1) * bs - 1, x * bs + 1:(x + 1) * bs - 1] = blk
    return vol


def _volume_scene(vol, bs, cheb, W, H, step, density, tf=D.README_TF):
    grid = tuple(s // bs for s in vol.shape)
    pool = D.pack(vol, bs, 1, 2)               # every brick, FP32 v / 255
    k, b = D.mapping(bs, 1)
    n = pool.shape[0]
    vaddr = np.arange(n, dtype=np.int32).reshape(grid)
    return O.VolumeScene(W, H, (0, 0, 0), grid[::-1], step, 1.0 / bs, tf, cheb, vaddr,
                         np.full(n, k), np.full((n, 3), b), pool, bs + 2, density=density)


@pytest.mark.parametrize("prec", [32, 64])
def test_skip_row_closed_form(prec):
    """Axis ray along a row whose chebyshev distances are 0,1,2,3,3,2,1,0 (non-empty bricks at
    both ends), step 1/64: brick 0 takes 64 samples; each skip moves to
    ceil((tfar + (cd - 1) cdu) / step) steps (cdu = 1 on an axis ray), landing on x = 2, 4
    and 7, and the caller's step passes the landing sample, so the last brick takes 63:
    127 samples.  The image equals the render without skipping, where the 384 samples of
    the empty bricks add nothing."""
    nonempty = np.zeros((3, 3, 8), bool)
    nonempty[1, 1, 0] = nonempty[1, 1, 7] = True
    cheb = _cheb_brute(nonempty)
    assert cheb[1, 1].tolist() == [0, 1, 2, 3, 3, 2, 1, 0]
    bs = 8
    vol = _shell_volume(nonempty, bs)
    # density 0.005 keeps alpha below the 0.85 escalation threshold over both bricks
    on = _volume_scene(vol, bs, cheb, 2, 2, 1.0 / 64, 0.005).render(_row_view(), prec, film=False, counts=True)
    off = _volume_scene(vol, bs, np.zeros_like(cheb), 2, 2, 1.0 / 64, 0.005).render(_row_view(), prec, film=False,
                                                                                      counts=True)
    assert on["counts"][1, 1] == 127 and off["counts"][1, 1] == 512
    assert on["rgba"][1, 1, 3] > 0.05
    assert np.array_equal(on["rgba"][1, 1], off["rgba"][1, 1])


def test_skip_equals_no_skip_random_views():
    """Clustered non-empty bricks (chebyshev up to 4) seen from several orbit views: the
    FP64 render with empty-space skipping equals the one without it to rounding, while
    composing far fewer samples.  Skipping one brick too far (cd instead of cd - 1 extra
    bricks) would drop samples of non-empty interiors."""
    rng = np.random.default_rng(4)
    g = 8
    nonempty = np.zeros((g, g, g), bool)
    for c in ([1, 1, 1], [6, 5, 6], [2, 6, 5]):
        nonempty[tuple(np.clip(np.array(c) + rng.integers(-1, 2, size=(4, 3)), 0, g - 1).T)] = True
    cheb = _cheb_brute(nonempty)
    assert cheb.max() >= 3
    bs = 8
    vol = _shell_volume(nonempty, bs, seed=5)
    W = H = 24
    step = (1.0 / bs) / 4 / 0.5
    for arm in [(25, 65, 2.2), (140, -30, 2.0), (300, 10, 2.5)]:
        pos = O.orbit_position(arm, 64)
        v = O.make_view(pos, (0, 0, 0), (0, 1, 0), 60.0, (g / 2,) * 3, (g,) * 3, 64)
        on = _volume_scene(vol, bs, cheb, W, H, step, 0.05).render(v, 64, film=False)
        off = _volume_scene(vol, bs, np.zeros_like(cheb), W, H, step, 0.05).render(v, 64, film=False)
        assert off["rgba"][..., 3].max() > 0.1
        assert np.abs(on["rgba"] - off["rgba"]).max() < 1e-9
        assert on["samples"] < 0.6 * off["samples"]


@pytest.mark.parametrize("prec", [32, 64])
def test_skip_after_escalation_closed_form(prec):
    """Escalation then skipping, on the chebyshev row 0,1,2,3,3,2,1,0 with a constant TF of
    opacity 0.1 per sample (step 1/64, budget 512):
      * brick 0: 19 samples at step 1 (1 - 0.9^19 > 0.85), then step 4 for 7 samples
        (1 - 0.9^26 > 0.93), then step 16: one more sample at x = 62/64;
      * skips with sstep = 16/64 from x = 78/64, 158/64 and 286/64 advance 4, 7 and 11
        sstep (ceil((tfar + (cd - 1)) / sstep)) and the caller adds one more, landing on
        x = 7 + 30/64: 3 samples in brick 7;
      * the skips subtract 4, 7 and 11 from the step budget, not 16 times that
        (volume_shader.cu:47 returns big steps), so 316 steps remain at the box exit and
        the ray samples on outside it, through the clamped brick lookup, for 20 more
        samples: 50 in all, ending at x = 13 + 14/64 with the budget at -4."""
    nonempty = np.zeros((3, 3, 8), bool)
    nonempty[1, 1, 0] = nonempty[1, 1, 7] = True
    cheb = _cheb_brute(nonempty)
    bs = 8
    vol = _shell_volume(nonempty, bs)
    tf = np.tile(np.float32([0.5, 0.5, 0.5, 1.0]), (9, 1))
    sc = _volume_scene(vol, bs, cheb, 2, 2, 1.0 / 64, 0.1, tf=tf)
    r = sc.render(_row_view(), prec, film=True, counts=True)
    assert r["counts"][1, 1] == 50
    film = r["film"][1, 1]
    assert film[0] == 13 + 14 / 64 and film[6] == -4 and film[7] == 16
    assert abs(film[11] - (1 - 0.9 ** 50)) < (1e-12 if prec == 64 else 2e-6)
