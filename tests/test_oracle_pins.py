"""CPU pins of the oracle against closed forms, brute force and invariants.

The reference ships no tests or golden vectors for the render path, so each
oracle function is pinned here by something a plausible mistake (dropped
term, wrong sign or index, transposed operand) would fail.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _rot(axis, deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


# ------------------------------------------------------------- camera (A2) --
@pytest.mark.parametrize("arm", [(25, 65, 2.2), (0, 0, 3.0), (130, -40, 1.5), (-75, 10, 4.0)])
def test_orbit_position_closed_form(arm):
    """position = R_y(arm.x) R_z(arm.y) (arm.z, 0, 0) (scene.hpp:68-81)."""
    ref = _rot("y", arm[0]) @ _rot("z", arm[1]) @ np.array([arm[2], 0, 0])
    assert np.allclose(O.orbit_position(arm, 64), ref, atol=1e-13)
    assert np.allclose(O.orbit_position(arm, 32), ref, atol=2e-6)


def test_readme_orbit_value():
    p = O.orbit_position((25, 65, 2.2), 64)
    assert np.allclose(p, [0.842649, 1.993877, -0.392934], atol=1e-6)


def test_inverse_is_inverse():
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = rng.normal(size=(4, 4)) + 3 * np.eye(4)
        inv = O.inverse(m.T.reshape(-1), 64).reshape(4, 4).T   # column-major in/out
        assert np.allclose(m @ inv, np.eye(4), atol=1e-10)


def test_exhibit_iet():
    """scale = compMax(max(|c - s|, |c|)), translation = c (scene.hpp:20-29)."""
    c, s = np.array([8.0, 4.0, 2.0]), np.array([16.0, 8.0, 4.0])
    m = O.exhibit_iet(c, s, 64).reshape(4, 4).T
    assert np.allclose(m[:3, :3], 8.0 * np.eye(3)) and np.allclose(m[:3, 3], c)


# ------------------------------------------------------------ rays (A4-A6) --
def _view(prec=32, arm=(25, 65, 2.2), fdim=1.0):
    pos = O.orbit_position(arm, prec)
    c = np.full(3, fdim / 2)
    return O.make_view(pos, (0, 0, 0), (0, 1, 0), 60.0, c, np.full(3, fdim), prec), pos


@pytest.mark.parametrize("prec,tol", [(64, 1e-12), (32, 2e-6)])
def test_centre_pixel_looks_at_target(prec, tol):
    """uv = (0, 0) at pixel (W/2, H/2): the ray is the optical axis, normalize(target - eye)."""
    v, pos = _view(prec)
    r = O.gen_ray(v, 256, 256, 512, 512, prec)
    want = -pos / np.linalg.norm(pos)
    assert np.allclose(r[3:], want, atol=tol)
    # origin = camera position mapped to object space (scale 0.5, centre 0.5)
    assert np.allclose(r[:3], pos * 0.5 + 0.5, atol=tol)


def test_pixel_corners_and_aspect():
    """uv = ((x, y) - res/2) * 2 / res.y with integer corners and y flipped (cpu_shading.cc:21-22)."""
    v, pos = _view(64)
    # build the camera basis in world space and compare a corner ray's direction
    f = -pos / np.linalg.norm(pos)
    s = np.cross(f, [0, 1, 0]); s /= np.linalg.norm(s)
    u = np.cross(s, f)
    ctg = 1 / math.tan(math.radians(30))
    W, H = 640, 480
    for (x, y) in [(0, 0), (639, 479), (100, 400)]:
        ux, uy = (x - W / 2) * 2 / H, (y - H / 2) * 2 / H
        want = ux * s - uy * u + ctg * f
        want /= np.linalg.norm(want)
        r = O.gen_ray(v, x, y, W, H, 64)
        assert np.allclose(r[3:], want, atol=1e-12)


def _slab_ref(o, d, bmin, bmax):
    tn, tf = -np.inf, np.inf
    for i in range(3):
        if d[i] == 0:
            if o[i] < bmin[i] or o[i] > bmax[i]:
                return False, None, None
            continue
        t1, t2 = (bmin[i] - o[i]) / d[i], (bmax[i] - o[i]) / d[i]
        tn, tf = max(tn, min(t1, t2)), min(tf, max(t1, t2))
    return tf > tn, tn, tf


def test_intersect_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(500):
        o = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        bmin = rng.uniform(-1, 0, 3); bmax = bmin + rng.uniform(0.1, 2, 3)
        hit, tn, tf = O.intersect(o, d, bmin, bmax, 64)
        rh, rn, rf = _slab_ref(o, d, bmin, bmax)
        assert hit == rh
        if hit:
            assert abs(tn - rn) < 1e-12 and abs(tf - rf) < 1e-12


def test_intersect_axis_aligned_inf():
    """Zero direction components give +-inf slabs, not NaN (glm_math.hpp:61-71)."""
    hit, tn, tf = O.intersect([2.0, 0.5, 0.5], [-1.0, 0.0, 0.0], [0, 0, 0], [1, 1, 1], 32)
    assert hit and tn == 1.0 and tf == 2.0
    hit, _, _ = O.intersect([2.0, 1.5, 0.5], [-1.0, 0.0, 0.0], [0, 0, 0], [1, 1, 1], 32)
    assert not hit


# ---------------------------------------------------------- sampling (A12) --
def test_trilinear_exact_on_affine_field():
    """On f = a.x + b.y + c.z + e trilinear interpolation is exact: at fx = n*u - 0.5
    (normalised coordinates, Linear) the sample equals f(fx) (cpu_sampler.hpp:88-104,142-166)."""
    n = 8
    a, b, c, e = 0.25, -0.5, 0.125, 1.0
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    data = (a * x + b * y + c * z + e).astype(np.float32)
    rng = np.random.default_rng(2)
    for _ in range(200):
        u = rng.uniform(0.6 / n, (n - 0.6) / n, 3)   # interior: no clamping
        fx = n * u - 0.5
        want = a * fx[0] + b * fx[1] + c * fx[2] + e
        got, taps = O.brick_sample(data, u, 64)
        assert abs(got - want) < 1e-12
        assert list(taps[:3]) == list(np.floor(fx).astype(int))
        assert list(taps[3:]) == list(np.ceil(fx).astype(int))
        got32, _ = O.brick_sample(data, u, 32)
        assert abs(got32 - want) < 2e-6


def test_trilinear_brute_force_with_clamp():
    """FP64 brute force on a random 4^3 brick including clamped border taps."""
    rng = np.random.default_rng(3)
    n = 4
    data = rng.random((n, n, n)).astype(np.float32)

    def ref(u):
        fx = n * np.asarray(u) - 0.5
        lo = np.floor(fx)
        w = fx - lo
        i = np.clip(lo.astype(int), 0, n - 1)
        j = np.clip(np.ceil(fx).astype(int), 0, n - 1)
        acc = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    ix = (i[0], j[0])[dx]; iy = (i[1], j[1])[dy]; iz = (i[2], j[2])[dz]
                    wt = (w[0] if dx else 1 - w[0]) * (w[1] if dy else 1 - w[1]) * (w[2] if dz else 1 - w[2])
                    acc += wt * float(data[iz, iy, ix])
        return acc

    for _ in range(300):
        u = rng.uniform(0.0, 0.999, 3)
        got, _ = O.brick_sample(data, u, 64)
        assert abs(got - ref(u)) < 1e-12


def test_wrap_mod_at_brick_edge():
    """Wrap: fx = mod(n*u, n); u = 1 wraps to 0 (the reference's Wrap clamps taps, Q11)."""
    n = 4
    data = np.arange(n ** 3, dtype=np.float32).reshape(n, n, n)
    a, _ = O.brick_sample(data, [1.0, 0.3, 0.3], 64)
    b, _ = O.brick_sample(data, [0.0, 0.3, 0.3], 64)
    assert a == b


# ----------------------------------------------------- transfer fn (A14) ----
def test_tf_knots_border_and_linearity():
    """Knots at (i + 0.5)/N hold entry i; 0 outside [0, 1] (Border); clamp to edge inside;
    linear between knots (transfer_fn.hpp:14-28)."""
    rng = np.random.default_rng(4)
    N = 9
    tf = rng.random((N, 4)).astype(np.float32)
    for i in range(N):
        assert np.allclose(O.tf_sample(tf, (i + 0.5) / N, 64), tf[i], atol=1e-15)
    assert np.all(O.tf_sample(tf, -1e-3, 32) == 0) and np.all(O.tf_sample(tf, 1.0001, 32) == 0)
    assert np.allclose(O.tf_sample(tf, 0.2 / N, 64), tf[0])            # clamp below knot 0
    assert np.allclose(O.tf_sample(tf, (N - 0.2) / N, 64), tf[N - 1])  # clamp above the last knot
    for i in range(N - 1):
        mid = O.tf_sample(tf, (i + 1.0) / N, 64)
        assert np.allclose(mid, 0.5 * (tf[i] + tf[i + 1]), atol=1e-12)
    assert np.allclose(O.tf_sample(tf, 1.0, 64), tf[N - 1]) and np.allclose(O.tf_sample(tf, 0.0, 64), tf[0])


def test_escalation_threshold_float_equivalence():
    """`v.w > 0.93` / `> 0.85` compare binary32 against double literals; the kernels use
    `v.w >= 0.93f` / `>= 0.85f`.  Exhaustive over every float within 2^16 ulps."""
    for t in (0.93, 0.85):
        c = np.float32(t)
        bits = c.view(np.int32) + np.arange(-(1 << 16), 1 << 16, dtype=np.int32)
        f = bits.view(np.float32)
        assert np.array_equal(f.astype(np.float64) > t, f >= c)


def test_u8_normalisation_table():
    """float(v) / 255.f is what Texture3D<u8> NormalizedFloat stores (texture_3d.hpp:78-86,
    glm_math.hpp:83-89); multiplying by 1/255 differs for some v, so the kernels use a table."""
    v = np.arange(256, dtype=np.float32)
    ref = v / np.float32(255)
    assert np.array_equal(ref, np.array([np.float32(i) / np.float32(255) for i in range(256)]))
    assert not np.array_equal(ref, v * (np.float32(1) / np.float32(255)))


# ----------------------------------------------- compositing / schedule (A7) --
def _axis_scene(n_vox, value_u8, tf_row, density, W=1, H=1):
    """Constant volume seen along -x: the centre pixel of a camera at world (3, 0, 0)
    looking at the origin is an exactly axis-aligned ray through the unit bbox."""
    pool = np.full((1, n_vox, n_vox, n_vox), np.float32(value_u8) / np.float32(255), np.float32)
    tf = np.tile(np.asarray(tf_row, np.float32), (4, 1))
    du = np.float32(1) / np.float32(n_vox)
    step = (du / np.float32(4)) / np.float32(0.5)
    v = O.make_view((3, 0, 0), (0, 0, 0), (0, 1, 0), 60.0, (0.5, 0.5, 0.5), (1, 1, 1), 32)
    sc = O.VolumeScene(2, 2, (0, 0, 0), (1, 1, 1), step, du, tf, np.zeros((1, 1, 1)), np.zeros((1, 1, 1)),
                       np.ones(1), np.zeros((1, 3)), pool, n_vox, density=density)
    return sc, v, step


@pytest.mark.parametrize("alpha,density", [(1.0, 0.01), (0.5, 0.2), (1.0, 1.0)])
def test_constant_field_closed_form(alpha, density):
    """Constant TF (c, a): after S samples v.w = 1 - (1 - a)^S, v.rgb = c * v.w, theta = v.rgb,
    phi = 1 - v.w; S follows from the escalation schedule (step 1 while v.w <= 0.85, 4 while
    <= 0.93, then 16) and the integer budget nsteps = int(chord / step)."""
    c = np.array([0.7, 0.2, 0.4])
    sc, v, step = _axis_scene(32, 100, [*c, alpha], density)
    r = sc.render(v, 64, film=True)
    px = r["film"][1, 1]
    d = px[3:6]
    assert np.allclose(d, [-1, 0, 0], atol=0)
    a = alpha * density
    budget = int(1.0 / float(step))
    S, used, vw = 0, 0, 0.0
    while used < budget:
        S += 1
        vw = 1 - (1 - a) ** S
        used += 16 if vw > 0.93 else (4 if vw > 0.85 else 1)
    assert r["counts"] is None or True
    assert abs(px[11] - (1 - (1 - a) ** S)) < 1e-12
    assert np.allclose(px[8:11], c * px[11], atol=1e-12)
    assert np.allclose(px[12:15], px[8:11], atol=1e-12) and abs(px[15] - (1 - px[11])) < 1e-12
    r2 = sc.render(v, 64, film=False, counts=True)
    assert r2["counts"][1, 1] == S


def test_density_one_is_reference_literal():
    """density 1 leaves the TF alpha untouched (x * 1 == x): same film as a TF pre-scaled by 1."""
    sc1, v, _ = _axis_scene(16, 200, [0.3, 0.6, 0.9, 0.05], 1.0)
    r = sc1.render(v, 32, film=True)
    a = np.float32(0.05)
    S = r["film"][1, 1, 6]  # nsteps after the loop is <= 0
    assert r["rgba"][1, 1, 3] > 0 and S <= 0


def test_fp32_vs_fp64_reference_schedule_is_discontinuous():
    """The reference schedule takes discrete decisions from float values (nsteps =
    int(chord / step), escalation thresholds), so FP32 and FP64 can differ by one whole
    sample on a few rays: bounded by the per-sample opacity (density * max TF alpha),
    while almost every pixel agrees to 1e-5."""
    from hydrant_amd import workloads as W
    from tests.helpers import oracle_view, oracle_volume

    sc = W.plain_volume(n=32, W=32, H=32)
    osc = oracle_volume(sc)
    ov = oracle_view(sc)
    r32 = osc.render(ov, 32, film=False)
    r64 = osc.render(ov, 64, film=False)
    diff = np.abs(r32["rgba"] - r64["rgba"]).max(axis=-1)
    assert diff.max() <= 1.01 * sc.density
    assert np.mean(diff < 1e-5) > 0.97
    assert r32["samples"] > 0


def test_lattice_schedule_is_continuous_fp32_vs_fp64():
    """The lattice schedule (midpoint samples, coverage-weighted last sample, lattice skips,
    ERT) has no whole-sample flips: FP32 and FP64 renders of the same data agree to ~1e-5."""
    from hydrant_amd import _lib as L
    from hydrant_amd import workloads as W
    from tests.helpers import oracle_view, oracle_volume

    sc = W.paged_volume(n=96, W=48, H=48, bs=32, resident=1.0, schedule=L.SCHED_LATTICE)
    sc.ert = 0.9995
    ov = oracle_view(sc)
    r32 = oracle_volume(sc, exact_mapping=True).render(ov, 32, film=False)
    r64 = oracle_volume(sc, exact_mapping=True).render(ov, 64, film=False)
    assert np.max(np.abs(r32["rgba"] - r64["rgba"])) < 2e-5


def test_lattice_coverage_weight_closed_form():
    """Constant field along an axis ray: L = chord / step samples-worth of opacity, i.e.
    v.w = 1 - (1 - a)^floor(L) * (1 - a * frac(L)) with a the per-sample opacity."""
    from hydrant_amd import _lib as L

    pool = np.full((1, 16, 16, 16), np.float32(0.5), np.float32)
    tf = np.tile(np.asarray([0.2, 0.4, 0.6, 1.0], np.float32), (4, 1))
    step = 0.013
    v = O.make_view((3, 0, 0), (0, 0, 0), (0, 1, 0), 60.0, (0.5, 0.5, 0.5), (1, 1, 1), 64)
    sc = O.VolumeScene(2, 2, (0, 0, 0), (1, 1, 1), step, 1 / 16, tf, np.zeros((1, 1, 1)), np.zeros((1, 1, 1)),
                       np.ones(1), np.zeros((1, 3)), pool, 16, density=0.05, schedule=L.SCHED_LATTICE)
    r = sc.render(v, 64, film=True, counts=True)
    lf = 1.0 / step
    a = 0.05
    want = 1 - (1 - a) ** np.floor(lf) * (1 - a * (lf - np.floor(lf)))
    assert abs(r["rgba"][1, 1, 3] - want) < 1e-12
    assert r["counts"][1, 1] == int(np.ceil(lf))
    assert abs(r["film"][1, 1, 6] - lf) < 1e-12


# ---------------------------------------------------------- resume (A16) ----
def test_oracle_resume_bit_exact():
    """Partial residency then full residency + resume == one uninterrupted march, bitwise."""
    from hydrant_amd import workloads as W
    from tests.helpers import oracle_view, oracle_volume

    sc = W.paged_volume(n=96, W=40, H=40, bs=32, resident=1.0)
    full = sc.bricked.vaddr
    part = full.copy()
    part[np.random.default_rng(9).random(part.shape) < 0.5] = -1
    ov = oracle_view(sc)
    ref = oracle_volume(sc).render(ov, 32, film=True)
    first = oracle_volume(sc, vaddr=part).render(ov, 32, film=True)
    res = oracle_volume(sc).resume(first["film"], 32)
    assert np.array_equal(res["film"], ref["film"])
    assert first["samples"] + res["samples"] == ref["samples"]


# ------------------------------------------------------ composite (A21) -----
def test_composite_is_over_and_associative():
    """Front-to-back composite of fetched partials equals 'over' (theta == rgb, phi == 1 - a),
    and compositing a pre-merged pair gives the same result (associativity)."""
    rng = np.random.default_rng(6)
    N, H, Wd = 3, 5, 7
    parts = np.zeros((N, H, Wd, 8), np.float32)
    for i in range(N):
        a = rng.random((H, Wd)).astype(np.float32)
        rgb = (rng.random((H, Wd, 3)) * a[..., None]).astype(np.float32)
        parts[i, ..., 0:3], parts[i, ..., 3], parts[i, ..., 4:7], parts[i, ..., 7] = rgb, 1 - a, rgb, a
    z = [1, 2, 0]
    rgba, _ = O.composite(parts, z)
    acc = np.zeros((H, Wd, 4))
    for k in z:
        p = parts[k].astype(np.float64)
        acc[..., :3] += (1 - acc[..., 3:4]) * p[..., 4:7]
        acc[..., 3] += (1 - acc[..., 3]) * p[..., 7]
    assert np.allclose(rgba, acc, atol=1e-6)
    pair, _ = O.composite(parts[[1, 2]], [0, 1])
    merged = np.zeros((H, Wd, 8), np.float32)
    merged[..., 4:8] = pair
    merged[..., 0:3] = pair[..., 0:3]
    merged[..., 3] = 1 - pair[..., 3]
    two = np.stack([merged, parts[0]])
    rgba2, _ = O.composite(two, [0, 1])
    assert np.allclose(rgba2, rgba, atol=1e-6)


def test_pixel_pass_saturate():
    rgba = np.array([[[0.5, 1.2, -0.1, 0.3], [0.2, 0.2, 0.2, 0.0]]], np.float32)
    out = O.pixel_pass(rgba, clear=(1, 2, 3))
    assert out[0, 0].tolist() == [127, 255, 0] and out[0, 1].tolist() == [1, 2, 3]


# ---- Paging renderer's residency view (paging_shader.cu:24-61) --------------
def _row_view(nx_origin=-0.5):
    """A view whose centre pixel (1, 1) of a 2 x 2 image looks along +x from
    (nx_origin, 1.5, 1.5): camera -z maps to +x, ctg 1."""
    v = O.OView()
    v.trans[:] = [0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, 0, nx_origin, 1.5, 1.5, 1]
    v.ctg_fovy_2 = 1.0
    v.ray_o[:] = [nx_origin, 1.5, 1.5]
    return v


def _paging_centre(cheb, vaddr, lowest):
    nz, ny, nx = cheb.shape
    rgba = O.blocks_render(_row_view(), 2, 2, (0, 0, 0), (nx, ny, nz), 1.0 / 64, cheb, None, mode=2,
                           vaddr=vaddr, lowest=lowest)
    return rgba[1, 1]


@pytest.mark.parametrize("prec", [32, 64])
def test_paging_view_closed_forms(prec):
    """One composite per resident brick along the ray (not per sample), green
    (0, .5, 0, .5) for pages >= lowest_blkcnt and red (.2, 0, 0, .2) below, front
    to back; a page fault stops the ray; empty bricks are skipped; the ray stops
    once alpha exceeds 0.95."""
    cheb = np.zeros((3, 3, 3), np.int32)
    vaddr = np.full((3, 3, 3), 5, np.int32)
    vaddr[1, 1, :] = [12, 3, 11]                      # green, red, green along the ray (lowest 10)
    np.testing.assert_allclose(_paging_centre(cheb, vaddr, 10), [0.1, 0.7, 0.0, 0.8], atol=1e-6)
    v2 = vaddr.copy()
    v2[1, 1, 1] = -1                                  # page fault in the middle brick
    np.testing.assert_allclose(_paging_centre(cheb, v2, 10), [0.0, 0.5, 0.0, 0.5], atol=1e-6)
    c2 = cheb.copy()
    c2[1, 1, 1] = 1                                   # empty middle brick: skipped, its page unread
    np.testing.assert_allclose(_paging_centre(c2, v2, 10), [0.0, 0.75, 0.0, 0.75], atol=1e-6)
    long_cheb = np.zeros((3, 3, 8), np.int32)
    long_vaddr = np.full((3, 3, 8), 20, np.int32)     # eight green bricks: stops after the fifth
    np.testing.assert_allclose(_paging_centre(long_cheb, long_vaddr, 10), [0.0, 0.96875, 0.0, 0.96875],
                               atol=1e-6)
    long_vaddr[1, 1, 5] = -1                          # a fault past the stop changes nothing
    np.testing.assert_allclose(_paging_centre(long_cheb, long_vaddr, 10), [0.0, 0.96875, 0.0, 0.96875],
                               atol=1e-6)
