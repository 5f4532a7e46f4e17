"""Frustum culling and realtime residency (host code of the library) against the numpy
restatement in oracle/culler.py, whose geometry is pinned here by properties that do
not restate it: points on rays through the image lie inside the frustum, bricks deep
inside the frustum are kept, a volume entirely in view is kept whole, and `limit`
keeps the nearest bricks."""
import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import dataset as D
from hydrant_amd.hip import make_view, orbit_camera
from hydrant_amd.residency import Residency, cull_bricks
from oracle import culler as OC

ARMS = [(25.0, 65.0, 2.2), (200.0, -30.0, 1.6), (90.0, 10.0, 0.9), (310.0, 80.0, 3.0), (0.0, 0.0, 0.45)]


def _grid(n=10, seed=3, frac_empty=0.35):
    rng = np.random.default_rng(seed)
    return (rng.random((n, n, n)) < frac_empty).astype(np.int32)   # [z][y][x], nonzero = empty


def _view(arm, n, fovy=60.0):
    g = D.geometry((8 * n,) * 3, 8)
    return make_view(orbit_camera(arm, (0.0, 0.0, 0.0), fovy), g.center, g.f_dim)


def _trans(v):
    return np.array(v.trans[:], np.float32)


def _dir(v, u, w):
    """World direction of the camera-space ray (u, w, -ctg)."""
    return np.array(OC.mat4_vec4(_trans(v), u, w, -v.ctg_fovy_2, 0.0)[:3], np.float64)


@pytest.mark.parametrize("arm", ARMS)
def test_cull_matches_restatement(arm):
    n = 10
    cheb = _grid(n)
    v = _view(arm, n)
    got = cull_bricks(v, cheb)
    ref, _ = OC.cull(_trans(v), v.ctg_fovy_2, cheb)
    assert np.array_equal(got, ref)
    for limit in (0, 1, 17, 10 ** 6):
        ref_l, _ = OC.cull(_trans(v), v.ctg_fovy_2, cheb, limit=limit)
        assert np.array_equal(cull_bricks(v, cheb, limit=limit), ref_l)
    box = ((2, 1, 3), (9, 7, 10))
    ref_b, _ = OC.cull(_trans(v), v.ctg_fovy_2, cheb, bbox=box)
    assert np.array_equal(cull_bricks(v, cheb, bbox=box), ref_b)


@pytest.mark.parametrize("arm", ARMS)
def test_restated_frustum_geometry(arm):
    """Points along rays through the image's interior are inside; points along rays well
    outside the screen rect are not."""
    v = _view(arm, 10)
    f = OC.Frustum(_trans(v), v.ctg_fovy_2)
    o = np.array(f.orig, np.float64)
    rng = np.random.default_rng(11)
    for _ in range(200):
        u, w = rng.uniform(-0.95, 0.95, 2)
        t = rng.uniform(0.1, 20.0)
        assert f.contains(o + t * _dir(v, u, w))
        u_out = rng.choice([-1, 1]) * rng.uniform(1.1, 3.0)
        assert not f.contains(o + t * _dir(v, u_out, w))
        assert not f.contains(o + t * _dir(v, w, u_out))


@pytest.mark.parametrize("arm", ARMS)
def test_restated_cull_keeps_bricks_deep_inside(arm):
    """A non-empty brick whose centre lies more than its half-diagonal inside all four
    planes has every corner inside, so the cull keeps it; empty bricks never appear."""
    n = 10
    cheb = _grid(n)
    v = _view(arm, n)
    kept, f = OC.cull(_trans(v), v.ctg_fovy_2, cheb)
    kept = set(kept.tolist())
    o = np.array(f.orig, np.float64)
    norms = [np.array(m, np.float64) / np.linalg.norm(m) for m in f.norm]
    for z in range(n):
        for y in range(n):
            for x in range(n):
                i = (z * n + y) * n + x
                if cheb[z, y, x]:
                    assert i not in kept
                    continue
                c = np.array([x + 0.5, y + 0.5, z + 0.5]) - o
                if min(float(m @ c) for m in norms) > 0.9:
                    assert i in kept, (x, y, z)


def test_restated_cull_keeps_a_volume_in_full_view():
    n = 8
    cheb = _grid(n, seed=9)
    v = _view((40.0, 30.0, 6.0), n, fovy=60.0)
    kept, _ = OC.cull(_trans(v), v.ctg_fovy_2, cheb)
    assert np.array_equal(kept, np.flatnonzero(cheb.ravel() == 0).astype(np.int32))


@pytest.mark.parametrize("arm", ARMS[:3])
def test_restated_limit_keeps_the_nearest(arm):
    n = 10
    cheb = _grid(n)
    v = _view(arm, n)
    full, f = OC.cull(_trans(v), v.ctg_fovy_2, cheb)
    o = np.array(f.orig, np.float64)

    def d(i):
        x, y, z = i % n, (i // n) % n, i // (n * n)
        return float(np.sum((np.array([x + 0.5, y + 0.5, z + 0.5]) - o) ** 2))

    for k in (1, 9, len(full) // 2):
        part, _ = OC.cull(_trans(v), v.ctg_fovy_2, cheb, limit=k)
        assert len(part) == k and set(part) <= set(full)
        assert max(d(i) for i in part) <= min(d(i) for i in set(full) - set(part)) + 1e-4


def _dist(f, n):
    return lambda i: OC.centre_dist2(f.orig, i % n, (i // n) % n, i // (n * n))


def _check_invariants(res, model, first, max_pages):
    va = res.vaddr().ravel()
    assert np.array_equal(va, model.vaddr)
    slots = [va[i] for i in model.present]
    assert len(set(slots)) == len(slots)
    assert all(first <= s < first + max_pages for s in slots)
    others = np.ones(va.size, bool)
    others[list(model.present)] = False
    assert np.array_equal(va[others], model.basic[others])
    assert res.stats() == (len(model.present), model.n_slots)


PATH = [(25.0 + 6 * k, 65.0 - 4 * k, 2.2 - 0.12 * k) for k in range(8)]   # a camera moving a little per frame


@pytest.mark.parametrize("max_pages,first", [(40, 0), (25, 7), (120, 2), (1000, 3)])
def test_residency_follows_the_reference_sets(max_pages, first):
    """Missing sets (nearest first), slot assignment (new slots, then the last redundant
    brick's slot with its entry reset to the basic one) and the abandoned / artifact cases
    follow the reference's bookkeeping, view after view."""
    n = 10
    cheb = _grid(n)
    basic = np.random.default_rng(4).integers(-1, 3, size=cheb.shape).astype(np.int32)
    res = Residency(cheb.shape, max_pages, first_slot=first, basic_vaddr=basic)
    model = OC.ResidencyModel(cheb.size, max_pages, first, basic)
    kept = swapped = 0
    for arm in PATH + ARMS + ARMS[:2]:
        v = _view(arm, n)
        missing = res.update(v, cheb)
        required, f = OC.cull(_trans(v), v.ctg_fovy_2, cheb, limit=max_pages)
        want = model.update(required, _dist(f, n))
        assert missing.tolist() == want
        # commit every missing brick, one already present brick, and a repeat
        kept += len(set(required.tolist()) & model.present)
        extra = [int(next(iter(model.present)))] if model.present else []
        for i in list(missing) + extra + list(missing[:1]):
            n_slots = model.n_slots
            s = model.commit(int(i))
            swapped += s >= 0 and model.n_slots == n_slots
            assert res.commit(int(i)) == s
        _check_invariants(res, model, first, max_pages)
    assert kept > 0                               # frames share resident bricks
    assert swapped > 0 or max_pages >= 1000      # and a full pool reuses redundant slots
    res.close()


def test_residency_artifact_when_nothing_can_be_freed():
    n = 10
    cheb = _grid(n)
    res = Residency(cheb.shape, 3)
    v = _view(ARMS[0], n)
    missing = res.update(v, cheb)
    assert len(missing) == 3
    assert [res.commit(int(i)) for i in missing] == [0, 1, 2]
    other = int(np.flatnonzero(cheb.ravel() == 0)[-1])
    if other in missing:
        other = int(np.flatnonzero(cheb.ravel() == 0)[0])
    assert res.commit(other) == -1          # pool full, nothing redundant
    assert res.update(v, cheb).size == 0    # the same view needs nothing new
    res.close()


def test_residency_rejects_bad_arguments():
    cheb = np.zeros((4, 4, 4), np.int32)
    with pytest.raises(L.HydError) as e:
        cull_bricks(_view(ARMS[0], 4), cheb, bbox=((0, 0, 0), (5, 4, 4)))
    assert e.value.code == L.HYD_ERR_RANGE
    with pytest.raises(L.HydError):
        cull_bricks(_view(ARMS[0], 4), cheb, bbox=((3, 0, 0), (2, 4, 4)))
    assert cull_bricks(_view(ARMS[0], 4), cheb, limit=-3).size == 0
    with pytest.raises(L.HydError):
        Residency((0, 4, 4), 8)
    res = Residency((4, 4, 4), 8)
    with pytest.raises(L.HydError):
        res.commit(64)
    with pytest.raises(ValueError):
        res.update(_view(ARMS[0], 4), np.zeros((4, 4, 5), np.int32))
    res.close()


def _sample(page: np.ndarray, k: float, b, x):
    """Sampler of one padded page at brick-local point x: u = k x + b, then the
    normalised-coordinate linear filter with clamped taps."""
    pad = page.shape[0]
    w, i, j = [], [], []
    for c in range(3):
        f = pad * (k * x[c] + b[c]) - 0.5
        fl = np.floor(f)
        w.append(f - fl)
        i.append(int(np.clip(fl, 0, pad - 1)))
        j.append(int(np.clip(np.ceil(f), 0, pad - 1)))
    v = lambda xx, yy, zz: float(page[zz, yy, xx])
    x00 = v(i[0], i[1], i[2]) * (1 - w[0]) + v(j[0], i[1], i[2]) * w[0]
    x10 = v(i[0], j[1], i[2]) * (1 - w[0]) + v(j[0], j[1], i[2]) * w[0]
    x01 = v(i[0], i[1], j[2]) * (1 - w[0]) + v(j[0], i[1], j[2]) * w[0]
    x11 = v(i[0], j[1], j[2]) * (1 - w[0]) + v(j[0], j[1], j[2]) * w[0]
    y0 = x00 * (1 - w[1]) + x10 * w[1]
    y1 = x01 * (1 - w[1]) + x11 * w[1]
    return y0 * (1 - w[2]) + y1 * w[2]


@pytest.mark.parametrize("level", [1, 2])
def test_lowest_level_mapping_reproduces_a_linear_field(level):
    """A box-averaged coarse level of a linear field, read through each level-0 brick's
    fallback page and mapping, gives the field's value at the level-0 position (the
    average of a linear function is its value at the box centre), away from the clamped
    volume border; a wrong scale, offset or block index would not."""
    n, bs, apron = 128, 8, 1
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    coef = (0.5, 1.25, 2.0)
    vol = (coef[0] * x + coef[1] * y + coef[2] * z).astype(np.float32)
    low = D.lowest_level(vol, bs, apron, L.FMT_F32, level)
    s = 1 << level
    g = n // bs
    assert low.pages.shape[0] == g ** 3 and low.scale == s
    rng = np.random.default_rng(2)
    margin = s   # bricks of coarse blocks off the volume border (whose aprons are not clamped)
    checked = 0
    for bz in range(margin, g - margin):
        for by in range(margin, g - margin):
            for bx in range(margin, g - margin):
                i = (bz * g + by) * g + bx
                for _ in range(3):
                    p = rng.uniform(0.0, 1.0, 3)
                    got = _sample(low.pages[i], float(low.page_k[i]), low.page_b[i], p)
                    pos = (np.array([bx, by, bz]) + p) * bs - 0.5   # level-0 voxel coordinate
                    want = float(np.dot(coef, pos))
                    assert abs(got - want) < 1e-3, (bx, by, bz, got, want)
                    checked += 1
    assert checked > 0
