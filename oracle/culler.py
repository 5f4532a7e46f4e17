"""Frustum culling and realtime residency bookkeeping restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
Follows the reference operation by operation in binary32 (no contraction, the
operation order of the product's host math):
  - Frustum::contains / contains_fast: include/hydrant/octree_culler.hpp:20-50
  - get_frustrum (default ScreenRect [-1, 1]^2): octree_culler.hpp:88-92,190-213
    (the border directions stay in camera space, as the reference leaves them)
  - OctreeCuller::cull: octree_culler.hpp:121-161 (non-empty bricks, nearest
    `limit` by squared centre distance, then increasing index; ties broken by
    index here, where the reference's nth_element leaves them unspecified)
  - Ray::intersect: include/hydrant/core/glm_math.hpp:58-72
  - RtBlockPagingServerImpl::update and the decode callback:
    src/slave/paging/rt_block_paging.cc:249-326
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def mat4_vec4(m, x, y, z, w):
    """glm mat4 * vec4 on a column-major 16-float matrix: (m0 x + m1 y) + (m2 z + m3 w)."""
    m = np.asarray(m, f32)
    x, y, z, w = f32(x), f32(y), f32(z), f32(w)
    return [f32(f32(m[r] * x) + f32(m[4 + r] * y)) + f32(f32(m[8 + r] * z) + f32(m[12 + r] * w)) for r in range(4)]


def dot3(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def cross3(a, b):
    return (f32(f32(a[1] * b[2]) - f32(b[1] * a[2])), f32(f32(a[2] * b[0]) - f32(b[2] * a[0])),
            f32(f32(a[0] * b[1]) - f32(b[0] * a[1])))


def normalize3(v):
    s = f32(f32(1.0) / f32(np.sqrt(dot3(v, v))))
    return (f32(v[0] * s), f32(v[1] * s), f32(v[2] * s))


def _gmin(x, y):
    return y if y < x else x


def _gmax(x, y):
    return y if x < y else x


def intersect_box(o, d, bmin, bmax):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = [f32(f32(1.0) / f32(d[i])) for i in range(3)]
        tbot = [f32(inv[i] * f32(f32(bmin[i]) - f32(o[i]))) for i in range(3)]
        ttop = [f32(inv[i] * f32(f32(bmax[i]) - f32(o[i]))) for i in range(3)]
    tmin = [_gmin(ttop[i], tbot[i]) for i in range(3)]
    tmax = [_gmax(ttop[i], tbot[i]) for i in range(3)]
    tnear = _gmax(_gmax(tmin[0], tmin[1]), tmin[2])
    tfar = _gmin(_gmin(tmax[0], tmax[1]), tmax[2])
    return bool(tfar > tnear), tnear, tfar


class Frustum:
    def __init__(self, trans, ctg):
        c = f32(ctg)
        self.border = [(f32(1), f32(1), -c), (f32(1), f32(-1), -c), (f32(-1), f32(-1), -c), (f32(-1), f32(1), -c)]
        o = mat4_vec4(trans, 0, 0, 0, 1)
        self.orig = (o[0], o[1], o[2])
        self.norm = []
        for i in range(4):
            n = cross3(self.border[i], self.border[(i + 1) % 4])
            w = mat4_vec4(trans, n[0], n[1], n[2], 0)
            self.norm.append((w[0], w[1], w[2]))

    def contains(self, p):
        q = tuple(f32(f32(p[i]) - self.orig[i]) for i in range(3))
        return all(dot3(n, q) > 0 for n in self.norm)

    def contains_fast(self, bmin, bmax):
        vp = (bmin, bmax)
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    if self.contains((vp[i][0], vp[j][1], vp[k][2])):
                        return True
        for b in self.border:
            hit, _, tfar = intersect_box(self.orig, normalize3(b), bmin, bmax)
            if hit and tfar > 0:
                return True
        return False


def centre_dist2(orig, x, y, z):
    d = (f32(orig[0] - f32(f32(x) + f32(0.5))), f32(orig[1] - f32(f32(y) + f32(0.5))),
         f32(orig[2] - f32(f32(z) + f32(0.5))))
    return dot3(d, d)


def cull(trans, ctg, cheb_zyx: np.ndarray, limit=None, bbox=None):
    """Linear brick indices (x fastest) in increasing order, and the frustum."""
    nz, ny, nx = cheb_zyx.shape
    lo, hi = bbox if bbox is not None else ((0, 0, 0), (nx, ny, nz))
    f = Frustum(trans, ctg)
    cand = []
    for z in range(lo[2], hi[2]):
        for y in range(lo[1], hi[1]):
            for x in range(lo[0], hi[0]):
                if cheb_zyx[z, y, x] != 0:
                    continue
                bmin = (f32(x), f32(y), f32(z))
                bmax = (f32(x + 1), f32(y + 1), f32(z + 1))
                if f.contains_fast(bmin, bmax):
                    cand.append((centre_dist2(f.orig, x, y, z), (z * ny + y) * nx + x))
    cand.sort()
    n = len(cand) if limit is None else max(0, min(int(limit), len(cand)))
    return np.array(sorted(i for _, i in cand[:n]), np.int32), f


class ResidencyModel:
    """The reference's present / missing / redundant sets and slot assignment."""

    def __init__(self, n_bricks, max_pages, first_slot=0, basic=None):
        self.max_pages, self.first_slot = max_pages, first_slot
        self.basic = np.full(n_bricks, -1, np.int32) if basic is None else np.asarray(basic, np.int32).ravel()
        self.vaddr = self.basic.copy()
        self.present, self.redundant, self.n_slots = set(), [], 0

    def update(self, required, dist):
        req = set(int(i) for i in required)
        missing = sorted(req - self.present, key=lambda i: (dist(i), i))
        self.redundant = sorted(self.present - req)
        return missing

    def commit(self, idx):
        if idx in self.present:
            return -1                       # abandoned
        if self.n_slots < self.max_pages:
            s = self.first_slot + self.n_slots
            self.n_slots += 1
        elif self.redundant:
            out = self.redundant.pop()      # the last redundant brick
            self.present.discard(out)
            s = int(self.vaddr[out])
            self.vaddr[out] = self.basic[out]
        else:
            return -1                       # artifact
        self.vaddr[idx] = s
        self.present.add(idx)
        return s
