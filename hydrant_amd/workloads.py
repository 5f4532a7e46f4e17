"""Seeded synthetic workloads for the render path.

Each builder returns a :class:`Scene` that holds the bricked dataset, the
camera/view and the shader parameters of one benchmark configuration, at
full size or scaled down (``n`` voxels per axis, ``W`` x ``H`` pixels) for
parity tests.  Geometry follows BasicRenderer::init/update
(include/hydrant/basic_renderer.hpp:57-84): object space is in block units,
bbox = [0, raw / block_size], du = 1 / block_size, step = du / 4 / sample_rate.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import dataset as D

SEED_BLOBS = 1001
SEED_PERLIN = 2002
SEED_PAGES = 2003
SEED_SDF = 3003
SEED_BATCH = 4000
README_ARM = (25.0, 65.0, 2.2)


@dataclass
class Scene:
    name: str
    bricked: D.Bricked
    tf: np.ndarray
    W: int
    H: int
    arm: tuple
    fovy: float
    geom: D.SceneGeometry
    density: float = 1.0
    max_steps: int = 4_000_000
    schedule: int = L.SCHED_REFERENCE
    ert: float = 0.0
    mode: int = L.MODE_DEFAULT
    rank: float = 0.0          # Partition mode colour parameter
    extra: dict = field(default_factory=dict)

    @property
    def bbox_min(self):
        return np.zeros(3, np.float32)

    @property
    def bbox_max(self):
        return self.geom.f_dim

    def camera(self):
        from .hip import orbit_camera
        return orbit_camera(self.arm, (0.0, 0.0, 0.0), self.fovy)

    def view(self):
        from .hip import make_view
        return make_view(self.camera(), self.geom.center, self.geom.f_dim)

    def volume_params(self):
        from .hip import volume_params
        return volume_params(self.bbox_min, self.bbox_max, self.geom.step, self.geom.du, self.max_steps,
                             self.mode, self.rank, self.density, self.schedule, self.ert)


def plain_volume(n: int = 256, W: int = 512, H: int = 512, fmt: int = L.FMT_F32, density: float = 1e-2,
                 sample_rate: float = 0.5, seed: int = SEED_BLOBS) -> Scene:
    """Plain Volume mode: n^3 u8 sum of 16 Gaussian blobs (sigma 8-40 voxels at n = 256,
    scaled with n; amplitude U(0.3, 1)), one brick (k = 1, b = 0), README TF and
    orbit camera, step 0.5 voxel, density 1e-2."""
    vol = D.gen_blobs(n, 16, seed, sig=(8.0, 40.0), sigma_scale=n / 256.0, amp=(0.3, 1.0))
    b = D.single_brick(vol, fmt)
    geom = D.geometry((n, n, n), n, sample_rate)
    return Scene("plain_volume", b, D.README_TF.copy(), W, H, README_ARM, 60.0, geom, density=density,
                 extra={"volume": vol})


def paged_volume(n: int = 1024, W: int = 512, H: int = 512, fmt: int = L.FMT_U8, bs: int = 64, apron: int = 1,
                 resident: float = 0.4, density: float = 1e-2, sample_rate: float = 0.5, seed: int = SEED_PERLIN,
                 page_seed: int = SEED_PAGES, schedule: int = L.SCHED_REFERENCE) -> Scene:
    """Blocks.Volume paged: n^3 u8 Perlin fBm (5 octaves, persistence 0.5, base period
    n/8 voxels) in bs^3 bricks + apron, `resident` of the bricks resident through
    a seeded page table, empty bricks (TF alpha zero over min/max) skipped."""
    vol = D.gen_perlin(n, 5, 0.5, n / 8.0, 2.0, seed)
    tf = D.README_TF.copy()
    nb = ((n + bs - 1) // bs) ** 3
    n_res = int(round(resident * nb))
    b = D.make_bricked(vol, bs, apron, fmt, D.volume_empty_fn(tf), resident=n_res, seed=page_seed)
    geom = D.geometry((n, n, n), bs, sample_rate)
    return Scene("paged_volume", b, tf, W, H, README_ARM, 60.0, geom, density=density, schedule=schedule,
                 extra={"volume": vol})


def isosurface(n: int = 512, W: int = 1024, H: int = 1024, bs: int = 64, apron: int = 1, n_spheres: int = 64,
               sample_rate: float = 1.0, seed: int = SEED_SDF, corrected: int = 1) -> Scene:
    """Isosurface: n^3 FP16 SDF of seeded spheres (radius 10-60 voxels at n = 512,
    scaled with n), smooth union k = 8, iso 0, bracket + 8 bisections."""
    s = n / 512.0
    sdf = D.gen_sdf_spheres(n, n_spheres, (10.0 * s, 60.0 * s), 8.0 * s, seed)
    b = D.make_bricked(sdf, bs, apron, L.FMT_F16, D.iso_empty_fn(0.0), resident=1.0, seed=1)
    geom = D.geometry((n, n, n), bs, sample_rate)
    sc = Scene("isosurface", b, np.zeros((1, 4), np.float32), W, H, README_ARM, 60.0, geom,
               extra={"volume": sdf, "isovalue": 0.0, "corrected": corrected, "bisect_steps": 8})
    return sc


def iso_params(sc: Scene, mode: int = L.ISO_COLOR):
    """IsosurfaceRenderer::offline_render_ctxed shader setup (isosurface.cc:108-113)."""
    import ctypes as C

    cam = sc.camera()
    c = np.asarray(sc.geom.center, np.float32)
    s = np.asarray(sc.geom.f_dim, np.float32)
    iet = np.zeros(16, np.float32)
    L.call("hyd_exhibit_iet", L.ptr(c, C.c_float), L.ptr(s, C.c_float), L.ptr(iet, C.c_float))
    to_world = np.zeros(16, np.float32)
    L.call("hyd_mat4_inverse", L.ptr(iet, C.c_float), L.ptr(to_world, C.c_float))
    pos = np.array(cam.position[:], np.float32)
    tgt = np.array(cam.target[:], np.float32)
    up = np.array(cam.up[:], np.float32)
    cr = np.array([tgt[1] * up[2] - up[1] * tgt[2], tgt[2] * up[0] - up[2] * tgt[0],
                   tgt[0] * up[1] - up[0] * tgt[1]], np.float32)
    light = ((pos + tgt) + up) + cr
    p = L.IsoParams()
    p.bbox_min = L.f3(0, 0, 0)
    p.bbox_max = L.f3(*[float(x) for x in sc.geom.f_dim])
    p.step = float(sc.geom.step)
    p.du = float(sc.geom.du)
    p.max_steps = sc.max_steps
    p.mode = mode
    p.to_world = L.f16(*[float(x) for x in to_world])
    p.eye_pos = L.f3(*[float(x) for x in pos])
    p.light_pos = L.f3(*[float(x) for x in light])
    p.surface_color = L.f3(1.0, 1.0, 1.0)
    p.isovalue = float(sc.extra["isovalue"])
    p.corrected = int(sc.extra["corrected"])
    p.bisect_steps = int(sc.extra["bisect_steps"])
    return p


def blocks(n: int = 256, W: int = 512, H: int = 512, bs: int = 32, density: float = 1e-2, seed: int = SEED_BLOBS,
           sample_rate: float = 1.0) -> Scene:
    """Blocks.Volume / Blocks.Solid over a blob volume in bs^3 bricks; a brick is
    empty when its maximum is below 1/255 (no visible data)."""
    vol = D.gen_blobs(n, 16, seed, sig=(8.0, 40.0), sigma_scale=n / 256.0, amp=(0.3, 1.0))
    b = D.make_bricked(vol, bs, 0, L.FMT_U8, lambda lo, hi: hi < (1.0 / 255.0), resident=1.0, seed=1)
    geom = D.geometry((n, n, n), bs, sample_rate)
    return Scene("blocks", b, D.README_TF.copy(), W, H, README_ARM, 60.0, geom, density=density)


@dataclass
class BatchedWorkload:
    """Many views of several volumes sharing one brick pool (the batched form)."""
    sizes: tuple
    fmt: int
    volumes: list            # per volume: dict(vol, bricked, geom, page_base)
    pool: np.ndarray         # all volumes' pages, page order
    views: list              # per view: dict(volume, W, H, arm)
    tf: np.ndarray
    density: float
    ert: float
    bs: int
    apron: int

    def volume_descs(self):
        out = []
        for v in self.volumes:
            b = v["bricked"]
            vaddr = b.vaddr.copy()
            vaddr[vaddr >= 0] += v["page_base"]
            out.append({"grid": b.grid, "bs": self.bs, "apron": self.apron, "bbox_max": v["geom"].f_dim,
                        "step": v["geom"].step, "value_scale": b.value_scale, "vaddr": vaddr, "cheb": b.cheb})
        return out

    def hip_views(self):
        from .hip import make_view, orbit_camera
        out = []
        for v in self.views:
            g = self.volumes[v["volume"]]["geom"]
            out.append((make_view(orbit_camera(v["arm"]), g.center, g.f_dim), v["W"], v["H"], v["volume"]))
        return out

    def scene(self, i: int) -> Scene:
        """View i as a single-view lattice Scene over its own volume (for the oracle)."""
        v = self.views[i]
        vol = self.volumes[v["volume"]]
        sc = Scene(f"batched_view{i}", vol["bricked"], self.tf, v["W"], v["H"], v["arm"], 60.0, vol["geom"],
                   density=self.density, schedule=L.SCHED_LATTICE, ert=self.ert,
                   extra={"volume": vol["vol"], "map_bs": self.bs})
        return sc


def batched(sizes=(256, 512, 768), n_views: int = 96, res_choices=(256, 384, 512, 768, 1024), fmt: int = L.FMT_F16,
            bs: int = 64, apron: int = 1, density: float = 1e-2, ert: float = 0.9999, sample_rate: float = 0.5,
            seed: int = SEED_BATCH, normalize: bool = False) -> BatchedWorkload:
    """Batched precision form: seeded Gaussian-blob volumes (sigma scaled by size/256) stored
    as fmt bricks of bs^3 + apron, n_views views with resolutions drawn from res_choices and
    seeded orbit angles (arm.x ~ U[0, 360), arm.y ~ U[-80, 80], distance 2.2).
    normalize=False stores the 8-bit counts (exact in FP16) and renders with scale 1/255;
    normalize=True stores float(v)/255 rounded to the brick format."""
    tf = D.README_TF.copy()
    vols = []
    pools = []
    base = 0
    for i, n in enumerate(sizes):
        vol = D.gen_blobs(n, 16, seed + i, sig=(8.0, 40.0), sigma_scale=n / 256.0, amp=(0.3, 1.0))
        b = D.make_bricked(vol, bs, apron, fmt, D.volume_empty_fn(tf), resident=1.0, seed=seed + 10 + i,
                           normalize=normalize)
        geom = D.geometry((n, n, n), bs, sample_rate)
        vols.append({"vol": vol, "bricked": b, "geom": geom, "page_base": base, "n": n})
        pools.append(b.pool)
        base += b.n_pages
    pool = np.concatenate(pools, axis=0)
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n_views):
        vi = int(rng.integers(0, len(sizes)))
        r = int(res_choices[int(rng.integers(0, len(res_choices)))])
        arm = (float(rng.uniform(0.0, 360.0)), float(rng.uniform(-80.0, 80.0)), 2.2)
        views.append({"volume": vi, "W": r, "H": r, "arm": arm})
    return BatchedWorkload(tuple(sizes), fmt, vols, pool, views, tf, density, ert, bs, apron)


SEED_DIFF = 5005


@dataclass
class DiffWorkload:
    """Differentiable form: fit a volume and a TF to target views of hidden ones."""
    n: int
    W: int
    H: int
    arms: list
    hidden_volume: np.ndarray
    hidden_tf: np.ndarray
    init_volume: np.ndarray
    init_tf: np.ndarray
    n_steps: int = 512
    ckpt_every: int = 64
    sigma: float = 8.0

    def views(self):
        from .hip import make_view, orbit_camera
        c = np.full(3, 0.5, np.float32)
        s = np.ones(3, np.float32)
        return [make_view(orbit_camera(a), c, s) for a in self.arms]


def differentiable(n: int = 128, n_views: int = 32, W: int = 256, H: int = 256, seed: int = SEED_DIFF,
                   n_steps: int = 512, ckpt_every: int = 64, sigma: float = 8.0) -> DiffWorkload:
    """Hidden n^3 FP32 blob volume (8 blobs, sigma 6-20 voxels at n = 128, scaled with n) and a
    hidden 256-entry TF of seeded smooth bumps; n_views seeded orbit views (arm.x ~ U[0, 360),
    arm.y ~ U[-60, 60], distance 2.2).  The fit starts from a grey translucent TF and a volume
    ramping linearly from 0.15 to 0.35 along x + y + z.  (A constant start is a stationary
    point of the exact gradient: every sample reads the same TF pair with equal weights, so the
    two entries stay equal and dL/dvolume stays zero; a deterministic adjoint stays there.)"""
    from .diff import smooth_tf

    vol = D.gen_blobs(n, 8, seed, sig=(6.0, 20.0), sigma_scale=n / 128.0, amp=(0.3, 1.0))
    hidden = vol.astype(np.float32) / np.float32(255.0)
    tf = smooth_tf(256, seed + 1)
    rng = np.random.default_rng(seed + 2)
    arms = [(float(rng.uniform(0.0, 360.0)), float(rng.uniform(-60.0, 60.0)), 2.2) for _ in range(n_views)]
    ramp = np.arange(n, dtype=np.float32) / np.float32(max(n - 1, 1))
    init_v = (np.float32(0.15) + np.float32(0.2 / 3.0) * (ramp[:, None, None] + ramp[None, :, None]
                                                         + ramp[None, None, :])).astype(np.float32)
    init_tf = np.tile(np.array([0.5, 0.5, 0.5, 0.2], np.float32), (256, 1))
    return DiffWorkload(n, W, H, arms, hidden, tf, init_v, init_tf, n_steps, ckpt_every, sigma)


def blocks_params(sc: Scene, mode: int = L.BLOCKS_VOLUME):
    p = L.BlocksParams()
    p.bbox_min = L.f3(0, 0, 0)
    p.bbox_max = L.f3(*[float(x) for x in sc.geom.f_dim])
    p.step = float(sc.geom.step)
    p.max_steps = sc.max_steps
    p.mode = mode
    p.density = float(sc.density)
    return p
