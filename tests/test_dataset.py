"""CPU tests of dataset preparation (include/hydrant_dataset.h) against numpy brute force."""
import ctypes as C

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import dataset as D


def test_f16_conversions_match_numpy():
    lib = L.lib()
    halves = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    want = halves.view(np.float16).astype(np.float32)
    got = np.array([lib.hyd_f16_to_f32(int(h)) for h in halves[::7]], np.float32)
    w = want[::7]
    finite = np.isfinite(w)
    assert np.array_equal(got[finite], w[finite])
    assert np.all(np.isnan(got[np.isnan(w)]))
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.normal(scale=s, size=4000).astype(np.float32) for s in (1e-6, 1e-4, 1.0, 300.0, 7e4)]
                        + [np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8,
                                     np.inf, -np.inf], np.float32)])
    got = np.array([lib.hyd_f32_to_f16(float(x)) for x in xs], np.uint16)
    assert np.array_equal(got, xs.astype(np.float16).view(np.uint16))


def _pack_ref(vol, bs, apron, brick):
    nz, ny, nx = vol.shape
    g = [(nx + bs - 1) // bs, (ny + bs - 1) // bs, (nz + bs - 1) // bs]
    bx, by, bz = brick % g[0], (brick // g[0]) % g[1], brick // (g[0] * g[1])
    pad = bs + 2 * apron
    ix = np.clip(bx * bs + np.arange(pad) - apron, 0, nx - 1)
    iy = np.clip(by * bs + np.arange(pad) - apron, 0, ny - 1)
    iz = np.clip(bz * bs + np.arange(pad) - apron, 0, nz - 1)
    return vol[np.ix_(iz, iy, ix)]


@pytest.mark.parametrize("fmt", [L.FMT_U8, L.FMT_F16, L.FMT_F32])
def test_pack_bricks_apron_clamp(fmt):
    rng = np.random.default_rng(1)
    vol = rng.integers(0, 256, size=(20, 17, 23), dtype=np.uint8)
    bricks = np.array([0, 5, 11, 3], np.int32)
    out = D.pack(vol, 8, 1, fmt, bricks)
    norm = vol.astype(np.float32) / np.float32(255)
    for p, b in enumerate(bricks):
        if fmt == L.FMT_U8:
            assert np.array_equal(out[p], _pack_ref(vol, 8, 1, b))
        elif fmt == L.FMT_F32:
            assert np.array_equal(out[p], _pack_ref(norm, 8, 1, b))
        else:
            assert np.array_equal(out[p], _pack_ref(norm, 8, 1, b).astype(np.float16).view(np.uint16))


def test_brick_stats_and_mean():
    rng = np.random.default_rng(2)
    vol = rng.integers(0, 256, size=(16, 16, 24), dtype=np.uint8)
    grid, bmin, bmax, bmean = D.brick_stats(vol, 8, 1)
    assert grid == [3, 2, 2]
    norm = vol.astype(np.float32) / np.float32(255)
    for bz in range(2):
        for by in range(2):
            for bx in range(3):
                b = (bz * 2 + by) * 3 + bx
                padded = _pack_ref(norm, 8, 1, b)
                inner = norm[bz * 8:(bz + 1) * 8, by * 8:(by + 1) * 8, bx * 8:(bx + 1) * 8]
                assert bmin[bz, by, bx] == padded.min() and bmax[bz, by, bx] == padded.max()
                assert abs(bmean[bz, by, bx] - inner.mean(dtype=np.float64)) < 1e-6


def test_chebyshev_brute_force():
    rng = np.random.default_rng(3)
    empty = (rng.random((5, 6, 7)) < 0.8).astype(np.uint8)
    cheb = D.chebyshev(empty)
    filled = np.argwhere(empty == 0)
    for idx in np.ndindex(empty.shape):
        if not empty[idx]:
            assert cheb[idx] == 0
        else:
            assert cheb[idx] == np.abs(filled - np.array(idx)).max(axis=1).min()
    allempty = np.ones((2, 2, 3), np.uint8)
    assert np.all(D.chebyshev(allempty) == 3)


def test_seeded_page_table():
    rng = np.random.default_rng(4)
    empty = (rng.random((6, 6, 6)) < 0.3).astype(np.uint8)
    vol = rng.integers(0, 256, size=(48, 48, 48), dtype=np.uint8)
    b = D.make_bricked(vol, 8, 1, L.FMT_U8, empty_fn=lambda lo, hi: empty.astype(bool), resident=50, seed=11)
    res = b.vaddr[b.vaddr >= 0]
    assert sorted(res.tolist()) == list(range(50))
    assert np.all(empty[b.vaddr >= 0] == 0)
    assert b.n_nonempty == int((empty == 0).sum())
    # page p holds brick order[p]
    for p in (0, 17, 49):
        assert np.array_equal(b.pool[p], _pack_ref(vol, 8, 1, int(b.order[p])))
        assert b.vaddr.reshape(-1)[b.order[p]] == p
    b2 = D.make_bricked(vol, 8, 1, L.FMT_U8, empty_fn=lambda lo, hi: empty.astype(bool), resident=50, seed=11)
    assert np.array_equal(b.vaddr, b2.vaddr)


def test_mapping_kb():
    """k = bs / pad, b = (apron / bs) * k in binary32 (lossless_block_paging.cc:21-22)."""
    k, b = D.mapping(64, 1)
    assert k == np.float32(64) / np.float32(66) and b == (np.float32(1) / np.float32(64)) * k
    assert D.mapping(256, 0) == (np.float32(1), np.float32(0))


def test_tf_alpha_zero_on_brute_force():
    from oracle import oracle as O

    tf = D.README_TF
    rng = np.random.default_rng(5)
    for _ in range(300):
        lo = rng.uniform(-0.1, 1.0)
        hi = min(lo + rng.uniform(0, 0.3), 1.0)
        s = np.linspace(lo, hi, 97)
        zero = all(O.tf_sample(tf, x, 64)[3] == 0 for x in s)
        if D.tf_alpha_zero_on(tf, lo, hi):
            assert zero, (lo, hi)   # conservative: never marks a brick with visible data empty
    assert D.tf_alpha_zero_on(tf, 0.0, 0.05) and D.tf_alpha_zero_on(tf, 0.95, 1.0)
    assert not D.tf_alpha_zero_on(tf, 0.0, 0.2)


def test_blobs_generator_closed_form():
    """One blob of amplitude a, sigma s at centre c: u8 = floor(255 * min(1, a*exp(-r^2/2s^2)) + 0.5)."""
    vol = D.gen_blobs((16, 12, 10), n_blobs=1, seed=42, sig=(3.0, 3.0), amp=(0.8, 0.8))
    # recover the drawn centre from the same SplitMix64 stream
    def splitmix(s):
        out = []
        for _ in range(5):
            s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
            out.append(((z ^ (z >> 31)) >> 11) * 2.0 ** -53)
        return out
    u = splitmix(42)
    c = np.array([u[0] * 16, u[1] * 12, u[2] * 10])
    z, y, x = np.meshgrid(np.arange(10), np.arange(12), np.arange(16), indexing="ij")
    r2 = (x + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (z + 0.5 - c[2]) ** 2
    f = 0.8 * np.exp(-r2 / (2 * 9.0))
    want = np.floor(255 * np.minimum(f, 1) + 0.5).astype(np.uint8)
    assert np.max(np.abs(vol.astype(int) - want.astype(int))) <= 1
    assert np.mean(vol == want) > 0.99


def test_sdf_generator_single_sphere():
    """One sphere: the stored half equals |p - c| - r rounded to half."""
    s = D.gen_sdf_spheres(24, n_spheres=1, r=(5.0, 5.0), k=2.0, seed=9).view(np.float16).astype(np.float32)
    inside = s < 0
    assert 0 < inside.sum() < 24 ** 3
    n_in = inside.sum()
    assert abs(n_in - 4 / 3 * np.pi * 125) / (4 / 3 * np.pi * 125) < 0.08


def test_generators_deterministic_and_hashed():
    import json
    from pathlib import Path

    man = json.loads((Path(__file__).parent / "golden" / "manifest.json").read_text())
    vol = D.gen_blobs(256, 16, 1001, sig=(8.0, 40.0), sigma_scale=1.0, amp=(0.3, 1.0))
    assert D.sha256(vol) == man["inputs"]["blobs_256_seed1001"]
    p = D.gen_perlin(128, 5, 0.5, 16.0, 2.0, 2002)
    assert D.sha256(p) == man["inputs"]["perlin_128_seed2002"]
    s = D.gen_sdf_spheres(128, 64, (2.5, 15.0), 2.0, 3003)
    assert D.sha256(s) == man["inputs"]["sdf_128_seed3003"]
