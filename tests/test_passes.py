"""CPU checks of the output passes: the composite's sRGB byte table against the oracle's
binary32 bytes, and the oracle's isosurface fetch and depth-min composite against
closed forms."""
import numpy as np

from hydrant_amd import hip
from oracle import oracle as O


def test_srgb8_table_matches_reference_bytes():
    """hyd_composite writes saturate(linear_to_srgb(x)) (volume.cc:175-181,284-288) as the
    count of thresholds at or below x: equal to the oracle's binary32 bytes for every float
    in [0, 2] and for negative, large, infinite values."""
    t = hip.srgb8_thresholds()
    assert t[0] == -np.inf and np.all(np.diff(t) > 0)
    n, first = O.srgb8_check(t)
    assert n == 0, hex(first)
    # x = 1 gives byte 254: 1.055f * 1 - 0.055f rounds to 1 - 2^-24
    assert t[255] > 1.0
    # the check sees a one-ulp error in a single threshold
    bad = t.copy()
    bad[128] = np.nextafter(bad[128], np.float32(np.inf), dtype=np.float32)
    assert O.srgb8_check(bad)[0] == 1


def test_iso_fetch_saturates_without_alpha_test():
    """IsosurfaceShaderKernel::fetch: bytes = trunc(clamp(v * 255, 0, 255)) for every pixel
    (no alpha test, unlike StdVec4Pixel::write_to), depth passed through."""
    rng = np.random.default_rng(2)
    px = np.zeros((5, 7, 6))
    px[..., :4] = rng.uniform(-0.2, 1.2, (5, 7, 4)).astype(np.float32)
    px[..., 3] = 0.0
    px[..., 4] = rng.uniform(0, 3, (5, 7)).astype(np.float32)
    px[0, 0, 4] = np.inf
    val, depth = O.iso_fetch(px)
    want = np.floor(np.clip(px[..., :3].astype(np.float32) * np.float32(255), 0, 255)).astype(np.uint8)
    assert np.array_equal(val, want)
    assert np.array_equal(depth, px[..., 4].astype(np.float32))


def test_composite_iso_is_first_nearest():
    """Depth-min composite (isosurface.cc:272-289): each pixel takes the partial with the
    smallest depth; among equal depths the one earliest in z order; misses (+inf) lose to
    any hit, and a pixel no partial hit keeps the first partial's bytes."""
    rng = np.random.default_rng(5)
    N, H, W = 4, 6, 9
    val = rng.integers(0, 256, (N, H, W, 3), dtype=np.uint8)
    depth = rng.choice(np.float32([0.5, 1.0, 1.5, np.inf]), (N, H, W))
    depth[:, 0, 0] = np.inf
    z = [2, 0, 3, 1]
    rgb, d = O.composite_iso(val, depth, z)
    dz = depth[z]
    pick = np.argmin(dz, axis=0)          # first index of the minimum in z order
    src = np.take_along_axis(np.array(z)[:, None, None] * np.ones((1, H, W), int), pick[None], 0)[0]
    want = val[src, np.arange(H)[:, None], np.arange(W)[None, :]]
    assert np.array_equal(rgb, want)
    assert np.array_equal(d, dz.min(axis=0))
    assert np.array_equal(rgb[0, 0], val[2, 0, 0])
