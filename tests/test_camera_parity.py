"""The library's host camera math against the oracle's binary32 restatement, bit for bit.

Every GPU parity test hands the library's view to the oracle, so the product's
matrices are checked here, on the CPU, against the oracle's own: orbit position
(Camera::update_params, scene.hpp:68-81), Exhibit::get_iet (scene.hpp:20-29), glm's
inverse, and the ViewArgs of Raycaster::fill_ray_emit_args (raycaster.hpp:126-135),
for orbit and position/target/up cameras.  The oracle's FP32 camera is itself pinned
by closed forms in test_oracle_pins.py.
"""
import ctypes as C

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import hip
from oracle import oracle as O


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _same(product, oracle_fp32):
    """oracle_fp32 holds binary32 values widened to double: compare the bit patterns."""
    o = np.asarray(oracle_fp32, np.float64)
    assert np.array_equal(o.astype(np.float32).astype(np.float64), o), "oracle value is not binary32"
    return np.array_equal(_bits(product), _bits(o))


ARMS = [(25, 65, 2.2), (0, 0, 3.0), (130, -40, 1.5), (-75, 10, 4.0), (359.5, 89.0, 2.2), (12.3, -79.9, 2.7)]


@pytest.mark.parametrize("arm", ARMS)
def test_orbit_position_bit_equal(arm):
    cam = hip.orbit_camera(arm)
    assert _same(cam.position[:], O.orbit_position(arm, 32))
    assert list(cam.target[:]) == [0, 0, 0] and list(cam.up[:]) == [0, 1, 0]


def _exhibits():
    return [((0.5, 0.5, 0.5), (1, 1, 1)), ((2, 2, 2), (4, 4, 4)), ((8, 4, 2), (16, 8, 4)),
            ((6, 8, 6), (12, 16, 12)), ((0.75, 1.25, 0.5), (1.5, 2.5, 1.0))]


@pytest.mark.parametrize("ex", _exhibits())
def test_exhibit_iet_bit_equal(ex):
    c = np.asarray(ex[0], np.float32)
    s = np.asarray(ex[1], np.float32)
    out = np.zeros(16, np.float32)
    L.call("hyd_exhibit_iet", L.ptr(c, C.c_float), L.ptr(s, C.c_float), L.ptr(out, C.c_float))
    assert _same(out, O.exhibit_iet(c.astype(np.float64), s.astype(np.float64), 32))


def test_inverse_bit_equal():
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = (rng.normal(size=16) + 3 * np.eye(4).reshape(-1)).astype(np.float32)
        out = np.zeros(16, np.float32)
        L.call("hyd_mat4_inverse", L.ptr(m, C.c_float), L.ptr(out, C.c_float))
        assert _same(out, O.inverse(m.astype(np.float64), 32))


def _check_view(cam, ex):
    v = hip.make_view(cam, ex[0], ex[1])
    pos = np.asarray(cam.position[:], np.float32).astype(np.float64)
    tgt = np.asarray(cam.target[:], np.float32).astype(np.float64)
    up = np.asarray(cam.up[:], np.float32).astype(np.float64)
    o = O.make_view(pos, tgt, up, float(np.float32(cam.fovy)), np.asarray(ex[0], np.float32).astype(np.float64),
                    np.asarray(ex[1], np.float32).astype(np.float64), 32)
    assert _same(v.trans[:], o.trans[:])
    assert _same([v.ctg_fovy_2], [o.ctg_fovy_2])
    assert _same(v.ray_o[:], o.ray_o[:])


@pytest.mark.parametrize("arm", ARMS)
@pytest.mark.parametrize("ex", _exhibits()[:3])
def test_orbit_view_bit_equal(arm, ex):
    _check_view(hip.orbit_camera(arm), ex)


def test_ptu_views_bit_equal():
    """Position / target / up cameras (CameraPosTargetUp), several fovy, seeded."""
    rng = np.random.default_rng(8)
    n = 0
    for _ in range(200):
        pos = rng.uniform(-5, 5, 3)
        tgt = rng.uniform(-1, 1, 3)
        up = rng.normal(size=3)
        f = pos - tgt
        if np.linalg.norm(f) < 0.5 or np.linalg.norm(np.cross(f, up)) < 0.1 * np.linalg.norm(f) * np.linalg.norm(up):
            continue
        fovy = float(rng.choice([30.0, 45.0, 60.0, 75.5, 120.0]))
        ex = _exhibits()[int(rng.integers(0, 5))]
        _check_view(hip.ptu_camera(pos, tgt, up, fovy), ex)
        n += 1
    assert n > 150
