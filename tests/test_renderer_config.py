"""The JSON renderer interface without a GPU: registry, RendererConfig and
CameraConfig parsing (include/hydrant_renderer.h), error behaviour."""
import ctypes as C

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import renderer as R

README_CONFIG = {
    "camera": {"orbit": {"arm": [25, 65, 2.2]}},
    "render": {
        "resolution": [512, 512],
        "renderer": "Volume",
        "params": {
            "mode": "Volume",
            "density": 1e-2,
            "transfer_fn": {"values": [[0, 0, 0, 0], [1, 0, 0, 1], [1, 0.5, 0, 1], [1, 1, 0, 1], [0, 1, 0, 1],
                                       [0, 1, 1, 1], [0, 0, 1, 1], [1, 0, 1, 1], [0, 0, 0, 0]]},
        },
    },
}


def test_registered_renderers():
    assert R.candidates() == ["Blocks", "Isosurface", "Paging", "Volume"]


def test_readme_config_and_renderer_config():
    assert R.parse_config(README_CONFIG) == ("Volume", 512, 512)
    assert R.parse_config(README_CONFIG["render"]) == ("Volume", 512, 512)
    assert R.parse_config({"renderer": "Blocks"}) == ("Blocks", 512, 512)   # default resolution
    assert R.parse_config({"renderer": "Isosurface", "resolution": [1024, 768]}) == ("Isosurface", 1024, 768)


def test_orbit_camera_matches_camera_call():
    cam = R.parse_camera(README_CONFIG)
    ref = L.Camera()
    arm = np.array([25, 65, 2.2], np.float32)
    L.call("hyd_orbit_camera", L.ptr(arm, C.c_float), None, 60.0, C.byref(ref))
    assert list(cam.position) == list(ref.position)
    assert list(cam.target) == list(ref.target) and list(cam.up) == list(ref.up)
    assert cam.fovy == 60.0
    # the closed form of the README camera (R_y(25) R_z(65) (2.2, 0, 0))
    assert np.allclose(list(cam.position), [0.842649, 1.993877, -0.392934], atol=1e-5)


def test_ptu_camera_and_fovy():
    cam = R.parse_camera({"ptu": {"position": [1, 2, 3]}, "perspective": {"fovy": 45}})
    assert list(cam.position) == [1.0, 2.0, 3.0]
    assert list(cam.target) == [0.0, 0.0, 0.0] and list(cam.up) == [0.0, 1.0, 0.0]
    assert cam.fovy == 45.0
    cam = R.parse_camera({"ptu": {"position": [0, 0, 4], "target": [0.5, 0.5, 0.5], "up": [0, 0, 1]}})
    assert list(cam.target) == [0.5, 0.5, 0.5] and list(cam.up) == [0.0, 0.0, 1.0]


@pytest.mark.parametrize("cfg, message", [
    ({"renderer": "Raytrace"}, b"unknown renderer"),
    ({"resolution": [512, 512]}, b"missing \"renderer\""),
    ({"renderer": "Volume", "resolution": [0, 512]}, b"resolution"),
    ({"renderer": "Volume", "resolution": [512]}, b"resolution"),
    ('{"renderer": "Volume",', b"JSON parse error"),
    ('{"renderer": "Volume"} trailing', b"trailing"),
    ('{"renderer": Volume}', b"expected a value"),
])
def test_config_errors(cfg, message):
    name = C.create_string_buffer(64)
    w, h = C.c_int32(), C.c_int32()
    text = cfg if isinstance(cfg, str) else __import__("json").dumps(cfg)
    code = L.lib().hyd_renderer_parse_config(text.encode(), name, 64, C.byref(w), C.byref(h))
    assert code != L.HYD_OK
    assert message in L.lib().hyd_last_error()


@pytest.mark.parametrize("cfg, message", [
    ({"orbit": {}}, b"missing arm"),
    ({"orbit": {"arm": [1, 2]}}, b"expected 3 numbers"),
    ({"ptu": {"target": [0, 0, 0]}}, b"missing position"),
    ({"perspective": {"fovy": 60}}, b"needs \"orbit\" or \"ptu\""),
    ({"orbit": {"arm": "far"}}, b"expected an array"),
])
def test_camera_errors(cfg, message):
    cam = L.Camera()
    code = L.lib().hyd_renderer_parse_camera(__import__("json").dumps(cfg).encode(), C.byref(cam))
    assert code == L.HYD_ERR_ARG
    assert message in L.lib().hyd_last_error()


def test_json_syntax():
    """Escapes, exponents, whitespace and nesting of the configuration parser."""
    text = ' {\n "render" : {"renderer": "Vol\\u0075me", "resolution": [6.4e1, 1.28E+2],\n'\
           ' "params": {"x": [true, false, null, {"y": "a\\"b\\\\c\\n"}]}}}\t'
    assert R.parse_config(text) == ("Volume", 64, 128)
    cam = R.parse_camera('{"orbit": {"arm": [2.5e1, 6.5E1, 22e-1]}}')
    assert np.allclose(list(cam.position), [0.842649, 1.993877, -0.392934], atol=1e-5)


@pytest.mark.parametrize("cam, message", [
    (dict(position=(1, 1, 1), target=(1, 1, 1)), "degenerate camera"),      # position on the target
    (dict(position=(0, 3, 0), up=(0, 1, 0)), "degenerate camera"),          # up along the view direction
    (dict(position=(0, 0, 3), fovy=0.0), "fovy"),
    (dict(position=(0, 0, 3), fovy=180.0), "fovy"),
])
def test_degenerate_views_are_refused(cam, message):
    from hydrant_amd.hip import make_view, ptu_camera

    with pytest.raises(L.HydError, match=message):
        make_view(ptu_camera(**cam), (0.5, 0.5, 0.5), (1, 1, 1))
    v = make_view(ptu_camera((0, 0, 3)), (0.5, 0.5, 0.5), (1, 1, 1))
    assert np.all(np.isfinite(list(v.trans)))
