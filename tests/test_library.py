"""The C-ABI library loads and exports every symbol the public headers declare, and
the host wrappers check their arguments (no compute calls: this runs without a GPU)."""
import re
from pathlib import Path

import pytest

from hydrant_amd import _lib as L

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_symbols():
    names = set()
    for h in INCLUDE.glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hyd_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_symbols():
    names = declared_symbols()
    assert {"hyd_create", "hyd_render_volume", "hyd_composite", "hyd_pack_bricks"} <= names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_headers():
    assert declared_symbols() <= set(L.SIGNATURES), declared_symbols() - set(L.SIGNATURES)


def test_host_only_calls_without_gpu():
    import ctypes as C
    import numpy as np

    cam = L.Camera()
    arm = np.array([25, 65, 2.2], np.float32)
    L.call("hyd_orbit_camera", L.ptr(arm, C.c_float), None, 60.0, C.byref(cam))
    assert abs(cam.position[0] - 0.842649) < 1e-5
    assert L.lib().hyd_version() == 1
    code = L.lib().hyd_orbit_camera(None, None, 60.0, None)
    assert code == L.HYD_ERR_ARG and b"null" in L.lib().hyd_last_error()


def test_diff_wrapper_rejects_mismatched_tensors():
    """The differentiable wrapper checks tensor sizes against its parameters before any
    device call (a short volume, a malformed TF, or wrong checkpoint / gradient shapes
    would otherwise reach the kernels)."""
    import torch

    from hydrant_amd import _lib as L
    from hydrant_amd.diff import DiffRenderer

    view = L.View()
    r = DiffRenderer(object(), [view, view], 16, 8, (4, 5, 6), n_steps=32, ckpt_every=8)
    vol, tf = torch.zeros(6, 5, 4), torch.zeros(16, 4)
    ck = torch.zeros(2, 8, 16, r.nck)
    g = torch.zeros(2, 8, 16, 4)
    with pytest.raises(ValueError, match="volume"):
        r.forward_raw(torch.zeros(6, 5, 3), tf)
    with pytest.raises(ValueError, match="tf"):
        r.forward_raw(vol, torch.zeros(16, 3))
    with pytest.raises(ValueError, match="tf"):
        r.forward_raw(vol, torch.zeros(300, 4))
    with pytest.raises(ValueError, match="checkpoints"):
        r.backward_raw(vol, tf, ck[:, :, :8], g)
    with pytest.raises(ValueError, match="grad_rgba"):
        r.backward_raw(vol, tf, ck, g[:1])
    with pytest.raises(ValueError, match="float32"):
        r.backward_raw(vol, tf, ck, g.half())
    with pytest.raises(ValueError, match="float32"):
        r.backward_raw(vol, tf, ck, g.double())


def test_output_check():
    """The wrappers refuse a caller's output buffer that is too small, of the wrong type or
    in host memory before it reaches a kernel."""
    import torch

    from hydrant_amd.hip import check_output

    with pytest.raises(ValueError, match="at least 400"):
        check_output(torch.zeros(99, 4), 400, "batched output")
    with pytest.raises(ValueError, match="float32"):
        check_output(torch.zeros(100, 4, dtype=torch.float64), 400, "batched output")
    with pytest.raises(ValueError, match="device"):
        check_output(torch.zeros(100, 4), 400, "batched output")


def test_image_input_check():
    """The film, RGBA and partial-image inputs of the resume, pixel, fetch and composite
    passes are checked for type, layout, device and size before they reach a kernel."""
    import torch

    from hydrant_amd.hip import check_image

    with pytest.raises(ValueError, match="shape"):
        check_image(torch.zeros(8, 16, 15), 16, "film", "cuda:0")
    with pytest.raises(ValueError, match="float32"):
        check_image(torch.zeros(8, 16, 16, dtype=torch.float16), 16, "film", "cuda:0")
    with pytest.raises(ValueError, match="contiguous"):
        check_image(torch.zeros(16, 8, 16).transpose(0, 1), 16, "film", "cuda:0")
    with pytest.raises(ValueError, match="device"):
        check_image(torch.zeros(8, 16, 16), 16, "film", "cuda:0")
    with pytest.raises(ValueError, match="shape"):
        check_image(torch.zeros(8, 16), 4, "rgba", "cuda:0")
