"""Golden fixtures (tests/golden/, written by make_golden.py): the oracle reproduces
them exactly on the CPU, and the HIP path matches them on the GPU."""
import json
from pathlib import Path

import numpy as np
import pytest

from hydrant_amd import _lib as L
from hydrant_amd import dataset as D

HERE = Path(__file__).parent / "golden"
MAN = json.loads((HERE / "manifest.json").read_text())


def _scene(name):
    sys_path_fix = __import__("tests.golden.make_golden", fromlist=["scenes"])
    sc = sys_path_fix.scenes()[name]
    if sc.schedule == L.SCHED_LATTICE:
        sc.ert = 0.99
    return sc


@pytest.mark.parametrize("name", sorted(MAN["fixtures"]))
def test_oracle_reproduces_fixture(name):
    from tests.helpers import oracle_view, oracle_volume

    sc = _scene(name)
    assert D.sha256(sc.extra["volume"]) == MAN["fixtures"][name]["volume_sha256"]
    fx = np.load(HERE / f"{name}.npz")
    r = oracle_volume(sc).render(oracle_view(sc), 32, film=True)
    assert np.array_equal(r["rgba"].astype(np.float32), fx["rgba"])
    assert r["samples"] == int(fx["samples"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MAN["fixtures"]))
def test_gpu_matches_fixture(gpu_ctx, name):
    sc = _scene(name)
    fx = np.load(HERE / f"{name}.npz")
    gpu_ctx.load(sc.bricked, sc.tf)
    out = gpu_ctx.render_volume(sc.volume_params(), sc.view(), sc.W, sc.H, film=True, count=True)
    gpu_ctx.synchronize()
    err = float(np.max(np.abs(out["rgba"].cpu().numpy() - fx["rgba"])))
    print(f"{name}: max-abs vs fixture {err:.3g}", flush=True)
    assert err <= 1e-4
    assert int(out["samples"].item()) == int(fx["samples"])
