"""Write the golden fixtures under tests/golden/.

The reference holds no test vectors for the render path (it has no tests), so
these fixtures are produced by the oracle on small seeded scenes.  They pin
regressions of the oracle and of the HIP path against a stored copy; they do
not pin correctness by themselves (test_oracle_pins.py does that).

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from hydrant_amd import _lib as L  # noqa: E402
from hydrant_amd import dataset as D  # noqa: E402
from hydrant_amd import workloads as W  # noqa: E402
from tests.helpers import oracle_view, oracle_volume  # noqa: E402


def scenes():
    return {
        "plain_n32_48x40": W.plain_volume(n=32, W=48, H=40),
        "plain_u8_n32_32x32_density1": W.plain_volume(n=32, W=32, H=32, fmt=L.FMT_U8, density=1.0),
        "paged_n128_bs32_64x64": W.paged_volume(n=128, W=64, H=64, bs=32, resident=0.4),
        "lattice_n96_bs32_48x48": W.paged_volume(n=96, W=48, H=48, bs=32, resident=1.0,
                                                 schedule=L.SCHED_LATTICE),
    }


def main():
    man = {"note": "oracle outputs (binary32) of small seeded scenes; sha256 of generator outputs",
           "inputs": {}, "fixtures": {}}
    man["inputs"]["blobs_256_seed1001"] = D.sha256(D.gen_blobs(256, 16, 1001, sig=(8.0, 40.0), sigma_scale=1.0,
                                                               amp=(0.3, 1.0)))
    man["inputs"]["perlin_128_seed2002"] = D.sha256(D.gen_perlin(128, 5, 0.5, 16.0, 2.0, 2002))
    man["inputs"]["sdf_128_seed3003"] = D.sha256(D.gen_sdf_spheres(128, 64, (2.5, 15.0), 2.0, 3003))
    for name, sc in scenes().items():
        if sc.schedule == L.SCHED_LATTICE:
            sc.ert = 0.99
        r = oracle_volume(sc).render(oracle_view(sc), 32, film=True)
        rgba = r["rgba"].astype(np.float32)
        np.savez_compressed(HERE / f"{name}.npz", rgba=rgba, samples=np.int64(r["samples"]),
                            nsteps=r["film"][..., 6].astype(np.int32))
        man["fixtures"][name] = {"samples": int(r["samples"]), "rgba_sha256": D.sha256(rgba),
                                 "volume_sha256": D.sha256(sc.extra["volume"]), "W": sc.W, "H": sc.H}
    (HERE / "manifest.json").write_text(json.dumps(man, indent=1, sort_keys=True) + "\n")
    print(json.dumps(man, indent=1))


if __name__ == "__main__":
    main()
