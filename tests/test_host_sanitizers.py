"""Host code under AddressSanitizer and UBSan: the renderer-config JSON parser,
the culling / residency bookkeeping and the dataset preparation, driven by the
fuzzers in tests/fuzz/.
CPU only: the drivers make no GPU call."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
FUZZ = ROOT / "tests" / "fuzz"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _run(cmd, exe, tmp_path):
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert run.returncode == 0, f"exit {run.returncode}\n{run.stdout[-1000:]}\n{run.stderr[-3000:]}"
    return run.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_json_parser_fuzz_asan(tmp_path):
    exe = tmp_path / "json_fuzz"
    out = _run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                f"-I{ROOT / 'hydrant_amd' / 'csrc'}", str(FUZZ / "json_fuzz.cpp"), "-o", str(exe)], exe, tmp_path)
    assert out.startswith("ok ")


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_residency_fuzz_asan(tmp_path):
    exe = tmp_path / "residency_fuzz"
    out = _run([HIPCC, "-std=c++17", "-O1", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
                "-fsanitize=undefined", f"-I{ROOT / 'include'}", f"-I{ROOT / 'hydrant_amd' / 'csrc'}",
                str(FUZZ / "residency_fuzz.cpp"), str(ROOT / "hydrant_amd" / "csrc" / "residency.cpp"), "-o", str(exe)],
               exe, tmp_path)
    assert out.startswith("culled ")


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_dataset_fuzz_asan(tmp_path):
    exe = tmp_path / "dataset_fuzz"
    out = _run([HIPCC, "-std=c++17", "-O1", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
                "-fsanitize=undefined", f"-I{ROOT / 'include'}", f"-I{ROOT / 'hydrant_amd' / 'csrc'}",
                str(FUZZ / "dataset_fuzz.cpp"), str(ROOT / "hydrant_amd" / "csrc" / "synth.cpp"), "-o", str(exe)],
               exe, tmp_path)
    assert out.startswith("bricks ")
