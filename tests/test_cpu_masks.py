"""The enumeration kernels' row-mask helpers (qubitserf_amd/csrc/masks.hpp), compiled for
the host and checked against brute force: colex unrank and successor on one- and two-word
masks (lists of up to 64 and 128 rows), bit ranges of 256-bit flag masks, and the set-bit
walk. The header is the same source the kernels include."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_mask_helpers(tmp_path):
    exe = tmp_path / "masks_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(ROOT / "qubitserf_amd" / "csrc"),
                    str(ROOT / "tests" / "masks_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "masks ok" in r.stdout
