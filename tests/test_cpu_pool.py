"""The engine's host worker pool (qubitserf_amd/csrc/pool.hpp), compiled for the host and
checked: every index of every job runs exactly once, exceptions come back (the first by
index) after every index ran, and concurrent callers are served one job at a time. The
header is the same source the library includes."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_pool(tmp_path):
    exe = tmp_path / "pool_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I", str(ROOT / "qubitserf_amd" / "csrc"),
                    str(ROOT / "tests" / "pool_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "pool ok" in r.stdout
