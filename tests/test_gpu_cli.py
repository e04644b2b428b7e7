"""The distfind tool end to end on the device: Pauli strings on stdin and 0/1 matrix
files, default / --z / --x / --zx outputs, against the known distances."""
import subprocess
from pathlib import Path

import pytest

from qubitserf_amd import codes
from qubitserf_amd.distfind import formats as F

pytestmark = pytest.mark.gpu
CLI = Path(__file__).resolve().parent.parent / "qubitserf_amd" / "_lib" / "distfind"


def _run(args, stdin=""):
    r = subprocess.run([str(CLI)] + args, input=stdin, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip()


def test_stdin_toric_and_surface(hip):
    assert _run([], F.to_pauli_lines(*codes.toric(4))) == "4"
    assert _run(["--zx"], F.to_pauli_lines(*codes.toric(3))) == "3 3"
    Hx, Hz = codes.surface(5)
    assert _run(["--z"], F.to_pauli_lines(Hx, Hz)) == "5"
    assert _run(["--x", "--gpu", "--bz", "--threads", "4"], F.to_pauli_lines(Hx, Hz)) == "5"


def test_matrix_files(hip, tmp_path):
    Hx, Hz = codes.bb_family(6)
    for name, H in (("hx.txt", Hx), ("hz.txt", Hz)):
        (tmp_path / name).write_text("\n".join(" ".join(str(int(b)) for b in row) for row in H) + "\n")
    assert _run(["--hx", str(tmp_path / "hx.txt"), "--hz", str(tmp_path / "hz.txt")]) == "6"
