"""The distfind tool's text formats and command line, without a device: the parsers
(cpp/distfind/cli/distfind_cli.cpp:100-188 behaviour: blank lines, ragged rows, Y and mixed
stabilisers rejected, identity rows dropped) and the argument checks of the binary."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from qubitserf_amd import codes
from qubitserf_amd.distfind import formats as F

CLI = Path(__file__).resolve().parent.parent / "qubitserf_amd" / "_lib" / "distfind"


def test_matrix_roundtrip_and_blank_lines():
    H = codes.hamming_parity(3)
    text = "\n\n".join(" ".join(str(int(b)) for b in row) for row in H) + "\n\n"
    assert np.array_equal(F.parse_matrix_text(text), H)
    assert np.array_equal(F.parse_matrix_text("1 0 1\r\n0\t1 1\n"), np.array([[1, 0, 1], [0, 1, 1]], np.uint8))


@pytest.mark.parametrize("text,msg", [("1 0 1\n1 1\n", "ragged"), ("1 2 0\n", "non-binary"), ("\n \n", "empty"),
                                      ("10 1\n", "non-binary")])
def test_matrix_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        F.parse_matrix_text(text)


def test_pauli_split_and_roundtrip():
    Hx, Hz = codes.toric(3)
    gx, gz = F.parse_pauli_css(F.to_pauli_lines(Hx, Hz))
    assert np.array_equal(gx, Hx) and np.array_equal(gz, Hz)
    gx, gz = F.parse_pauli_css("X.X_\nz.Z.\nIIII\n\nXXXX\n")   # blank line ends; identity row dropped
    assert gx.tolist() == [[1, 0, 1, 0]] and gz.tolist() == [[1, 0, 1, 0]]


@pytest.mark.parametrize("text,msg", [("XY\n", "contains a Y"), ("XZ\n", "mixes X and Z"), ("XX\nZ\n", "differing lengths"),
                                      ("", "no stabilisers"), ("XQ\n", "unrecognised character")])
def test_pauli_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        F.parse_pauli_css(text)


def _run(args, stdin=""):
    return subprocess.run([str(CLI)] + args, input=stdin, capture_output=True, text=True, timeout=60)


def test_cli_usage_errors():
    assert _run(["-h"]).returncode == 0 and "Usage: distfind" in _run(["-h"]).stdout
    for args, msg in ((["--zx", "--z"], "--zx cannot be combined"), (["--hx", "a.txt"], "must be given together"),
                      (["--cc"], "not provided"), (["--mitm"], "not provided"), (["--cpu"], "reference tool"),
                      (["--threads", "x"], "non-negative integer"), (["--bogus"], "unrecognised argument")):
        r = _run(args)
        assert r.returncode == 2 and msg in r.stderr, (args, r.stderr)
