"""Text formats of the distfind tool (cpp/distfind/cli/distfind_cli.cpp:100-188), parsed
by the library (distfind_parse_matrix_text / distfind_parse_pauli_css)."""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native


def _lib():
    L = _native.load()
    if not getattr(L, "_fmt_bound", False):
        ip = ctypes.POINTER(ctypes.c_int)
        L.distfind_parse_matrix_text.argtypes = [ctypes.c_char_p, _native.u8p, ctypes.c_int64, ip, ip]
        L.distfind_parse_matrix_text.restype = ctypes.c_int
        L.distfind_parse_pauli_css.argtypes = [ctypes.c_char_p, _native.u8p, _native.u8p, ctypes.c_int64, ip, ip, ip]
        L.distfind_parse_pauli_css.restype = ctypes.c_int
        L._fmt_bound = True
    return L


def parse_matrix_text(text: str) -> np.ndarray:
    """Whitespace-separated 0/1 rows (blank lines skipped) -> uint8 matrix; ValueError on
    non-binary tokens, ragged rows or an empty matrix."""
    L = _lib()
    raw = text.encode()
    buf = np.zeros(len(raw) + 1, np.uint8)
    r, c = ctypes.c_int(), ctypes.c_int()
    if L.distfind_parse_matrix_text(raw, buf.ctypes.data_as(_native.u8p), buf.size, ctypes.byref(r), ctypes.byref(c)):
        raise ValueError(_native.last_error())
    return buf[: r.value * c.value].reshape(r.value, c.value).copy()


def parse_pauli_css(text: str):
    """Pauli stabiliser lines (I . _ space, X, Z; blank line ends) -> (Hx, Hz) uint8;
    ValueError for Y, mixed X/Z lines, differing lengths or no stabilisers."""
    L = _lib()
    raw = text.encode()
    hx = np.zeros(len(raw) + 1, np.uint8)
    hz = np.zeros(len(raw) + 1, np.uint8)
    a, b, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if L.distfind_parse_pauli_css(raw, hx.ctypes.data_as(_native.u8p), hz.ctypes.data_as(_native.u8p), hx.size,
                                  ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)):
        raise ValueError(_native.last_error())
    return (hx[: a.value * n.value].reshape(a.value, n.value).copy(),
            hz[: b.value * n.value].reshape(b.value, n.value).copy())


def to_pauli_lines(Hx, Hz) -> str:
    """The inverse: X-type rows then Z-type rows as Pauli strings."""
    out = []
    for H, ch in ((Hx, "X"), (Hz, "Z")):
        for row in np.asarray(H) & 1:
            out.append("".join(ch if b else "I" for b in row))
    return "\n".join(out) + "\n"
