"""Low-weight codeword collection on the gfx950 engine (the reference's codeaut_bz_collect,
cpp/codeaut/bz.cpp:58-136, called from python/qubitserf/codeaut/lowweight.py:54-78)."""
from __future__ import annotations

import numpy as np

from .. import _native


def bz_collect(G, p: int, keep_weight: int, backend: int = 1):
    """Every distinct codeword of weight in [1, keep_weight] that is the XOR of 1..p rows of
    G, as a (count, n) uint8 array sorted lexicographically by packed words; plus the
    number of subset combinations evaluated. Raises if the engine could not complete."""
    L = _native.load()
    a = np.ascontiguousarray(np.asarray(G, dtype=np.uint8) & 1)
    m, n = a.shape
    h = L.codeaut_bz_collect(a.ctypes.data_as(_native.u8p), m, n, int(p), int(keep_weight), int(backend), 0)
    try:
        if not L.codeaut_bz_ok(h):
            raise RuntimeError("codeaut_bz_collect failed: " + _native.last_error())
        cnt = L.codeaut_bz_count(h)
        out = np.zeros((cnt, n), dtype=np.uint8)
        if cnt:
            L.codeaut_bz_copy_rows(h, out.ctypes.data_as(_native.u8p))
        return out, int(L.codeaut_bz_combos(h))
    finally:
        L.codeaut_bz_free(h)
