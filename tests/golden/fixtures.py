"""Encoding of 0/1 matrices in the golden fixtures: {"shape": [m, n], "rows": [hex, ...]},
row r as the hex of the integer whose bit j is entry (r, j)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def enc(M) -> dict:
    M = np.asarray(M, dtype=np.uint8) & 1
    rows = []
    for r in M:
        v = 0
        for j in np.nonzero(r)[0]:
            v |= 1 << int(j)
        rows.append(format(v, "x"))
    return {"shape": [int(M.shape[0]), int(M.shape[1])], "rows": rows}


def dec(d) -> np.ndarray:
    m, n = d["shape"]
    M = np.zeros((m, n), dtype=np.uint8)
    for i, h in enumerate(d["rows"]):
        v = int(h, 16)
        for j in range(n):
            if (v >> j) & 1:
                M[i, j] = 1
    return M


def load(name: str = "reference_cases.json") -> dict:
    return json.loads((HERE / name).read_text())
