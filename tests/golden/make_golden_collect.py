"""Writes tests/golden/reference_collect.json: the reference CPU library's
codeaut_bz_collect (oracle/_ref, backend cpu) on the code ensemble of the reference's
completeness test (tests/codeaut/test_codeaut.py:166-187 of the reference: numpy
default_rng(3), 40 draws of k ~ U{8..13}, n ~ U{k+1..2k+4}, B k x n Bernoulli(1/2), the
full-rank ones kept). Each code's generator is its systematic form (reduced row echelon
form of B, as a BZ information set); cases: p = 3 and 5 with keep = p (complete: every
codeword of weight <= p is the XOR of at most p rows of a systematic generator) and
keep = p + 3. Stored: inputs, row count, combination count, SHA-256 of the sorted rows.

    make -f oracle/ref.mk REF=/root/reference && python tests/golden/make_golden_collect.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from qubitserf_amd import codes  # noqa: E402
from tests import oracle as O  # noqa: E402
from tests.golden.fixtures import enc  # noqa: E402


def ensemble():
    """(index, k, n, systematic generator) of the full-rank draws, in draw order."""
    rng = np.random.default_rng(3)
    out = []
    for i in range(40):
        k = int(rng.integers(8, 14))
        n = int(rng.integers(k + 1, 2 * k + 5))
        B = rng.integers(0, 2, size=(k, n), dtype=np.uint8)
        if codes.gf2_rank(B) != k:
            continue
        out.append((i, k, n, codes.gf2_rref(B)))
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libqubitserf_ref.so missing")
    cases = []
    for i, k, n, G in ensemble():
        for p in (3, 5):
            for keep in (p, p + 3):
                rows, combos = O.ref_collect(G, p, keep)
                cases.append(dict(name=f"draw{i}_p{p}_keep{keep}", G=enc(G), p=p, keep=keep,
                                  ref=dict(count=int(rows.shape[0]), combos=int(combos),
                                           sha256=hashlib.sha256(np.packbits(rows, axis=1).tobytes()).hexdigest())))
    (HERE / "reference_collect.json").write_text(json.dumps(
        {"meta": {"source": "reference CPU library (oracle/_ref), codeaut_bz_collect, backend cpu"}, "collect": cases},
        indent=1))
    print(f"{len(cases)} cases")


if __name__ == "__main__":
    main()
