"""Writes tests/golden/*.json: inputs and the reference's own answers for them.

The answers come from the reference's CPU library compiled from its sources by
oracle/ref.mk into oracle/_ref/libqubitserf_ref.so (distfind_* / codeaut_bz_* C entry
points, backend "cpu"). Run in the container that has the reference tree:

    make -f oracle/ref.mk REF=/root/reference && python tests/golden/make_golden.py

Matrices are stored row by row as hex strings of the packed little-endian bit row
(bit j of a row is bit j of the integer), with their shape.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from qubitserf_amd import codes  # noqa: E402
from tests import oracle as O  # noqa: E402
from tests.golden.fixtures import enc  # noqa: E402

HW = os.cpu_count() or 1   # the reference seeds its random information sets per hardware thread


def dist_record(r):
    return dict(distance=r.distance, lower_bound=r.lower_bound, proven=r.proven, levels=r.levels)


def classical_cases():
    out = []
    for seed in range(8):
        H = codes.config1_parity(seed)
        out.append(dict(name=f"config1_seed{seed}", H=enc(H)))
    rng = np.random.default_rng(3)
    for i in range(40):
        H = codes.random_rate_code(rng, 40, 100)
        out.append(dict(name=f"rate_n40_100_{i}", H=enc(H)))
    for n in range(3, 9):
        out.append(dict(name=f"rep{n}", H=enc(codes.repetition_parity(n))))
    for r in (3, 4):
        out.append(dict(name=f"hamming{r}", H=enc(codes.hamming_parity(r))))
    return out


def css_cases():
    fam = [("steane", codes.steane()), ("shor", codes.shor())]
    fam += [(f"toric{L}", codes.toric(L)) for L in range(3, 8)]
    fam += [(f"surface{L}", codes.surface(L)) for L in range(3, 8)]
    fam += [("bb_l6", codes.bb_family(6)), ("bb_l9", codes.bb_family(9))]
    fam += [("qrm1_6", codes.quantum_reed_muller(1, 6)), ("hgp_rep2_rep5",
            codes.hypergraph_product(codes.repetition_parity(2), codes.repetition_parity(5)))]
    return [dict(name=n, Hx=enc(Hx), Hz=enc(Hz)) for n, (Hx, Hz) in fam]


def five_qubit():
    z = np.array([0, 1, 1, 0, 0]); x = np.array([1, 0, 0, 1, 0])
    return np.array([np.concatenate([np.roll(z, s), np.roll(x, s)]) for s in range(4)], dtype=np.uint8)


def random_stabilizer(rng, q, m):
    """m commuting random Paulis on q qubits (isotropic subspace of the symplectic space),
    built greedily: a random vector is kept when it commutes with every kept one and is
    independent of them."""
    rows = []
    while len(rows) < m:
        v = rng.integers(0, 2, 2 * q, dtype=np.uint8)
        if not v.any():
            continue
        ok = all((int(v[:q] @ r[q:]) + int(v[q:] @ r[:q])) % 2 == 0 for r in rows)
        if ok and codes.gf2_rank(np.array(rows + [v])) == len(rows) + 1:
            rows.append(v)
    return np.array(rows, dtype=np.uint8)


def stabilizer_cases():
    out = [dict(name="five_qubit", S=enc(five_qubit()))]
    rng = np.random.default_rng(0x5AB)
    for i in range(8):
        q = int(rng.integers(6, 11))
        m = int(rng.integers(q // 2, q - 1))
        out.append(dict(name=f"random_isotropic_{i}_q{q}_m{m}", S=enc(random_stabilizer(rng, q, m))))
    return out


def collect_cases():
    out = []
    rng = np.random.default_rng(3)
    for i in range(12):
        k = int(rng.integers(8, 14))
        n = int(rng.integers(k + 4, 2 * k + 8))
        G = np.concatenate([np.eye(k, dtype=np.uint8), rng.integers(0, 2, (k, n - k), dtype=np.uint8)], axis=1)
        p = int(rng.integers(2, min(k, 6) + 1))
        keep = int(rng.integers(2, n // 2 + 2))
        out.append(dict(name=f"collect_{i}", G=enc(G), p=p, keep=keep))
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libqubitserf_ref.so missing: make -f oracle/ref.mk REF=/root/reference")
    from tests.golden.fixtures import dec

    meta = dict(source="reference CPU library (oracle/_ref), backend cpu", hw_threads=HW)
    t0 = time.time()
    cl = classical_cases()
    for c in cl:
        c["ref"] = dist_record(O.ref_classical_distance(dec(c["H"])))
    css = css_cases()
    for c in css:
        Hx, Hz = dec(c["Hx"]), dec(c["Hz"])
        c["ref"] = {w: dist_record(O.ref_css_distance(Hx, Hz, w)) for w in ("M", "Z", "X")}
    st = stabilizer_cases()
    for c in st:
        c["ref"] = dist_record(O.ref_stabilizer_distance(dec(c["S"])))
    co = collect_cases()
    for c in co:
        rows, combos = O.ref_collect(dec(c["G"]), c["p"], c["keep"])
        c["ref"] = dict(count=int(rows.shape[0]), combos=int(combos),
                        sha256=hashlib.sha256(np.packbits(rows, axis=1).tobytes()).hexdigest(),
                        rows=enc(rows) if rows.shape[0] <= 256 else None)
    data = dict(meta=meta, classical=cl, css=css, stabilizer=st, collect=co)
    (HERE / "reference_cases.json").write_text(json.dumps(data, indent=1))
    print(f"wrote {HERE / 'reference_cases.json'} in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
