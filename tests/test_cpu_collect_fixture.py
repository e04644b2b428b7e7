"""The collect ensemble fixture (tests/golden/reference_collect.json, the reference's
codeaut_bz_collect on the code ensemble of its completeness test) against brute force:
with a systematic generator and keep = p the collection is every codeword of weight 1..p,
exactly; with keep = p + 3 it lies between the weight <= p codewords and the weight <= keep
ones. Rows in the reference's order: lexicographic over the packed u64 words."""
import hashlib

import numpy as np
import pytest

from tests.golden.fixtures import dec, load

CASES = load("reference_collect.json")["collect"]


def _codewords(G):
    k, n = G.shape
    idx = np.arange(1, 1 << k, dtype=np.int64)
    U = ((idx[:, None] >> np.arange(k)) & 1).astype(np.uint8)
    return (U @ G.astype(np.int64)) % 2


def _sorted_packed(C):
    """Rows ordered by their packed word (bit j = column j; n <= 64 here), as packbits bytes."""
    key = (C.astype(np.uint64) << np.arange(C.shape[1], dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    C = C[np.argsort(key, kind="stable")]
    return np.packbits(C.astype(np.uint8), axis=1)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_fixture_against_brute_force(case):
    G = dec(case["G"])
    assert G.shape[1] <= 64
    C = np.unique(_codewords(G), axis=0)
    w = C.sum(axis=1)
    p, keep = case["p"], case["keep"]
    exact = C[(w >= 1) & (w <= p)]
    if keep == p:
        got = hashlib.sha256(_sorted_packed(exact).tobytes()).hexdigest()
        assert case["ref"]["count"] == exact.shape[0] and case["ref"]["sha256"] == got
    else:
        assert exact.shape[0] <= case["ref"]["count"] <= int(((w >= 1) & (w <= keep)).sum())
    k = G.shape[0]
    from math import comb
    assert case["ref"]["combos"] == sum(comb(k, s) for s in range(1, min(p, k) + 1))
