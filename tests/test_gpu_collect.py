"""Collect mode at scale on the device.

* The reference's completeness ensemble (tests/golden/reference_collect.json: its
  codeaut_bz_collect on the codes of tests/codeaut/test_codeaut.py:166-187, pinned against
  brute force by tests/test_cpu_collect_fixture.py): codeaut_bz_collect here returns the
  same rows in the same order (SHA-256) and the same combination count, for every case.
* qsf_hip_collect_upto (every level 1..d in one device pass, wave-aggregated appends, device
  radix sort and dedup) equals the sorted, deduplicated union of the per-level
  collections and of the oracle's, including collections larger than its first device
  buffer (the overflow re-run), on the tiled and row-sweep kernels."""
import hashlib

import numpy as np
import pytest

from tests import oracle as O
from tests.golden.fixtures import dec, load

pytestmark = pytest.mark.gpu
CASES = load("reference_collect.json")["collect"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_ensemble_matches_reference(hip, case):
    from qubitserf_amd.codeaut import bz_collect

    rows, combos = bz_collect(dec(case["G"]), case["p"], case["keep"])
    assert rows.shape[0] == case["ref"]["count"] and combos == case["ref"]["combos"]
    assert hashlib.sha256(np.packbits(rows, axis=1).tobytes()).hexdigest() == case["ref"]["sha256"]


def _lex_unique(a):
    a = np.ascontiguousarray(a, np.uint64)
    if a.shape[0] == 0:
        return a
    a = np.unique(a, axis=0)   # numpy orders rows lexicographically by column = by word, unsigned
    return a


@pytest.mark.parametrize("K,stride,d,keep,policy", [(40, 1, 4, 14, "auto"), (60, 2, 4, 40, "auto"),
                                                    (34, 2, 5, 48, "tiled"), (30, 1, 5, 12, "sweep"),
                                                    (200, 4, 2, 110, "auto")])
def test_collect_upto_equals_union_of_levels(hip, K, stride, d, keep, policy):
    rng = np.random.default_rng(K * 100 + d)
    n = 64 * stride
    gamma = rng.integers(0, 2**64, size=(1, K, stride), dtype=np.uint64)
    print(f"collect_upto K={K} stride={stride} d={d} keep={keep} {policy}", flush=True)
    with hip.kernel_policy(policy):
        got = hip.collect_upto(gamma, n, d, keep)
        st = hip.last_collect_stats()
        per = [hip.collect_level(gamma, n, dd, keep) for dd in range(1, d + 1)]
    want = _lex_unique(np.concatenate(per, axis=0))
    assert got.shape == want.shape and np.array_equal(got, want), (got.shape, want.shape)
    assert st["appended"] >= got.shape[0]
    if K <= 60:
        orc = _lex_unique(np.concatenate([O.collect_plan(gamma, n, dd, keep) for dd in range(1, d + 1)], axis=0))
        assert np.array_equal(got, orc)
    # rows are sorted strictly increasing (lexicographic by word, unsigned)
    for i in range(1, got.shape[0]):
        assert tuple(got[i - 1]) < tuple(got[i])


def test_collect_upto_overflow_rerun(hip):
    """More rows than the first device buffer (65536): the run repeats with room for all."""
    rng = np.random.default_rng(5)
    K, n = 24, 64
    gamma = rng.integers(0, 2**64, size=(1, K, 1), dtype=np.uint64)
    got = hip.collect_upto(gamma, n, 6, 64)   # every XOR of <= 6 rows: 190 050 distinct with high probability
    st = hip.last_collect_stats()
    from math import comb
    total = sum(comb(K, s) for s in range(1, 7))
    assert st["appended"] == total > 65536
    want = _lex_unique(np.concatenate([hip.collect_level(gamma, n, dd, 64) for dd in range(1, 7)], axis=0))
    assert np.array_equal(got, want)
