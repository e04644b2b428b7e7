"""The reference's level schedule can certify a distance above the true one, and the
sound schedule does not (oracle on the CPU).

The reference enumerates an information set of rank r < K only at levels d with
(d+1) - (K-r) > 0 (cpp/distfind/bz.cpp:142-144) but adds max(0, (d+1) - (K-r)) to the
lower bound at every level (bz.cpp:159-161). That contribution is valid only if every
combination of at most d rows of the set was enumerated, including levels 1..K-r-1, which
were skipped. The batched workload's code 32 (random [94, 49] code, seed 0) has its
second set at rank 45 (K - r = 4): level 3 of that set holds a codeword of weight 11, the
reference certifies 12, and the sound schedule certifies 11.
"""
import numpy as np
import pytest

from qubitserf_amd import codes, engine
from tests import oracle as O


def _code32():
    return codes.config3_batch(0, 256)[32]


def test_reference_certificate_refuted_on_code_32():
    H = _code32()
    n = H.shape[1]
    gamma, ranks, check, _ = O.side_plan(0, H)
    assert ranks == [49, 45]
    # a codeword of weight 11 among the 3-row combinations of the second set
    rows = O.collect_plan(gamma[1:2], n, 3, 11)
    assert rows.shape[0] >= 1
    x = engine.unpack_rows(rows[:1], n)[0]
    assert x.sum() == 11 and not (H.astype(np.int64) @ x % 2).any()   # H x = 0: a codeword
    # the reference's schedule (oracle restatement) certifies 12 ...
    r = O.classical_distance(H)
    assert (r.distance, r.proven) == (12, 1)
    if O.ref_available():   # ... as the reference's own library does
        rr = O.ref_classical_distance(H)
        assert (rr.distance, rr.proven) == (12, 1)
    # the sound schedule enumerates the second set from level 1 and certifies 11
    with O.sound():
        s = O.classical_distance(H)
    assert (s.distance, s.proven) == (11, 1), (s.distance, s.lower_bound, s.proven, s.levels)


@pytest.mark.parametrize("seed", range(4))
def test_sound_schedule_matches_brute_force(seed):
    """Random codes with rank-deficient information sets: the sound schedule's certified
    distance equals exhaustive enumeration over all 2^K messages."""
    rng = np.random.default_rng(700 + seed)
    checked = deficient = 0
    while checked < 12:
        n = int(rng.integers(18, 30))
        m = int(rng.integers(n // 3, n // 2 + 2))
        H = codes.random_full_rank(rng, m, n)
        K = n - m
        if K > 18:
            continue
        _, ranks, _, _ = O.side_plan(0, H)
        deficient += any(r < K for r in ranks)
        with O.sound():
            s = O.classical_distance(H)
        b = O.bruteforce_distance(0, H)
        assert s.proven and s.distance == b, (seed, checked, s.distance, b, ranks)
        checked += 1
    assert deficient > 0


def test_sound_schedule_agrees_with_reference_where_reference_holds():
    """On the golden classical cases the sound schedule certifies the reference's distance
    (none of them has a lighter codeword in a skipped level)."""
    from tests.golden.fixtures import dec, load

    for case in load()["classical"]:
        H = dec(case["H"])
        with O.sound():
            s = O.classical_distance(H)
        assert (s.distance, s.proven) == (case["ref"]["distance"], case["ref"]["proven"]), case["name"]
