"""Batched solves (distfind_batch_distance): every problem's result equals the single-code
entry point's (distance, lower_bound, proven, levels) and the oracle's distance; a work cap
leaves valid bounds; a proposal's column order changes only the information sets."""
import numpy as np
import pytest

from qubitserf_amd import codes
from tests import oracle as O

pytestmark = pytest.mark.gpu


def _mixed_problems():
    rng = np.random.default_rng(5)
    probs = [("classical", codes.config1_parity(s)) for s in range(4)]
    probs += [("classical", codes.random_rate_code(rng, 40, 90)) for _ in range(10)]
    probs += [("classical", codes.hamming_parity(4)), ("classical", codes.repetition_parity(9))]
    probs += [("css", *codes.toric(5), "min"), ("css", *codes.surface(6), "z"), ("css", *codes.steane(), "x"),
              ("css", *codes.bb_family(6), "min"), ("css", *codes.toric(4), "x"), ("css", *codes.shor(), "min")]
    return probs


def _single(p):
    from qubitserf_amd import distfind as df
    if p[0] == "classical":
        return df.api.classical_distance_raw(p[1], "bz", "gpu")
    which = {"min": "M", "z": "Z", "x": "X"}[p[3]]
    return df.api.css_distance_raw(p[1], p[2], "bz", which, "gpu")


def _oracle(p):
    if p[0] == "classical":
        return O.classical_distance(p[1])
    return O.css_distance(p[1], p[2], which={"min": "M", "z": "Z", "x": "X"}[p[3]])


def test_batch_equals_single_and_oracle(hip):
    from qubitserf_amd import distfind as df
    probs = _mixed_problems()
    got = df.batch_distance(probs)
    st = df.last_batch_stats()
    assert st["capped"] == 0 and st["rounds"] >= 1
    for p, g in zip(probs, got):
        s = _single(p)
        o = _oracle(p)
        key = (g.distance, g.lower_bound, g.proven, g.levels)
        assert key == (s.distance, s.lower_bound, s.proven, s.levels), (p[0], key)
        assert (g.distance, g.proven) == (o.distance, o.proven)
        assert g.backend.decode() == "hip"
    # rounds: at most the longest single solve's step count, not the sum over problems
    assert st["rounds"] <= max(2 * g.levels for g in got)


def test_batch_mixed_widths(hip):
    """One batch holding codes with K <= 128 and K > 128 (the two tiled mask widths share
    the rounds): each result equals the single-code one and the oracle's."""
    from qubitserf_amd import distfind as df
    rng = np.random.default_rng(45)
    probs = [("classical", codes.random_rate_code(rng, 150, 220, 0.82, 0.9)) for _ in range(4)]
    probs += [("classical", codes.random_rate_code(rng, 60, 100)) for _ in range(4)]
    got = df.batch_distance(probs)
    for p, g in zip(probs, got):
        s = _single(p)
        o = _oracle(p)
        assert (g.distance, g.lower_bound, g.proven, g.levels) == (s.distance, s.lower_bound, s.proven, s.levels)
        assert (g.distance, g.proven) == (o.distance, o.proven), p[1].shape


def test_batch_repeatable(hip):
    from qubitserf_amd import distfind as df
    probs = _mixed_problems()[:8]
    a = [(r.distance, r.lower_bound, r.levels) for r in df.batch_distance(probs)]
    b = [(r.distance, r.lower_bound, r.levels) for r in df.batch_distance(probs[::-1])][::-1]
    assert a == b


def test_work_cap_gives_valid_bounds(hip):
    from qubitserf_amd import distfind as df
    rng = np.random.default_rng(17)
    probs = [("classical", codes.random_rate_code(rng, 70, 110)) for _ in range(8)]
    full = df.batch_distance(probs)
    capped = df.batch_distance(probs, work_cap=200_000)
    assert df.last_batch_stats()["capped"] >= 1
    for f, c in zip(full, capped):
        assert f.proven
        if c.proven:
            assert (c.distance, c.levels) == (f.distance, f.levels)
        else:
            assert c.lower_bound <= f.distance <= c.distance


def test_order_changes_sets_not_distance(hip):
    from qubitserf_amd import distfind as df
    rng = np.random.default_rng(23)
    probs, want = [], []
    for L in (5, 6):
        H = codes.quasi_cyclic(rng, L)
        n = H.shape[1]
        want.append(O.classical_distance(H).distance)
        probs.append(("classical", H, rng.permutation(n)))
        probs.append(("classical", H, np.arange(n)[::-1].copy()))
        want.append(want[-1])
    got = df.batch_distance(probs)
    assert [g.distance for g in got] == want
    assert all(g.proven for g in got)


def test_empty_and_degenerate(hip):
    from qubitserf_amd import distfind as df
    assert df.batch_distance([]) == []
    full = np.eye(6, dtype=np.uint8)              # k = 0: no codewords
    r = df.batch_distance([("classical", full), ("classical", codes.repetition_parity(5))])
    assert r[0].distance == -1 and r[1].distance == 5


def test_session_matches_batch(hip):
    """Continuous batching: problems added at different rounds finish with the same
    results as a one-shot batch."""
    from qubitserf_amd import distfind as df
    probs = _mixed_problems()
    want = [(r.distance, r.lower_bound, r.proven, r.levels) for r in df.batch_distance(probs)]
    s = df.Session()
    ids = s.add(probs[:5])
    got = {}
    rounds = 0
    pending = list(range(5, len(probs)))
    while s.active or pending:
        if pending and rounds % 2 == 1:
            ids += s.add([probs[pending.pop(0)]])
        if s.active:
            s.step()
        rounds += 1
        for i, r in s.poll():
            got[i] = (r.distance, r.lower_bound, r.proven, r.levels)
    assert [got[i] for i in ids] == want
    s.close()
