"""Lookahead rounds (QSA_LOOKAHEAD=1: every level a problem can still need, in one batched
device call, each level on its own bound, replayed in the level loop's order) against
the default one-level rounds: identical distance, lower bound, proven flag and level count
on every problem, under a work cap and without one, for classical batches and for CSS
pairs (both sides interleaved)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _key(r):
    return (r.distance, r.lower_bound, r.proven, r.levels)


def _both(problems, **kw):
    from qubitserf_amd import distfind as df

    old = os.environ.get("QSA_LOOKAHEAD")
    try:
        os.environ["QSA_LOOKAHEAD"] = "0"
        plain = df.batch_distance(problems, **kw)
        os.environ["QSA_LOOKAHEAD"] = "1"
        ahead = df.batch_distance(problems, **kw)
    finally:
        if old is None:
            os.environ.pop("QSA_LOOKAHEAD", None)
        else:
            os.environ["QSA_LOOKAHEAD"] = old
    return plain, ahead


@pytest.mark.parametrize("cap", [int(1e9), None])
def test_lookahead_classical_batch_equals_level_rounds(hip, cap):
    from qubitserf_amd import codes

    if cap:   # n ~ U{64..160}: most stop at the cap
        Hs = codes.config3_batch(3, 48)
    else:     # n ~ U{40..72}: every code certifies within seconds
        rng = np.random.default_rng(11)
        Hs = [codes.random_rate_code(rng, 40, 72) for _ in range(32)]
    print(f"lookahead: {len(Hs)} classical codes, cap {cap}", flush=True)
    plain, ahead = _both([("classical", H) for H in Hs], **({"work_cap": cap} if cap else {}))
    for i, (a, b) in enumerate(zip(plain, ahead)):
        assert _key(a) == _key(b), (i, _key(a), _key(b))
    assert sum(r.proven for r in plain) >= 1


def test_lookahead_css_pairs_equal_level_rounds(hip):
    from qubitserf_amd import codes

    probs = []
    for l in (6, 9):
        Hx, Hz = codes.bb_family(l)
        for which in ("min", "x", "z"):
            probs.append(("css", Hx, Hz, which))
    rng = np.random.default_rng(7)
    for _ in range(6):
        probs.append(("classical", codes.random_rate_code(rng, 40, 64)))
    print(f"lookahead: {len(probs)} mixed problems", flush=True)
    plain, ahead = _both(probs)
    for i, (a, b) in enumerate(zip(plain, ahead)):
        assert _key(a) == _key(b), (i, _key(a), _key(b))
