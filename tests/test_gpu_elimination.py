"""Device elimination (qsf_hip_eliminate) vs the oracle, bit-exact: the greedy
information-set peel and per-part systematic forms (generators, ranks, set counts), and
the random information-set upper bound (both sides draw trial t's column order from the
same stream). Covers the LDS and global-scratch layouts, many problems per launch,
batches of trials, detector rows, set-capacity overflow and empty inputs."""
import numpy as np
import pytest

from qubitserf_amd import codes
from qubitserf_amd.engine import pack_rows
from tests import oracle as O

pytestmark = pytest.mark.gpu


def _basis(rng, K, n):
    G = rng.integers(0, 2, size=(K, n), dtype=np.uint8)
    return pack_rows(G)


@pytest.mark.parametrize("K,n,trials", [(1, 3, 4), (5, 17, 256), (40, 100, 256), (64, 64, 31), (65, 200, 256),
                                        (130, 300, 64), (200, 1024, 16), (400, 1024, 8), (300, 2000, 4)])
def test_random_codes_match_oracle(hip, K, n, trials):
    rng = np.random.default_rng(K * 7919 + n)
    G = _basis(rng, K, n)     # not always full rank: dependent rows become zero rows, skipped
    got = hip.is_bound(G, n, trials=trials)
    assert got == O.is_bound_trials(G, n, trials=trials), (K, n, trials)
    chk = _basis(rng, 3, n)
    got = hip.is_bound(G, n, check=chk, trials=trials, seed=12345)
    assert got == O.is_bound_trials(G, n, check=chk, trials=trials, seed=12345), (K, n, trials)


def test_trial_batches(hip):
    """More trials than one launch holds: batches continue the trial numbering."""
    rng = np.random.default_rng(3)
    G = _basis(rng, 90, 400)     # 90 x 7 words: scratch layout, 1024 trials per launch
    assert hip.is_bound(G, 400, trials=1500) == O.is_bound_trials(G, 400, trials=1500)


def test_init_and_empty(hip):
    rng = np.random.default_rng(4)
    G = _basis(rng, 10, 50)
    assert hip.is_bound(G, 50, trials=64, init=1) == 1
    assert hip.is_bound(G, 50, trials=0, init=77) == 77
    assert hip.is_bound(np.zeros((0, 1), np.uint64), 50, trials=64, init=9) == 9
    assert hip.is_bound(np.zeros((3, 1), np.uint64), 50, trials=64) == O.WEIGHT_NONE


@pytest.mark.parametrize("name,HxHz", [("toric6", codes.toric(6)), ("bb72", codes.bb_family(6)),
                                       ("gross", codes.gross())])
def test_css_sides_match_oracle(hip, name, HxHz):
    for kind in (1, 2):
        gamma, ranks, check, even = O.side_plan(kind, *HxHz)
        n = HxHz[0].shape[1]
        assert hip.is_bound(gamma[0], n, check=check) == O.is_bound_trials(gamma[0], n, check=check), (name, kind)


def test_gross_bound_is_a_logical_weight(hip):
    """The gross code's bound is >= its distance 12 (a weight of some logical)."""
    gamma, ranks, check, even = O.side_plan(1, *codes.gross())
    b = hip.is_bound(gamma[0], 144, check=check)
    assert 12 <= b < O.WEIGHT_NONE


def _problem_cases():
    return [("toric6", 1, codes.toric(6)), ("surface5", 2, codes.surface(5)), ("bb72", 1, codes.bb_family(6)),
            ("gross_z", 1, codes.gross()), ("gross_x", 2, codes.gross()), ("qrm1_8", 1, codes.quantum_reed_muller(1, 8)),
            ("hamming5", 0, (codes.hamming_parity(5), None)), ("rep7", 0, (codes.repetition_parity(7), None))]


@pytest.mark.parametrize("name,kind,mats", _problem_cases(), ids=[c[0] for c in _problem_cases()])
def test_info_sets_match_oracle(hip, name, kind, mats):
    gen, check = O.problem_gen(kind, *mats)
    n = mats[0].shape[1]
    want_sets, want_ranks = O.greedy_sets(gen, n)
    got = hip.eliminate([dict(gen=gen, n=n, check=check)], what=hip.ELIM_SETS)[0]
    assert got["ranks"] == want_ranks, name
    assert np.array_equal(got["sets"], want_sets), name


def test_info_sets_random_and_batched(hip):
    """Many problems in one launch (LDS and scratch layouts mixed) equal one-by-one oracle runs."""
    rng = np.random.default_rng(21)
    probs = []
    for K, n in [(1, 9), (12, 30), (40, 100), (64, 64), (70, 300), (90, 150), (200, 1024), (400, 1000), (3, 4000)]:
        G = rng.integers(0, 2, size=(K, n), dtype=np.uint8)
        G[:, rng.integers(0, n, size=max(1, n // 10))] = 0      # some zero columns
        probs.append(dict(gen=pack_rows(G), n=n))
    got = hip.eliminate([dict(p) for p in probs], what=hip.ELIM_SETS)
    for p, g in zip(probs, got):
        ws, wr = O.greedy_sets(p["gen"], p["n"])
        assert g["ranks"] == wr and np.array_equal(g["sets"], ws), (p["gen"].shape, p["n"])


def test_info_sets_capacity(hip):
    G = pack_rows(np.ones((1, 20), np.uint8))    # repetition-like: one set per column
    ws, wr = O.greedy_sets(G, 20)
    got = hip.eliminate([dict(gen=G, n=20, set_cap=5)], what=hip.ELIM_SETS)[0]
    assert got["nsets"] == len(wr) == 20
    assert got["ranks"] == wr[:5] and np.array_equal(got["sets"], ws[:5])


def test_part_sets_match_oracle(hip):
    rng = np.random.default_rng(8)
    for K, n in [(10, 40), (50, 130), (300, 700)]:
        G = pack_rows(rng.integers(0, 2, size=(K, n), dtype=np.uint8))
        perm = rng.permutation(n).tolist()
        cuts = sorted(rng.choice(np.arange(1, n), size=3, replace=False).tolist())
        parts = [perm[a:b] for a, b in zip([0] + cuts, cuts + [n])]
        ws, wr = O.part_sets(G, n, parts)
        got = hip.eliminate([dict(gen=G, n=n, parts=parts)], what=hip.ELIM_SETS)[0]
        assert got["ranks"] == wr and np.array_equal(got["sets"], ws), (K, n)


def test_bounds_batched_match_oracle(hip):
    rng = np.random.default_rng(9)
    probs, want = [], []
    for K, n in [(5, 17), (40, 100), (120, 300), (400, 1024)]:
        G = _basis(rng, K, n)
        chk = _basis(rng, 2, n) if K % 2 == 0 else None
        probs.append(dict(gen=G, n=n, check=chk))
        want.append(O.is_bound_trials(G, n, check=chk, trials=40, seed=77))
    got = hip.eliminate(probs, what=hip.ELIM_BOUND, trials=40, seed=77)
    assert [g["best"] for g in got] == want
    st = hip.last_elim_stats()
    assert st["jobs"] == 4 and st["trials"] == 40 and st["bound_ms"] > 0
