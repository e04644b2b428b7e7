"""Random information-set upper bound on the device (qsf_hip_is_bound) vs its CPU
restatement in the oracle: bit-exact, since both draw trial t's column order from the same
stream. Covers the LDS and global-scratch layouts, batches of trials, detector rows, empty
inputs, and the bound's role in a solve (the distance never depends on it)."""
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
