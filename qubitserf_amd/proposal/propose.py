"""Using the scorer: proposal items for the batched solver, and the work model that ranks
candidate information-set families (the same formula the C++ set-up uses to pick among a
proposal's candidates)."""
from __future__ import annotations

import math

import numpy as np


def predicted_work(K: int, ranks, even: bool, upper: int) -> float:
    """Codewords BZ enumerates before its lower bound reaches `upper`: sum over levels d of
    C(K, d) x (sets active at d), the bound after level d being sum_g max(0, d+1-(K-r_g)),
    rounded up to even for an even code (cpp/distfind/bz.cpp:142-162)."""
    work = 0
    for d in range(1, K + 1):
        act = sum(1 for r in ranks if (d + 1) - (K - r) > 0)
        outer = sum(max(0, (d + 1) - (K - r)) for r in ranks)
        if even:
            outer += outer & 1
        work += math.comb(K, d) * act
        if upper <= outer:
            break
    return float(work)


def choose_orders(Hs, scorer=None, candidates=8, seed=0):
    """Batch items for distfind.batch_distance: each classical code with the proposal's
    column scores (scorer=None: all-zero scores, i.e. uniformly random candidate orders,
    the proposal without the network)."""
    scores = scorer.scores(Hs) if scorer is not None else [np.zeros(H.shape[1], np.float32) for H in Hs]
    return [("classical", H, None, (s, candidates, seed + 1000003 * i)) for i, (H, s) in enumerate(zip(Hs, scores))]
