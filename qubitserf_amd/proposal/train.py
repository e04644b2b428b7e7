"""Training the scorer on traces of the enumerator itself.

For every training code, a few column orders (the natural one and random ones) each get a
capped batched BZ run (distfind.batch_distance with that order and a work cap): the lower
bound each run reaches within the cap is the order's gain. The device peel of each order
(qsf_hip_eliminate, set_of_col) says which columns formed its first information set. A
column's target is the gain-weighted share of orders whose first set took it:
    y_j = sum_o softmax(gain / tau)_o [j in first set of o]
and the scorer learns y_j by binary cross-entropy (forward, backward and Adam on the HIP
kernels). Columns a good order puts in its first set then score high, and the proposal's
orders put them first.
"""
from __future__ import annotations

import numpy as np

from .. import codes, engine
from ..distfind import api as df
from .model import MPScorer, pack_batch


def training_codes(count=2000, seed=0, n_lo=64, n_hi=200):
    """Half Bernoulli(1/2) codes (n ~ U{n_lo..n_hi}, rate ~ U[0.3, 0.6]), half quasi-cyclic
    codes (3 x 6 blocks of weight-3 circulants, n = 6L in the same range)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 2 == 0:
            out.append(codes.random_rate_code(rng, n_lo, n_hi))
        else:
            L = int(rng.integers(max(4, (n_lo + 5) // 6), n_hi // 6 + 1))
            out.append(codes.quasi_cyclic(rng, L))
    return out


def traces(Hs, orders=8, work_cap=1_000_000, seed=0, tau=1.0, chunk=1024, is_trials=16):
    """Per code: (targets y (n,), gains per order). Runs the capped solves in one batch."""
    rng = np.random.default_rng(seed + 7)
    items, meta = [], []
    for c, H in enumerate(Hs):
        n = H.shape[1]
        ords = [np.arange(n, dtype=np.int32)] + [rng.permutation(n).astype(np.int32) for _ in range(orders - 1)]
        for o in ords:
            items.append(("classical", H, o))
            meta.append((c, o))
    res = []
    for s in range(0, len(items), chunk):   # bounded batches: each holds its problems' device tables
        res += df.batch_distance(items[s:s + chunk], work_cap=work_cap, is_trials=is_trials)
    gens = [engine.pack_rows(codes.gf2_kernel(H)) for H in Hs]
    probs = [dict(gen=gens[c], n=Hs[c].shape[1], order=o, set_cap=1) for c, o in meta]
    sets = engine.eliminate(probs, what=engine.ELIM_SETS)
    out = []
    k = 0
    for c, H in enumerate(Hs):
        n = H.shape[1]
        gains = np.array([res[k + t].lower_bound for t in range(orders)], np.float64)
        w = np.exp((gains - gains.max()) / tau)
        w /= w.sum()
        y = np.zeros(n, np.float32)
        for t in range(orders):
            y += w[t] * (sets[k + t]["set_of_col"] == 0)
        out.append((y, gains))
        k += orders
    return out


def train(scorer: MPScorer, Hs, targets, epochs=4, batch=32, lr=2e-3, seed=0, log=None):
    """Minibatch Adam on the BCE of the column targets; returns the mean loss per epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    idx = np.arange(len(Hs))
    for ep in range(epochs):
        rng.shuffle(idx)
        tot, cnt = 0.0, 0
        for s in range(0, len(idx), batch):
            bi = idx[s:s + batch]
            X, Ac, Av, mask = pack_batch([Hs[i] for i in bi])
            T = np.zeros_like(mask)
            for r, i in enumerate(bi):
                T[r, : len(targets[i])] = targets[i]
            logits, cache = scorer.forward(X, Ac, Av)
            loss, ds = scorer.loss_grad(logits, T, mask)
            g = scorer.backward(cache, ds)
            scorer.adam(g, lr=lr)
            tot += float(loss.item()) * len(bi)
            cnt += len(bi)
        hist.append(tot / max(1, cnt))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4f}")
    return hist
