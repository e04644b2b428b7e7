"""Train the scorer and compare proposals, on the device:

    python -m qubitserf_amd.proposal.fit --codes 2000 --out qubitserf_amd/proposal/weights.npz

Training codes: training_codes(count, seed) (n in [64, 200]); traces from capped batched
BZ runs of the natural and random column orders; targets and loss as in train.py.
Evaluation: quasi-cyclic codes (3 x 6 circulant blocks of row weight 3, circulant size
L_lo..L_hi, seeded), comparing, per code, the predicted BZ work of the natural peel, of the
best of `candidates` orders sampled from the trained scores, and of the best of as many
uniformly random orders. Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .. import codes, engine
from .model import MPScorer
from .propose import predicted_work
from .train import train, training_codes, traces


def family_work(Hs, orders_per_code, uppers):
    """Predicted work of each (code, order) peel, from the device's set ranks."""
    probs, meta = [], []
    for c, H in enumerate(Hs):
        gen = engine.pack_rows(codes.gf2_kernel(H))
        for o in orders_per_code[c]:
            probs.append(dict(gen=gen, n=H.shape[1], order=o, set_cap=H.shape[1] + 2))
            meta.append(c)
    Ks = [H.shape[1] - codes.gf2_rank(H) for H in Hs]
    out = [[] for _ in Hs]
    for s in range(0, len(probs), 512):
        res = engine.eliminate(probs[s:s + 512], what=engine.ELIM_SETS)
        for r, c in zip(res, meta[s:s + 512]):
            K = Ks[c]
            ranks = sorted(r["ranks"], reverse=True)
            out[c].append(predicted_work(K, ranks, False, uppers[c]))
    return out


def sample_orders(scores, count, rng):
    return [np.argsort(-(scores + rng.gumbel(size=scores.shape)), kind="stable").astype(np.int32) for _ in range(count)]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--codes", type=int, default=2000)
    ap.add_argument("--orders", type=int, default=8)
    ap.add_argument("--cap", type=float, default=1e6)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--L-lo", type=int, default=20)
    ap.add_argument("--L-hi", type=int, default=40)
    ap.add_argument("--candidates", type=int, default=16)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.time()
    Hs = training_codes(a.codes, seed=0)
    tr = traces(Hs, orders=a.orders, work_cap=int(a.cap))
    t_tr = time.time() - t0
    log(f"traces: {len(Hs)} codes x {a.orders} orders in {t_tr:.1f} s")
    targets = [y for y, _ in tr]
    sc = MPScorer(seed=0)
    t1 = time.time()
    hist = train(sc, Hs, targets, epochs=a.epochs, batch=32, lr=2e-3, log=log)
    t_fit = time.time() - t1
    if a.out:
        sc.save(a.out)
    # evaluation
    ev = codes.config4_eval(0, a.eval, a.L_lo, a.L_hi)
    from ..distfind import api as df
    ub = [r.distance for r in df.batch_distance([("classical", H) for H in ev], work_cap=1)]
    rng = np.random.default_rng(1)
    s_learned = sc.scores(ev)
    ords = []
    for H, s in zip(ev, s_learned):
        n = H.shape[1]
        ords.append([np.arange(n, dtype=np.int32)] + sample_orders(s, a.candidates, rng)
                    + sample_orders(np.zeros(n), a.candidates, rng))
    w = family_work(ev, ords, ub)
    rows = []
    for c, H in enumerate(ev):
        nat = w[c][0]
        lw = min([nat] + w[c][1:1 + a.candidates])
        rw = min([nat] + w[c][1 + a.candidates:])
        rows.append(dict(L=H.shape[1] // 6, upper=ub[c], natural=nat, learned=lw, random=rw))
    gm = lambda xs: float(np.exp(np.mean(np.log(np.maximum(1.0, xs)))))
    out = dict(train_codes=a.codes, orders=a.orders, cap=a.cap, traces_s=t_tr, fit_s=t_fit, loss=hist,
               eval=rows, geo_gain_learned=gm([r["natural"] / r["learned"] for r in rows]),
               geo_gain_random=gm([r["natural"] / r["random"] for r in rows]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
