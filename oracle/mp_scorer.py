"""ORACLE — test infrastructure only. CPU restatement (numpy) of the learned proposal's
Tanner-graph message-passing scorer, forward and backward, used by tests/ as the checker
of the device kernels (qubitserf_amd/csrc/mp_hip.hip). Nothing in the product imports it.

The network has no counterpart in the reference (which builds information sets greedily,
cpp/distfind/css.cpp:177-203); its definition is the product's own
(qubitserf_amd/proposal/model.py), restated here:

    Ac = H / deg_c (rows),  Av = H^T / deg_v (rows)           adjacency, per code
    h0 = relu(X Win + bin)
    for l = 1..3:
        cm = Ac h_{l-1};  cg = relu(cm Wc_l + bc_l)
        va = Av cg;       h_l = relu(h_{l-1} Ws_l + va Wn_l + b_l)
    s = h3 wo + bo        (one logit per column)
    loss = mean over real columns of BCE(s, target)

With bf16=True every GEMM operand is rounded to bfloat16 (round to nearest even) where the
device stores it in BF16, and GEMMs accumulate in float64 (the device: float32), so the two
agree to accumulation-order and BF16-tie noise. With bf16=False it is plain float64 math,
which tests pin by finite differences.
"""
from __future__ import annotations

import numpy as np

LAYERS = 3


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def param_names():
    names = ["Win", "bin"]
    for l in range(LAYERS):
        names += [f"Wc{l}", f"bc{l}", f"Ws{l}", f"Wn{l}", f"b{l}"]
    return names + ["wo", "bo"]


def _q(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(params, X, Ac, Av, bf=True):
    """X: (B, n, F); Ac: (B, m, n); Av: (B, n, m). Returns (logits (B, n), cache)."""
    P = {k: _q(v, bf) for k, v in params.items() if k.startswith(("W", "w"))}
    P.update({k: np.asarray(v, np.float64) for k, v in params.items() if k.startswith("b")})
    X, Ac, Av = _q(X, bf), _q(Ac, bf), _q(Av, bf)
    relu = lambda z: np.maximum(z, 0.0)
    c = {"X": X, "Ac": Ac, "Av": Av, "P": P}
    h = _q(relu(X @ P["Win"] + P["bin"]), bf)
    c["h0"] = h
    for l in range(LAYERS):
        cm = _q(Ac @ h, bf)
        cg = _q(relu(cm @ P[f"Wc{l}"] + P[f"bc{l}"]), bf)
        va = _q(Av @ cg, bf)
        hn = _q(relu(h @ P[f"Ws{l}"] + va @ P[f"Wn{l}"] + P[f"b{l}"]), bf)
        c[f"cm{l}"], c[f"cg{l}"], c[f"va{l}"], c[f"h{l + 1}"] = cm, cg, va, hn
        h = hn
    s = (h @ P["wo"])[..., 0] + P["bo"][0]
    c["s"] = s
    return s, c


def loss_grad(s, target, mask):
    """Mean masked BCE with logits and its gradient."""
    cnt = max(1.0, float(mask.sum()))
    p = 1.0 / (1.0 + np.exp(-s))
    loss = float((mask * (np.maximum(s, 0) - s * target + np.log1p(np.exp(-np.abs(s))))).sum() / cnt)
    return loss, mask * (p - target) / cnt


def backward(c, ds, bf=True):
    """Gradients of every parameter given d loss / d logits (B, n)."""
    P = c["P"]
    g = {}
    flat = lambda a: a.reshape(-1, a.shape[-1])
    h = c[f"h{LAYERS}"]
    dsq = _q(ds[..., None], bf)
    g["wo"] = flat(h).T @ flat(dsq)
    g["bo"] = np.array([ds.sum()])
    dh = dsq @ P["wo"].T
    for l in reversed(range(LAYERS)):
        hn, hp = c[f"h{l + 1}"], c[f"h{l}"]
        dpre = dh * (hn > 0)
        g[f"b{l}"] = flat(dpre).sum(0)
        dpq = _q(dpre, bf)
        g[f"Ws{l}"] = flat(hp).T @ flat(dpq)
        g[f"Wn{l}"] = flat(c[f"va{l}"]).T @ flat(dpq)
        dh_prev = dpq @ P[f"Ws{l}"].T
        dva = _q(dpq @ P[f"Wn{l}"].T, bf)
        dcg = np.swapaxes(c["Av"], 1, 2) @ dva
        dcm = dcg * (c[f"cg{l}"] > 0)
        g[f"bc{l}"] = flat(dcm).sum(0)
        dcq = _q(dcm, bf)
        g[f"Wc{l}"] = flat(c[f"cm{l}"]).T @ flat(dcq)
        dcm_in = _q(dcq @ P[f"Wc{l}"].T, bf)
        dh = dh_prev + np.swapaxes(c["Ac"], 1, 2) @ dcm_in
    dpre = dh * (c["h0"] > 0)
    g["bin"] = flat(dpre).sum(0)
    g["Win"] = flat(c["X"]).T @ flat(_q(dpre, bf))
    return g
