"""The scorer's CPU restatement (oracle/mp_scorer.py) pinned by finite differences in
float64, and its BF16 rounding by the bit patterns bfloat16 is defined by."""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
_spec = importlib.util.spec_from_file_location("mp_scorer", ROOT / "oracle" / "mp_scorer.py")
M = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(M)


def _net(seed=0, B=2, n=9, m=5, F=16, D=64):
    rng = np.random.default_rng(seed)
    shapes = {"Win": (F, D), "bin": (D,), "wo": (D, 1), "bo": (1,)}
    for l in range(M.LAYERS):
        shapes.update({f"Wc{l}": (D, D), f"bc{l}": (D,), f"Ws{l}": (D, D), f"Wn{l}": (D, D), f"b{l}": (D,)})
    P = {k: rng.standard_normal(s) * (0.3 if k.startswith(("W", "w")) else 0.1) for k, s in shapes.items()}
    H = (rng.random((B, m, n)) < 0.4).astype(np.float64)
    Ac = H / np.maximum(H.sum(2, keepdims=True), 1)
    Av = np.swapaxes(H / np.maximum(H.sum(1, keepdims=True), 1), 1, 2)
    X = rng.standard_normal((B, n, F))
    T = (rng.random((B, n)) < 0.5).astype(np.float64)
    mask = np.ones((B, n))
    mask[1, -2:] = 0
    return P, X, Ac, Av, T, mask


def test_backward_matches_finite_differences():
    P, X, Ac, Av, T, mask = _net()

    def L(P):
        s, _ = M.forward(P, X, Ac, Av, bf=False)
        return M.loss_grad(s, T, mask)[0]

    s, c = M.forward(P, X, Ac, Av, bf=False)
    _, ds = M.loss_grad(s, T, mask)
    g = M.backward(c, ds, bf=False)
    rng = np.random.default_rng(1)
    for k in M.param_names():
        flat = P[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            eps = 1e-6
            old = flat[idx]
            flat[idx] = old + eps
            lp = L(P)
            flat[idx] = old - eps
            lm = L(P)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = g[k].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (k, idx, fd, an)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.0, 3.0e38, 1e-40, 1 + 2 ** -8, 1 + 3 * 2 ** -8], np.float32)
    y = M.bf16(x)
    assert y[0] == 1.0 and y[3] == -2.0
    assert y[1] == 1.0                    # 1 + 2^-8 is a tie: rounds to even (1.0)
    assert y[2] == np.float32(1.0078125)  # above the tie: rounds up
    assert y[7] == np.float32(1.015625)   # 1 + 3 2^-8: tie between 1+2^-7 and 1+2^-6 -> even (1+2^-6)
    assert (y.view(np.uint32) & 0xFFFF == 0).all()


def test_bf16_mode_close_to_exact():
    P, X, Ac, Av, T, mask = _net(3)
    s64, _ = M.forward(P, X, Ac, Av, bf=False)
    s16, _ = M.forward(P, X, Ac, Av, bf=True)
    assert np.abs(s64 - s16).max() <= 0.05 * (1 + np.abs(s64).max())


def test_pack_batch_bits_equals_pack_batch():
    """The batched bit-pattern packing used for inference equals pack_batch followed by
    the float32 -> bfloat16 rounding, bit for bit, on ragged codes."""
    from qubitserf_amd import codes
    from qubitserf_amd.proposal.devarray import bf16_bits
    from qubitserf_amd.proposal.model import pack_batch, pack_batch_bits

    Hs = codes.config3_batch(5, 12) + [np.eye(3, 7, dtype=np.uint8), np.zeros((2, 5), np.uint8)]
    seeds = list(range(100, 100 + len(Hs)))
    X, Ac, Av, mask = pack_batch(Hs, seeds)
    X2, A2, V2, m2 = pack_batch_bits(Hs, seeds)
    assert np.array_equal(X, X2) and np.array_equal(mask, m2)
    assert A2.dtype == np.uint16 and np.array_equal(bf16_bits(Ac), A2) and np.array_equal(bf16_bits(Av), V2)
    # without seeds every code draws from seed 0 (one draw, cut per code, for the batch)
    X, Ac, Av, mask = pack_batch(Hs)
    X2, A2, V2, m2 = pack_batch_bits(Hs)
    assert np.array_equal(X, X2) and np.array_equal(mask, m2)
    assert np.array_equal(bf16_bits(Ac), A2) and np.array_equal(bf16_bits(Av), V2)
