"""qsf_hip_enumerate_floor_noalloc (the entry point that runs on the library's static
scratch array and allocates nothing) against qsf_hip_enumerate_floor, both through the C
ABI, on the same plans: two calls in a row, the second with a smaller input, for each
kernel path (small-level sweep, VALU tiled kernel, int8 matrix-core kernel, row sweep).
Integer results: exact equality, and equality with the oracle's enumerator."""
import ctypes

import numpy as np
import pytest

from qubitserf_amd import _native, engine
from tests import oracle as O

pytestmark = pytest.mark.gpu
NONE = 1 << 30


def rand_plan(rng, K, stride, ng, kcheck):
    gamma = rng.integers(0, 2**64, size=(ng, K, stride), dtype=np.uint64)
    for g in range(ng):   # systematic on K random columns, like an information set
        cols = rng.choice(stride * 64, size=K, replace=False)
        for r, c in enumerate(cols):
            w, b = divmod(int(c), 64)
            gamma[g, :, w] &= ~np.uint64(1 << b)
            gamma[g, r, w] |= np.uint64(1 << b)
    check = rng.integers(0, 2**64, size=(kcheck, stride), dtype=np.uint64) if kcheck else None
    return gamma, check


def call(entry, gamma, check, n, d, best, floor, key):
    p, keep = engine._plan(gamma, check, n, d, None, best, key)
    out = ctypes.c_int(-7)
    rc = getattr(_native.load(), entry)(ctypes.byref(p), int(floor), ctypes.byref(out))
    assert rc == 0, f"{entry} rc={rc}: {_native.load().qsf_hip_last_error().decode()}"
    del keep
    return out.value, engine.last_stats()


# (label, K, stride, d, generators, detector rows, bound, floor); each pair: larger, then
# smaller. floor "min": the level's true minimum, which the floor contract requires as a
# lower bound.
CASES = [
    ("tiled", [(60, 3, 5, 2, 2, 40, -1), (34, 2, 4, 1, 1, 30, -1)]),
    ("small", [(24, 2, 3, 2, 2, NONE, -1), (12, 1, 2, 1, 0, NONE, -1)]),
    ("floor", [(56, 2, 5, 1, 2, NONE, "min"), (40, 2, 4, 1, 1, NONE, "min")]),
    ("rowsweep", [(300, 6, 3, 1, 1, NONE, -1), (270, 5, 3, 1, 0, NONE, -1)]),
]


@pytest.mark.parametrize("label,pair", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_enumerate_floor(hip, label, pair):
    rng = np.random.default_rng(sum(map(ord, label)))
    for i, (K, stride, d, ng, kcheck, best, floor) in enumerate(pair):
        print(f"noalloc case {label}[{i}] K={K} stride={stride} d={d} ng={ng}", flush=True)
        gamma, check = rand_plan(rng, K, stride, ng, kcheck)
        n = stride * 64
        oracle = O.enumerate_plan(gamma, n, d, check=check, current_best=best) if K <= 60 else None
        if floor == "min":   # the caller's guarantee: nothing in the solve is lighter than the floor
            floor = oracle
        key = engine.new_key()
        got, st_na = call("qsf_hip_enumerate_floor_noalloc", gamma, check, n, d, best, floor, key)
        want, st = call("qsf_hip_enumerate_floor", gamma, check, n, d, best, floor, engine.new_key())
        assert got == want, (label, i, got, want)
        assert st_na["kernel"] == st["kernel"] and st_na["codewords"] == st["codewords"], (st_na, st)
        if oracle is not None:
            assert got == oracle, (label, i)
        # the same solve again (its tables stay resident in the scratch array)
        again, _ = call("qsf_hip_enumerate_floor_noalloc", gamma, check, n, d, best, floor, key)
        assert again == want


def test_noalloc_matrix_kernel(hip, monkeypatch):
    """The int8 matrix-core kernel on scratch (the size rule is lowered so that the oracle
    can check the level), then a smaller level on the VALU kernel."""
    monkeypatch.setenv("QSA_MATRIX_MIN_WORK", "0")
    rng = np.random.default_rng(3)
    for K, stride, d, best in [(80, 3, 5, 7), (48, 2, 4, 9)]:
        print(f"noalloc matrix case K={K} d={d}", flush=True)
        gamma, check = rand_plan(rng, K, stride, 1, 2)
        n = stride * 64
        got, st_na = call("qsf_hip_enumerate_floor_noalloc", gamma, check, n, d, best, -1, engine.new_key())
        want, st = call("qsf_hip_enumerate_floor", gamma, check, n, d, best, -1, engine.new_key())
        assert got == want == O.enumerate_plan(gamma, n, d, check=check, current_best=best)
        assert st_na["mfma"] == st["mfma"], (st_na, st)


def test_noalloc_rejects_oversize(hip):
    """Tables larger than the scratch array: return code 4 with a reason, and the next call
    still works."""
    L = _native.load()
    cap = int(L.qsf_hip_noalloc_scratch_bytes())
    assert cap == 16 << 20
    rng = np.random.default_rng(11)
    K, stride, ng = 256, 16, 80   # gamma alone: 80 x 256 x 16 x 8 bytes = 2.6 MB; tables ~ 3x
    gamma = rng.integers(0, 2**64, size=(ng, K, stride), dtype=np.uint64)
    reps = 1 + cap // gamma.nbytes
    big = np.concatenate([gamma] * reps, axis=0)
    p, keep = engine._plan(big, None, stride * 64, 1, 1, NONE, engine.new_key())
    out = ctypes.c_int(0)
    rc = L.qsf_hip_enumerate_floor_noalloc(ctypes.byref(p), -1, ctypes.byref(out))
    assert rc == 4, rc
    assert b"no-allocation scratch" in L.qsf_hip_last_error()
    del keep
    gamma, check = rand_plan(rng, 30, 1, 1, 1)
    got, _ = call("qsf_hip_enumerate_floor_noalloc", gamma, check, 64, 3, NONE, -1, engine.new_key())
    want, _ = call("qsf_hip_enumerate_floor", gamma, check, 64, 3, NONE, -1, engine.new_key())
    assert got == want
