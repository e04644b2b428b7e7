"""ctypes access to the checkers: the oracle restatement (oracle/_build/liboracle.so) and,
when it has been built, the reference's own CPU library (oracle/_ref/libqubitserf_ref.so).
Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_LIB = ROOT / "oracle" / "_build" / "liboracle.so"
REF_LIB = ROOT / "oracle" / "_ref" / "libqubitserf_ref.so"
# the reference's sources with the gfx950 plugin in the accelerator slot (oracle/ref.mk)
REF_HIP_LIB = ROOT / "oracle" / "_ref" / "libqubitserf_ref_hip.so"
WEIGHT_NONE = 1 << 30

u8p = ctypes.POINTER(ctypes.c_uint8)
u64p = ctypes.POINTER(ctypes.c_uint64)


class OracleResult(ctypes.Structure):
    _fields_ = [("distance", ctypes.c_int), ("lower_bound", ctypes.c_int), ("proven", ctypes.c_int),
                ("levels", ctypes.c_int), ("seconds", ctypes.c_double)]


class RefResult(ctypes.Structure):
    _fields_ = [("distance", ctypes.c_int), ("lower_bound", ctypes.c_int), ("proven", ctypes.c_int),
                ("levels", ctypes.c_int), ("seconds", ctypes.c_double), ("backend", ctypes.c_char * 16)]


_o = None
_r = None
_rh = None


def lib():
    global _o
    if _o is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(str(ORACLE_LIB))
        i, c = ctypes.c_int, ctypes.c_char
        L.oracle_enumerate.argtypes = [i, i, i, i, i, u64p, i, u64p, i, i]
        L.oracle_enumerate_brute.argtypes = [i, i, i, i, i, u64p, i, u64p, i]
        L.oracle_collect.argtypes = [i, i, i, i, i, u64p, i, u64p, ctypes.c_int64, i]
        L.oracle_collect.restype = ctypes.c_int64
        rp = ctypes.POINTER(OracleResult)
        L.oracle_classical_distance.argtypes = [u8p, i, i, i, i, rp]
        L.oracle_css_distance.argtypes = [u8p, i, i, u8p, i, i, c, i, i, rp]
        L.oracle_stabilizer_distance.argtypes = [u8p, i, i, c, i, i, rp]
        L.oracle_bruteforce_distance.argtypes = [i, u8p, i, i, u8p, i, i]
        L.oracle_mitm_distance.argtypes = [i, u8p, i, i, u8p, i, i, i]
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_problem_info.argtypes = [i, u8p, i, i, u8p, i, i, ip, ip, ip, ip, ip, i]
        L.oracle_problem_gen.argtypes = [i, u8p, i, i, u8p, i, i, u64p, ctypes.c_int64, u64p, ctypes.c_int64, ip, ip]
        L.oracle_greedy_sets.argtypes = [u64p, i, i, u64p, i, ip]
        L.oracle_part_sets.argtypes = [u64p, i, i, ip, ip, i, u64p, ip]
        L.oracle_part_sets.restype = None
        L.oracle_is_bound_trials.argtypes = [u64p, i, i, u64p, i, i, i, i, ctypes.c_uint64, i, u64p, ip]
        _o = L
    return _o


def ref_available() -> bool:
    return REF_LIB.exists()


def _bind_ref(L):
    i, c, s = ctypes.c_int, ctypes.c_char, ctypes.c_char_p
    rp = ctypes.POINTER(RefResult)
    L.distfind_css_distance.argtypes = [u8p, i, i, u8p, i, i, s, c, s, i, i, rp]
    L.distfind_classical_distance.argtypes = [u8p, i, i, s, s, i, i, rp]
    L.distfind_stabilizer_distance.argtypes = [u8p, i, i, s, c, s, i, i, rp]
    L.distfind_backend_available.argtypes = [s]
    L.codeaut_bz_collect.restype = ctypes.c_void_p
    L.codeaut_bz_collect.argtypes = [u8p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_int32]
    for f in ("codeaut_bz_count", "codeaut_bz_combos"):
        getattr(L, f).restype = ctypes.c_int64
        getattr(L, f).argtypes = [ctypes.c_void_p]
    for f in ("codeaut_bz_ok",):
        getattr(L, f).restype = ctypes.c_int32
        getattr(L, f).argtypes = [ctypes.c_void_p]
    L.codeaut_bz_backend.restype = s
    L.codeaut_bz_backend.argtypes = [ctypes.c_void_p]
    L.codeaut_bz_copy_rows.argtypes = [ctypes.c_void_p, u8p]
    L.codeaut_bz_free.argtypes = [ctypes.c_void_p]
    L.codeaut_backend_available.restype = ctypes.c_int32
    L.codeaut_backend_available.argtypes = [s]
    return L


def ref_hip_available() -> bool:
    return REF_HIP_LIB.exists()


def ref_hip():
    """The reference's own drivers and C ABI with the gfx950 plugin as its GPU backend."""
    global _rh
    if _rh is None:
        L = _bind_ref(ctypes.CDLL(str(REF_HIP_LIB)))
        i = ctypes.c_int
        L.ref_hip_plugin_enumerate.argtypes = [i, i, i, i, i, u64p, i, u64p, i]
        L.ref_hip_plugin_collect.restype = ctypes.c_int64
        L.ref_hip_plugin_collect.argtypes = [i, i, i, i, u64p, i, u64p, ctypes.c_int64, ctypes.POINTER(i)]
        L.ref_hip_plugin_name.restype = ctypes.c_char_p
        _rh = L
    return _rh


def ref():
    global _r
    if _r is None:
        _r = _bind_ref(ctypes.CDLL(str(REF_LIB)))
    return _r


def _u8(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8) & 1)


def _p(a):
    return a.ctypes.data_as(u8p)


# ---------------------------------------------------------------- oracle restatement

class sound:
    """The oracle's sound level schedule (oracle_set_sound) inside a with-block."""

    def __init__(self, on=True):
        self.on = 1 if on else 0

    def __enter__(self):
        L = lib()
        L.oracle_set_sound.argtypes = [ctypes.c_int]
        self.prev = L.oracle_set_sound(self.on)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_sound(self.prev)


def classical_distance(H, threads=0, is_threads=0) -> OracleResult:
    a = _u8(H)
    r = OracleResult()
    assert lib().oracle_classical_distance(_p(a), a.shape[0], a.shape[1], threads, is_threads, ctypes.byref(r)) == 0
    return r


def css_distance(Hx, Hz, which="M", threads=0, is_threads=0) -> OracleResult:
    ax, az = _u8(Hx), _u8(Hz)
    r = OracleResult()
    assert lib().oracle_css_distance(_p(ax), ax.shape[0], ax.shape[1], _p(az), az.shape[0], az.shape[1],
                                     which.encode()[:1], threads, is_threads, ctypes.byref(r)) == 0
    return r


def stabilizer_distance(S, which="M", threads=0, is_threads=0) -> OracleResult:
    a = _u8(S)
    r = OracleResult()
    assert lib().oracle_stabilizer_distance(_p(a), a.shape[0], a.shape[1], which.encode()[:1], threads,
                                            is_threads, ctypes.byref(r)) == 0
    return r


def _kind_args(kind, A, B):
    a = _u8(A)
    b = _u8(B) if B is not None else np.zeros((1, 1), np.uint8)
    return a, b


def bruteforce_distance(kind, A, B=None) -> int:
    """kind 0: classical parity check A; 1/2: CSS Z/X side of (Hx=A, Hz=B). 2^K messages."""
    a, b = _kind_args(kind, A, B)
    return lib().oracle_bruteforce_distance(kind, _p(a), a.shape[0], a.shape[1], _p(b), b.shape[0], b.shape[1])


def mitm_distance(kind, A, B=None, wmax=64) -> int:
    a, b = _kind_args(kind, A, B)
    return lib().oracle_mitm_distance(kind, _p(a), a.shape[0], a.shape[1], _p(b), b.shape[0], b.shape[1], wmax)


def problem_info(kind, A, B=None):
    a, b = _kind_args(kind, A, B)
    K, kc, ev, su = (ctypes.c_int() for _ in range(4))
    ranks = (ctypes.c_int * 512)()
    ng = lib().oracle_problem_info(kind, _p(a), a.shape[0], a.shape[1], _p(b), b.shape[0], b.shape[1],
                                   ctypes.byref(K), ctypes.byref(kc), ctypes.byref(ev), ctypes.byref(su), ranks, 512)
    return dict(K=K.value, kcheck=kc.value, even=bool(ev.value), seed_upper=su.value, ranks=list(ranks[:ng]))


def _plan_arrays(gamma, check):
    g = np.ascontiguousarray(gamma, dtype=np.uint64)
    ngt, K, stride = g.shape
    c = np.ascontiguousarray(check if check is not None else np.zeros((0, stride), np.uint64), dtype=np.uint64)
    cp = c.ctypes.data_as(u64p) if c.shape[0] else None
    return g, c, cp, ngt, K, stride


def enumerate_plan(gamma, n, d, check=None, num_gamma=None, current_best=WEIGHT_NONE, threads=0) -> int:
    g, c, cp, ngt, K, stride = _plan_arrays(gamma, check)
    ng = ngt if num_gamma is None else num_gamma
    return lib().oracle_enumerate(n, stride, K, d, ng, g.ctypes.data_as(u64p), c.shape[0], cp, current_best, threads)


def enumerate_brute(gamma, n, d, check=None, num_gamma=None, current_best=WEIGHT_NONE) -> int:
    g, c, cp, ngt, K, stride = _plan_arrays(gamma, check)
    ng = ngt if num_gamma is None else num_gamma
    return lib().oracle_enumerate_brute(n, stride, K, d, ng, g.ctypes.data_as(u64p), c.shape[0], cp, current_best)


def collect_plan(gamma, n, d, keep, num_gamma=None, threads=0) -> np.ndarray:
    g, _, _, ngt, K, stride = _plan_arrays(gamma, None)
    ng = ngt if num_gamma is None else num_gamma
    cap = 1 << 16
    while True:
        out = np.zeros((cap, stride), np.uint64)
        rows = lib().oracle_collect(n, stride, K, d, ng, g.ctypes.data_as(u64p), keep, out.ctypes.data_as(u64p),
                                    cap, threads)
        if rows <= cap:
            return out[:rows]
        cap = rows


def problem_gen(kind, A, B=None):
    """(gen[K, stride], check[kcheck, stride] or None when every nonzero codeword counts)."""
    a, b = _kind_args(kind, A, B)
    n = a.shape[1]
    stride = (n + 63) // 64
    gen = np.zeros((n, stride), np.uint64)
    chk = np.zeros((n + 1, stride), np.uint64)
    K, kc = ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_problem_gen(kind, _p(a), a.shape[0], a.shape[1], _p(b), b.shape[0], b.shape[1],
                                  gen.ctypes.data_as(u64p), gen.size, chk.ctypes.data_as(u64p), chk.size,
                                  ctypes.byref(K), ctypes.byref(kc))
    assert rc == 0
    return gen[: K.value].copy(), (chk[: kc.value].copy() if kc.value else None)


def greedy_sets(gen, n):
    """(sets[ns, K, stride], ranks) of the greedy peel, in peel order."""
    g = np.ascontiguousarray(gen, np.uint64)
    K = g.shape[0]
    stride = (n + 63) // 64
    cap = n + 2
    out = np.zeros((cap, max(K, 1), stride), np.uint64)
    ranks = np.zeros(cap, np.int32)
    ns = lib().oracle_greedy_sets(g.ctypes.data_as(u64p), K, n, out.ctypes.data_as(u64p), cap,
                                  ranks.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out[:ns, :K], ranks[:ns].tolist()


def part_sets(gen, n, parts):
    g = np.ascontiguousarray(gen, np.uint64)
    K = g.shape[0]
    stride = (n + 63) // 64
    cols = np.array([c for p in parts for c in p] or [0], np.int32)
    off = np.cumsum([0] + [len(p) for p in parts]).astype(np.int32)
    out = np.zeros((len(parts), max(K, 1), stride), np.uint64)
    ranks = np.zeros(len(parts), np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib().oracle_part_sets(g.ctypes.data_as(u64p), K, n, cols.ctypes.data_as(ip), off.ctypes.data_as(ip), len(parts),
                           out.ctypes.data_as(u64p), ranks.ctypes.data_as(ip))
    return out[:, :K], ranks.tolist()


def is_bound_trials(gen, n, check=None, trials=256, seed=0x9E3779B9, init=WEIGHT_NONE, first=0) -> int:
    """The engine's random information-set bound restated on the CPU (oracle_is_bound_trials)."""
    g = np.ascontiguousarray(gen, np.uint64)
    K = g.shape[0]
    stride = (n + 63) // 64
    c = np.ascontiguousarray(check if check is not None else np.zeros((1, stride), np.uint64), np.uint64)
    kcheck = 0 if check is None else c.shape[0]
    return lib().oracle_is_bound_trials(g.ctypes.data_as(u64p) if K else None, K, n, c.ctypes.data_as(u64p), kcheck,
                                        1 if check is None else 0, first, trials, seed, init, None, None)


def is_trial(gen, n, t, seed=0x9E3779B9):
    """(systematic form K x stride, column order) of trial t of the engine's stream."""
    g = np.ascontiguousarray(gen, np.uint64)
    K = g.shape[0]
    stride = (n + 63) // 64
    rows = np.zeros((K, stride), np.uint64)
    perm = np.zeros(n, np.int32)
    c = np.zeros((1, stride), np.uint64)
    lib().oracle_is_bound_trials(g.ctypes.data_as(u64p), K, n, c.ctypes.data_as(u64p), 0, 1, t, 1, seed, WEIGHT_NONE,
                                 rows.ctypes.data_as(u64p), perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return rows, perm


# ---------------------------------------------------------------- reference CPU library

# L: the library to call (default: the reference's CPU build); backend: its backend name
# ("cpu", or "gpu" to run the plugin in the ref_hip() build).

def ref_css_distance(Hx, Hz, which="M", threads=0, L=None, backend="cpu") -> RefResult:
    ax, az = _u8(Hx), _u8(Hz)
    r = RefResult()
    assert (L or ref()).distfind_css_distance(_p(ax), ax.shape[0], ax.shape[1], _p(az), az.shape[0], az.shape[1],
                                              b"bz", which.encode()[:1], backend.encode(), threads, 0,
                                              ctypes.byref(r)) == 0
    return r


def ref_classical_distance(H, threads=0, L=None, backend="cpu") -> RefResult:
    a = _u8(H)
    r = RefResult()
    assert (L or ref()).distfind_classical_distance(_p(a), a.shape[0], a.shape[1], b"bz", backend.encode(), threads,
                                                    0, ctypes.byref(r)) == 0
    return r


def ref_stabilizer_distance(S, which="M", threads=0, method="bz", L=None, backend="cpu") -> RefResult:
    a = _u8(S)
    r = RefResult()
    assert (L or ref()).distfind_stabilizer_distance(_p(a), a.shape[0], a.shape[1], method.encode(),
                                                     which.encode()[:1], backend.encode(), threads, 0,
                                                     ctypes.byref(r)) == 0
    return r


def ref_collect(G, p, keep, threads=0, L=None, backend=0, with_meta=False):
    """(rows, combos) of the reference's codeaut_bz_collect; with_meta adds (ok, backend name)."""
    a = _u8(G)
    L = L or ref()
    h = L.codeaut_bz_collect(_p(a), a.shape[0], a.shape[1], p, keep, backend, threads)
    try:
        if with_meta:
            meta = (L.codeaut_bz_ok(h), L.codeaut_bz_backend(h).decode())
        cnt = L.codeaut_bz_count(h)
        out = np.zeros((cnt, a.shape[1]), np.uint8)
        if cnt:
            L.codeaut_bz_copy_rows(h, _p(out))
        return (out, L.codeaut_bz_combos(h)) + (meta if with_meta else ())
    finally:
        L.codeaut_bz_free(h)


def side_plan(kind, A, B=None):
    """(gamma[ng,K,stride], ranks, check[kcheck,stride], even) of a subproblem's information sets."""
    L = lib()
    a, b = _kind_args(kind, A, B)
    n = a.shape[1]
    stride = (n + 63) // 64
    cap = 64 * n * stride * 8 + 64
    gamma = np.zeros(cap, np.uint64)
    ranks = (ctypes.c_int * 256)()
    check = np.zeros(n * stride + 1, np.uint64)
    K, kc, ev = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.oracle_side_plan.restype = ctypes.c_int
    L.oracle_side_plan.argtypes = [ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int, ctypes.c_int,
                                   u64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int), ctypes.c_int, u64p,
                                   ctypes.c_int64] + [ctypes.POINTER(ctypes.c_int)] * 3
    ng = L.oracle_side_plan(kind, _p(a), a.shape[0], a.shape[1], _p(b), b.shape[0], b.shape[1],
                            gamma.ctypes.data_as(u64p), cap, ranks, 256, check.ctypes.data_as(u64p), check.size,
                            ctypes.byref(K), ctypes.byref(kc), ctypes.byref(ev))
    assert ng >= 0
    k = K.value
    return (gamma[: ng * k * stride].reshape(ng, k, stride).copy(), list(ranks[:ng]),
            check[: kc.value * stride].reshape(kc.value, stride).copy(), bool(ev.value))


def ref_enumerate_plan(gamma, n, d, check=None, num_gamma=None, current_best=WEIGHT_NONE, threads=0) -> int:
    """The reference's own CPU enumerator on one level (oracle/ref_shim.cpp)."""
    L = ref()
    f = L.ref_enumerate_plan
    f.restype = ctypes.c_int
    i = ctypes.c_int
    f.argtypes = [i, i, i, i, i, u64p, i, u64p, i, i]
    g, c, cp, ngt, K, stride = _plan_arrays(gamma, check)
    ng = ngt if num_gamma is None else num_gamma
    return f(n, stride, K, d, ng, g.ctypes.data_as(u64p), c.shape[0], cp, current_best, threads)
