"""One weight level at a time through the plugin-level C ABI (include/qsf_hip.h).

``enumerate_level`` / ``collect_level`` take numpy arrays shaped like the reference's
EnumPlan (cpp/qsf/backend.hpp:21-42): ``gamma`` is (num_gamma_total, K, stride) uint64,
``check`` (kcheck, stride) uint64.
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _native

WEIGHT_NONE = 1 << 30
_keys = itertools.count(0x7e57000000000000)


def new_key() -> int:
    return next(_keys)


def _plan(gamma, check, n, d, num_gamma, current_best, key):
    g = np.ascontiguousarray(gamma, dtype=np.uint64)
    if g.ndim != 3:
        raise ValueError("gamma must be (num_gamma_total, K, stride)")
    ngt, K, stride = g.shape
    c = np.ascontiguousarray(check if check is not None else np.zeros((0, stride), np.uint64), dtype=np.uint64)
    p = _native.EnumPlan()
    p.n, p.stride, p.K, p.d = int(n), int(stride), int(K), int(d)
    p.num_gamma = int(ngt if num_gamma is None else num_gamma)
    p.num_gamma_total = int(ngt)
    p.gamma = g.ctypes.data_as(_native.u64p)
    p.kcheck = int(c.shape[0])
    p.check = c.ctypes.data_as(_native.u64p) if c.shape[0] else None
    p.binom = None
    p.current_best = int(current_best)
    p.buffers_key = int(key if key is not None else new_key())
    return p, (g, c)


def _err(rc, name):
    raise RuntimeError(f"{name} failed (rc={rc}): {(_native.load().qsf_hip_last_error() or b'').decode()}")


def available() -> bool:
    return bool(_native.load().qsf_hip_available())


def enumerate_level(gamma, n, d, *, check=None, num_gamma=None, current_best=WEIGHT_NONE, key=None, floor=-1):
    """min(current_best, lightest logical XOR of d rows of one of the first num_gamma generators)."""
    L = _native.load()
    p, keep = _plan(gamma, check, n, d, num_gamma, current_best, key)
    out = ctypes.c_int(0)
    rc = L.qsf_hip_enumerate_floor(ctypes.byref(p), int(floor), ctypes.byref(out))
    if rc:
        _err(rc, "qsf_hip_enumerate")
    del keep
    return out.value


def plan_variant(gamma, n, d, *, check=None, num_gamma=None, current_best=WEIGHT_NONE, factor=0.0, at_p=0.0):
    """(fd, xd, rpl, mm) of the variant the automatic kernel choice picks for this level
    (host work only). factor/at_p: a structure factor as if measured at a level with fair
    per-codeword pass probability at_p (0: none measured)."""
    L = _native.load()
    p, keep = _plan(gamma, check, n, d, num_gamma, current_best, None)
    out = (ctypes.c_int * 4)()
    rc = L.qsf_hip_plan_variant(ctypes.byref(p), float(factor), float(at_p), out)
    if rc:
        _err(rc, "qsf_hip_plan_variant")
    del keep
    return tuple(out)


def is_bound(gen, n, *, check=None, trials=256, seed=0x9E3779B9, init=WEIGHT_NONE) -> int:
    """Random information-set upper bound on the device: min(init, lightest nonzero logical
    row over `trials` systematic forms of the packed generator `gen` (K x stride uint64)).
    check=None: every nonzero codeword is a logical (cpp/distfind/bz.cpp:24-70)."""
    L = _native.load()
    g = np.ascontiguousarray(gen, np.uint64)
    stride = (n + 63) // 64
    if g.ndim != 2 or (g.shape[0] and g.shape[1] != stride):
        raise ValueError(f"gen must be K x {stride} uint64")
    c = np.ascontiguousarray(check if check is not None else np.zeros((1, stride), np.uint64), np.uint64)
    kcheck = 0 if check is None else c.shape[0]
    out = ctypes.c_int(0)
    rc = L.qsf_hip_is_bound(g.ctypes.data_as(_native.u64p) if g.size else None, g.shape[0], int(n),
                            c.ctypes.data_as(_native.u64p), kcheck, 1 if check is None else 0, int(trials),
                            int(seed), int(init), ctypes.byref(out))
    if rc:
        _err(rc, "qsf_hip_is_bound")
    return out.value


ELIM_SETS, ELIM_BOUND = 1, 2


def eliminate(problems, *, what=ELIM_SETS | ELIM_BOUND, trials=256, seed=0x9E3779B9):
    """Batched device elimination (qsf_hip_eliminate). Each problem is a dict with `gen`
    (K x stride uint64), `n`, optional `check` (None: every nonzero codeword counts),
    `parts` (list of column lists), `order` (the peel's column order), `init`. Returns, per problem, a dict with `best` and,
    with ELIM_SETS, `sets` (ns x K x stride, peel/part order) and `ranks`."""
    L = _native.load()
    jobs = (_native.ElimJob * max(1, len(problems)))()
    keep = []
    for j, p in enumerate(problems):
        n = int(p["n"])
        stride = (n + 63) // 64
        g = np.ascontiguousarray(p["gen"], np.uint64).reshape(-1, stride)
        K = g.shape[0]
        q = jobs[j]
        q.gen = g.ctypes.data_as(_native.u64p) if K else None
        q.K, q.n = K, n
        chk = p.get("check")
        if chk is not None:
            c = np.ascontiguousarray(chk, np.uint64).reshape(-1, stride)
            q.check, q.kcheck, q.every_nonzero = c.ctypes.data_as(_native.u64p), c.shape[0], 0
            keep.append(c)
        else:
            q.check, q.kcheck, q.every_nonzero = None, 0, 1
        parts = p.get("parts")
        if parts:
            cols = np.array([c for part in parts for c in part] or [0], np.int32)
            off = np.cumsum([0] + [len(part) for part in parts]).astype(np.int32)
            q.part_cols = cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            q.part_off = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            q.nparts = len(parts)
            keep += [cols, off]
        order = p.get("order")
        if order is not None:
            o = np.ascontiguousarray(order, np.int32)
            q.order = o.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            keep.append(o)
        soc = np.full(n, -2, np.int32)
        q.set_of_col = soc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        keep.append(soc)
        cap = len(parts) if parts else int(p.get("set_cap", n + 2))
        sets = np.zeros((max(cap, 1), max(K, 1), stride), np.uint64)
        ranks = np.zeros(max(cap, 1), np.int32)
        q.sets = sets.ctypes.data_as(_native.u64p)
        q.set_cap = max(cap, 1)
        q.ranks = ranks.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        q.init = int(p.get("init", WEIGHT_NONE))
        keep += [g, sets, ranks]
        p["_out"] = (sets, ranks, soc)
    rc = L.qsf_hip_eliminate(jobs, len(problems), int(trials), int(seed), int(what))
    if rc:
        _err(rc, "qsf_hip_eliminate")
    out = []
    for j, p in enumerate(problems):
        sets, ranks, soc = p.pop("_out")
        q = jobs[j]
        r = {"best": q.best}
        if what & ELIM_SETS:
            ns = min(q.nsets, q.set_cap)
            r.update(nsets=q.nsets, sets=sets[:ns, : q.K].copy(), ranks=ranks[:ns].tolist(), set_of_col=soc.copy())
        out.append(r)
    del keep
    return out


def last_elim_stats() -> dict:
    s = _native.ElimStats()
    _native.load().qsf_hip_last_elim_stats(ctypes.byref(s))
    return {f: getattr(s, f) for f, _ in s._fields_}


def collect_level(gamma, n, d, keep_weight, *, num_gamma=None, key=None):
    """(rows, stride) uint64 array of every XOR of d rows with weight in [1, keep_weight] (dups allowed)."""
    L = _native.load()
    p, keep = _plan(gamma, None, n, d, num_gamma, WEIGHT_NONE, key)
    buf = _native.u64p()
    rows = ctypes.c_int64(0)
    rc = L.qsf_hip_collect(ctypes.byref(p), int(keep_weight), ctypes.byref(buf), ctypes.byref(rows))
    if rc:
        _err(rc, "qsf_hip_collect")
    stride = p.stride
    try:
        if rows.value == 0:
            return np.zeros((0, stride), np.uint64)
        arr = np.ctypeslib.as_array(buf, shape=(rows.value * stride,)).copy()
        return arr.reshape(rows.value, stride)
    finally:
        L.qsf_hip_free(ctypes.cast(buf, ctypes.c_void_p))
        del keep


def collect_upto(gamma, n, d, keep_weight, *, num_gamma=None, key=None):
    """(rows, stride) uint64 array of every distinct XOR of 1..d rows with weight in
    [1, keep_weight], sorted lexicographically by word (qsf_hip_collect_upto: one device
    pass over all levels, device sort and dedup)."""
    L = _native.load()
    p, keep = _plan(gamma, None, n, d, num_gamma, WEIGHT_NONE, key)
    buf = _native.u64p()
    rows = ctypes.c_int64(0)
    rc = L.qsf_hip_collect_upto(ctypes.byref(p), int(keep_weight), ctypes.byref(buf), ctypes.byref(rows))
    if rc:
        _err(rc, "qsf_hip_collect_upto")
    stride = p.stride
    try:
        if rows.value == 0:
            return np.zeros((0, stride), np.uint64)
        return np.ctypeslib.as_array(buf, shape=(rows.value * stride,)).copy().reshape(rows.value, stride)
    finally:
        L.qsf_hip_free(ctypes.cast(buf, ctypes.c_void_p))
        del keep


def last_collect_stats() -> dict:
    """Rows the last collect_upto appended before deduplication, and its device sort ms."""
    a, ms = ctypes.c_uint64(0), ctypes.c_double(0)
    _native.load().qsf_hip_last_collect_stats(ctypes.byref(a), ctypes.byref(ms))
    return {"appended": int(a.value), "sort_ms": float(ms.value)}


class kernel_policy:
    """Context manager forcing the kernel choice: "auto", "sweep" or "tiled"."""

    _codes = {"auto": 0, "sweep": 1, "tiled": 2}

    def __init__(self, name: str):
        self.code = self._codes[name]

    def __enter__(self):
        self.prev = _native.load().qsf_hip_set_kernel_policy(self.code)
        return self

    def __exit__(self, *exc):
        _native.load().qsf_hip_set_kernel_policy(self.prev)


def last_stats() -> dict:
    s = _native.HipStats()
    _native.load().qsf_hip_last_stats(ctypes.byref(s))
    return dict(codewords=int(s.codewords), kernel_ms=float(s.kernel_ms), kernel=int(s.kernel), tiles=int(s.tiles),
                fd=int(s.fd), xd=int(s.xd), rpl=int(s.rpl), slow_hits=int(s.slow_hits),
                medium_hits=int(s.medium_hits), ops_per_codeword=float(s.ops_per_codeword), mfma=int(s.mfma),
                side_ms=float(s.side_ms), side_codewords=int(s.side_codewords),
                side_ops_per_codeword=float(s.side_ops_per_codeword))


def pack_rows(M: np.ndarray) -> np.ndarray:
    """Pack a (rows, n) 0/1 matrix into (rows, ceil(n/64)) little-endian uint64 words."""
    M = np.asarray(M, dtype=np.uint8) & 1
    rows, n = M.shape
    stride = (n + 63) // 64
    padded = np.zeros((rows, stride * 64), dtype=np.uint8)
    padded[:, :n] = M
    bits = padded.reshape(rows, stride, 64).astype(np.uint64)
    return (bits << np.arange(64, dtype=np.uint64)).sum(axis=2, dtype=np.uint64)


def unpack_rows(P: np.ndarray, n: int) -> np.ndarray:
    P = np.asarray(P, dtype=np.uint64)
    bits = (P[:, :, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)
    return bits.reshape(P.shape[0], -1)[:, :n].astype(np.uint8)
