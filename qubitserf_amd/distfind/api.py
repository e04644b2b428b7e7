"""Distance entry points with the reference's signatures and conventions
(python/qubitserf/distfind/api.py:66-247): every call returns a plain ``int``, ``-1`` when
the distance is undefined (no logicals); ``which`` is one of min/z/x; ``backend`` must be
one of auto/cpu/gpu (anything else is a ``ValueError``). This engine serves "gpu" and
"auto" on a gfx950 device; asking for it without one — or asking for "cpu", which stays
with the reference library — raises ``RuntimeError``. Only ``method="bz"`` runs here.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native

_WHICH = {"min": "M", "m": "M", "z": "Z", "x": "X"}
_BACKENDS = {"auto", "cpu", "gpu"}


def _normalize_backend(backend) -> str:
    b = "auto" if backend is None else str(backend).lower()
    if b not in _BACKENDS:
        raise ValueError("backend must be one of auto/cpu/gpu")
    if b == "cpu":
        raise RuntimeError("backend 'cpu' is served by the reference library, not the gfx950 engine")
    if not _native.load().distfind_backend_available(b.encode()):
        raise RuntimeError("no GPU backend is available on this machine")
    return b


def _as_u8(mat) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(mat, dtype=np.uint8) & 1)
    if a.ndim != 2:
        raise ValueError("expected a 2D 0/1 matrix")
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_native.u8p)


def _as_css(Hx, Hz):
    if Hz is None:
        obj = Hx
        if isinstance(obj, (tuple, list)) and len(obj) == 2:
            Hx, Hz = obj
        elif hasattr(obj, "Hx") and hasattr(obj, "Hz"):
            Hx, Hz = obj.Hx, obj.Hz
        else:
            raise TypeError("css_distance expects (Hx, Hz) or an object with .Hx and .Hz")
    return _as_u8(Hx), _as_u8(Hz)


def _check(rc: int, name: str):
    if rc == 0:
        return
    msg = _native.last_error()
    if rc in (3, 4):
        raise ValueError(f"{name}: {msg}")
    raise RuntimeError(f"{name} failed (rc={rc}): {msg}")


def _which(which) -> str:
    w = _WHICH.get(str(which).lower())
    if w is None:
        raise ValueError("which must be one of min/z/x")
    return w


def css_distance_raw(Hx, Hz, method="bz", which="M", backend="gpu", threads=0, verbose=False):
    L = _native.load()
    ax, az = _as_u8(Hx), _as_u8(Hz)
    out = _native.DistfindResult()
    rc = L.distfind_css_distance(_ptr(ax), ax.shape[0], ax.shape[1], _ptr(az), az.shape[0], az.shape[1],
                                 method.encode(), which.encode()[:1] or b"M", backend.encode(),
                                 int(threads), 1 if verbose else 0, ctypes.byref(out))
    _check(rc, "distfind_css_distance")
    return out


def classical_distance_raw(H, method="bz", backend="gpu", threads=0, verbose=False):
    L = _native.load()
    a = _as_u8(H)
    out = _native.DistfindResult()
    rc = L.distfind_classical_distance(_ptr(a), a.shape[0], a.shape[1], method.encode(), backend.encode(),
                                       int(threads), 1 if verbose else 0, ctypes.byref(out))
    _check(rc, "distfind_classical_distance")
    return out


def stabilizer_distance_raw(S, method="bz", which="M", backend="gpu", threads=0, verbose=False):
    L = _native.load()
    a = _as_u8(S)
    if a.shape[1] % 2 != 0:
        raise ValueError("symplectic stabilizer matrix must have 2n (even) columns")
    out = _native.DistfindResult()
    rc = L.distfind_stabilizer_distance(_ptr(a), a.shape[0], a.shape[1], method.encode(),
                                        which.encode()[:1] or b"M", backend.encode(), int(threads),
                                        1 if verbose else 0, ctypes.byref(out))
    _check(rc, "distfind_stabilizer_distance")
    return out


def css_distance(Hx, Hz=None, *, method="bz", which="min", backend="auto", threads=0, verbose=False) -> int:
    """Distance of a CSS code from its X- and Z-check matrices (api.py:75-96)."""
    Hx, Hz = _as_css(Hx, Hz)
    w = _which(which)
    b = _normalize_backend(backend)
    return int(css_distance_raw(Hx, Hz, method, w, b, threads, verbose).distance)


def classical_distance(H, *, method="bz", backend="auto", threads=0, verbose=False) -> int:
    """Minimum distance of the classical code with parity-check matrix H (api.py:234-239)."""
    b = _normalize_backend(backend)
    return int(classical_distance_raw(H, method, b, threads, verbose).distance)


def stabilizer_distance(S, *, method="bz", which="min", backend="auto", threads=0, verbose=False) -> int:
    """Distance of a stabilizer code given as an (m, 2n) [z|x] matrix (api.py:152-182). CSS
    input is routed to the CSS solver; a non-CSS code runs BZ on its weight-doubling
    isometric image and the result is halved."""
    a = np.ascontiguousarray(np.asarray(S, dtype=np.uint8) & 1)
    if a.ndim != 2 or a.shape[1] % 2 != 0:
        raise ValueError("expected a 2D symplectic matrix with 2n (even) columns in [z | x] order")
    w = _which(which)
    b = _normalize_backend(backend)
    return int(stabilizer_distance_raw(a, method, w, b, threads, verbose).distance)


def available_backends():
    """Usable backends of this engine: ["gpu"] on a gfx950 machine, else []."""
    return ["gpu"] if _native.load().distfind_backend_available(b"gpu") else []


def version() -> str:
    return _native.load().distfind_version().decode()


def last_trace():
    """Per-level records of the last solve on this thread, as dicts with the fields of
    DistfindLevel (include/distfind_capi.h): side, d, active, ngamma, K, inner, outer,
    kernel, fd, xd, rpl, mfma, codewords, medium_hits, slow_hits, ms, kernel_ms, ops_per_codeword,
    side_ms, side_codewords, side_ops_per_codeword."""
    L = _native.load()
    n = L.distfind_last_trace(None, 0)
    buf = (_native.DistfindLevel * max(n, 1))()
    L.distfind_last_trace(buf, n)
    out = []
    for i in range(n):
        r = buf[i]
        rec = {f: getattr(r, f) for f, _ in _native.DistfindLevel._fields_}
        rec["side"] = chr(rec["side"])
        out.append(rec)
    return out


def _batch_items(problems):
    items = (_native.DistfindBatchItem * max(1, len(problems)))()
    keep = []
    for i, p in enumerate(problems):
        it = items[i]
        if isinstance(p, dict):
            it.work_cap = int(p.get("work_cap", 0))
            p = p["problem"]
        if p[0] == "classical":
            H = _as_u8(p[1])
            it.kind, it.A, it.ar, it.ac = 0, _ptr(H), H.shape[0], H.shape[1]
            keep.append(H)
            if len(p) > 2 and p[2] is not None:
                o = np.ascontiguousarray(p[2], np.int32)
                it.order = o.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
                keep.append(o)
            if len(p) > 3 and p[3] is not None:
                sc, cands, seed = p[3]
                sv = np.ascontiguousarray(sc, np.float32)
                if sv.shape != (H.shape[1],):
                    raise ValueError("scores must hold one value per column")
                it.scores = sv.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
                it.candidates, it.proposal_seed = int(cands), int(seed)
                keep.append(sv)
        elif p[0] == "css":
            Hx, Hz = _as_css(p[1], p[2])
            it.kind, it.A, it.ar, it.ac = 1, _ptr(Hx), Hx.shape[0], Hx.shape[1]
            it.B, it.br, it.bc = _ptr(Hz), Hz.shape[0], Hz.shape[1]
            it.which = _which(p[3] if len(p) > 3 else "min").encode()
            keep += [Hx, Hz]
        else:
            raise ValueError("problem kind must be 'classical' or 'css'")
    return items, keep


def batch_distance(problems, *, work_cap=0, is_trials=0, verbose=False):
    """Distances of many codes in one batched solve (``distfind_batch_distance``).

    Each problem is ``("classical", H)``, ``("classical", H, order)``,
    ``("classical", H, None, (scores, candidates, seed))`` or ``("css", Hx, Hz, which)``;
    ``order`` is a column order for the information-set peel, ``scores`` a proposal's
    column scores from which ``candidates`` orders are sampled. A problem may also be a
    dict ``{"problem": <tuple>, "work_cap": cap}`` with its own work cap.
    Returns a list of ``DistfindResult`` (distance, lower_bound, proven, levels), identical
    to the single-code entry points' results; with ``work_cap`` > 0 a problem whose next
    level would exceed that many codewords stops with ``proven == 0``; ``is_trials`` sets the
    random information-set trials per problem (0: 256)."""
    L = _native.load()
    _normalize_backend("gpu")
    items, keep = _batch_items(problems)
    out = (_native.DistfindResult * max(1, len(problems)))()
    rc = L.distfind_batch_distance(items, len(problems), int(work_cap), int(is_trials), 1 if verbose else 0, out)
    _check(rc, "batch_distance")
    del keep
    return [out[i] for i in range(len(problems))]


class Session:
    """Continuous batching over ``distfind_session_*``: ``add`` problems (same forms as
    ``batch_distance``) at any time, ``step`` runs one round (one level of every unfinished
    problem), ``poll`` returns ``(id, DistfindResult)`` of the problems finished since."""

    def __init__(self, work_cap=0, is_trials=0):
        _normalize_backend("gpu")
        self.L = _native.load()
        self.h = self.L.distfind_session_new(int(work_cap), int(is_trials))
        if not self.h:
            raise RuntimeError(f"session: {_native.last_error()}")

    def add(self, problems):
        items, keep = _batch_items(problems)
        ids = (ctypes.c_int64 * max(1, len(problems)))()
        _check(self.L.distfind_session_add(self.h, items, len(problems), ids), "session.add")
        del keep
        return [ids[i] for i in range(len(problems))]

    def step(self) -> int:
        a = ctypes.c_int(0)
        _check(self.L.distfind_session_step(self.h, ctypes.byref(a)), "session.step")
        return a.value

    def poll(self):
        out = []
        while True:
            ids = (ctypes.c_int64 * 64)()
            res = (_native.DistfindResult * 64)()
            n = self.L.distfind_session_poll(self.h, ids, res, 64)
            if n < 0:
                raise RuntimeError(f"session.poll: {_native.last_error()}")
            out += [(ids[i], res[i]) for i in range(n)]
            if n < 64:
                return out

    @property
    def active(self) -> int:
        return self.L.distfind_session_active(self.h)

    @property
    def device_bytes(self) -> int:
        """Device bytes the unfinished problems hold now (measured solve arenas)."""
        return int(self.L.distfind_session_device_bytes(self.h))

    def stats(self) -> dict:
        s = _native.DistfindBatchStats()
        self.L.distfind_session_stats(self.h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self):
        if self.h:
            self.L.distfind_session_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class sound_bounds:
    """Context manager (or call set_sound_bounds) for the sound level schedule: every
    information set that enters the lower bound is enumerated from level 1. Off (the
    default) is the reference's schedule, which skips a rank-deficient set's first K - r
    levels yet counts it in the bound (cpp/distfind/bz.cpp:142-144, 159-161)."""

    def __init__(self, on=True):
        self.on = bool(on)

    def __enter__(self):
        self.prev = set_sound_bounds(self.on)
        return self

    def __exit__(self, *exc):
        set_sound_bounds(self.prev)


def set_sound_bounds(on: bool) -> bool:
    """Switch the sound level schedule for later solves; returns the previous setting."""
    return bool(_native.load().distfind_set_sound_bounds(1 if on else 0))


def device_memory():
    """(free, total) device bytes (hipMemGetInfo)."""
    L = _native.load()
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(L.qsf_hip_mem_info(ctypes.byref(f), ctypes.byref(t)), "device_memory")
    return int(f.value), int(t.value)


def last_batch_stats() -> dict:
    s = _native.DistfindBatchStats()
    _native.load().distfind_last_batch_stats(ctypes.byref(s))
    return {f: getattr(s, f) for f, _ in s._fields_}
