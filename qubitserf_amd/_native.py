"""Loads the engine library ``qubitserf_amd/_lib/libqubitserf_amd.so`` (built in-tree by
``__graft_entry__.build()`` / ``make -C qubitserf_amd/csrc``) and attaches ctypes
signatures for the C ABI in ``include/``.

There is no fallback: if the library is missing, importing the API raises.
The ``distfind_*`` signatures are the reference's (python/qubitserf/distfind/_native.py:39-107).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("QUBITSERF_AMD_LIB", _PKG / "_lib" / "libqubitserf_amd.so"))

u8p = ctypes.POINTER(ctypes.c_uint8)
u64p = ctypes.POINTER(ctypes.c_uint64)


class DistfindResult(ctypes.Structure):
    _fields_ = [
        ("distance", ctypes.c_int),
        ("lower_bound", ctypes.c_int),
        ("proven", ctypes.c_int),
        ("levels", ctypes.c_int),
        ("seconds", ctypes.c_double),
        ("backend", ctypes.c_char * 16),
    ]


class DistfindOpResult(ctypes.Structure):
    _fields_ = [
        ("z_weight", ctypes.c_int),
        ("x_weight", ctypes.c_int),
        ("proven", ctypes.c_int),
        ("seconds", ctypes.c_double),
        ("backend", ctypes.c_char * 16),
    ]


class EnumPlan(ctypes.Structure):
    """Layout of qsf_enum_plan (include/qsf_hip.h) = qsf::EnumPlan (cpp/qsf/backend.hpp:21-42)."""

    _fields_ = [
        ("n", ctypes.c_int),
        ("stride", ctypes.c_int),
        ("K", ctypes.c_int),
        ("d", ctypes.c_int),
        ("num_gamma", ctypes.c_int),
        ("num_gamma_total", ctypes.c_int),
        ("gamma", u64p),
        ("kcheck", ctypes.c_int),
        ("check", u64p),
        ("binom", u64p),
        ("binom_maxN", ctypes.c_int),
        ("binom_maxK", ctypes.c_int),
        ("current_best", ctypes.c_int),
        ("buffers_key", ctypes.c_uint64),
    ]


class HipStats(ctypes.Structure):
    _fields_ = [
        ("codewords", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("kernel", ctypes.c_int),
        ("tiles", ctypes.c_int),
        ("fd", ctypes.c_int),
        ("slow_hits", ctypes.c_uint64),
        ("xd", ctypes.c_int),
        ("medium_hits", ctypes.c_uint64),
        ("rpl", ctypes.c_int),
        ("ops_per_codeword", ctypes.c_double),
        ("mfma", ctypes.c_int),
        ("side_ms", ctypes.c_double),
        ("side_codewords", ctypes.c_uint64),
        ("side_ops_per_codeword", ctypes.c_double),
    ]


class ElimJob(ctypes.Structure):
    """Layout of qsf_elim_job (include/qsf_hip.h)."""
    _fields_ = [
        ("gen", ctypes.POINTER(ctypes.c_uint64)),
        ("K", ctypes.c_int),
        ("n", ctypes.c_int),
        ("check", ctypes.POINTER(ctypes.c_uint64)),
        ("kcheck", ctypes.c_int),
        ("every_nonzero", ctypes.c_int),
        ("part_cols", ctypes.POINTER(ctypes.c_int)),
        ("part_off", ctypes.POINTER(ctypes.c_int)),
        ("nparts", ctypes.c_int),
        ("init", ctypes.c_int),
        ("best", ctypes.c_int),
        ("nsets", ctypes.c_int),
        ("sets", ctypes.POINTER(ctypes.c_uint64)),
        ("set_cap", ctypes.c_int),
        ("ranks", ctypes.POINTER(ctypes.c_int)),
        ("order", ctypes.POINTER(ctypes.c_int)),
        ("set_of_col", ctypes.POINTER(ctypes.c_int)),
    ]


class ElimStats(ctypes.Structure):
    _fields_ = [("sets_ms", ctypes.c_double), ("bound_ms", ctypes.c_double), ("sets_bytes", ctypes.c_uint64),
                ("jobs", ctypes.c_int), ("trials", ctypes.c_int)]


class DistfindBatchItem(ctypes.Structure):
    """Layout of DistfindBatchItem (include/distfind_capi.h)."""
    _fields_ = [("kind", ctypes.c_int), ("A", ctypes.POINTER(ctypes.c_uint8)), ("ar", ctypes.c_int),
                ("ac", ctypes.c_int), ("B", ctypes.POINTER(ctypes.c_uint8)), ("br", ctypes.c_int),
                ("bc", ctypes.c_int), ("which", ctypes.c_char), ("order", ctypes.POINTER(ctypes.c_int)),
                ("scores", ctypes.POINTER(ctypes.c_float)), ("candidates", ctypes.c_int),
                ("proposal_seed", ctypes.c_uint64), ("work_cap", ctypes.c_int64)]


class DistfindBatchStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_int), ("capped", ctypes.c_int), ("codewords", ctypes.c_uint64),
                ("device_ms", ctypes.c_double), ("setup_s", ctypes.c_double), ("seconds", ctypes.c_double)]


class DistfindLevel(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int32) for f in ("side", "d", "active", "ngamma", "K", "inner", "outer", "kernel",
                                              "fd", "xd", "rpl", "mfma")] + [
        ("codewords", ctypes.c_uint64), ("medium_hits", ctypes.c_uint64), ("slow_hits", ctypes.c_uint64),
        ("ms", ctypes.c_double), ("kernel_ms", ctypes.c_double), ("ops_per_codeword", ctypes.c_double),
        ("side_ms", ctypes.c_double), ("side_codewords", ctypes.c_uint64), ("side_ops_per_codeword", ctypes.c_double)]


_lib = None


def load() -> ctypes.CDLL:
    """The engine library, loaded once. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C qubitserf_amd/csrc`"
        )
    L = ctypes.CDLL(str(LIB_PATH))
    rp = ctypes.POINTER(DistfindResult)
    orp = ctypes.POINTER(DistfindOpResult)
    c_int, c_char_p, c_char = ctypes.c_int, ctypes.c_char_p, ctypes.c_char

    def sig(name, res, args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args

    sig("distfind_css_distance", c_int,
        [u8p, c_int, c_int, u8p, c_int, c_int, c_char_p, c_char, c_char_p, c_int, c_int, rp])
    sig("distfind_classical_distance", c_int, [u8p, c_int, c_int, c_char_p, c_char_p, c_int, c_int, rp])
    sig("distfind_stabilizer_distance", c_int, [u8p, c_int, c_int, c_char_p, c_char, c_char_p, c_int, c_int, rp])
    sig("distfind_operator_weight", c_int,
        [u8p, c_int, c_int, u8p, c_int, c_int, u8p, u8p, c_int, c_char_p, c_char_p, c_int, c_int, orp])
    sig("distfind_subsystem_distance", c_int,
        [u8p, c_int, c_int, u8p, c_int, c_int, c_char_p, c_char, c_char_p, c_int, c_int, rp])
    sig("distfind_subsystem_stabilizer_distance", c_int,
        [u8p, c_int, c_int, c_char_p, c_char, c_char_p, c_int, c_int, rp])
    sig("distfind_stabilizer_operator_weight", c_int,
        [u8p, c_int, c_int, u8p, c_int, c_char_p, c_char_p, c_int, c_int, rp])
    sig("distfind_backend_available", c_int, [c_char_p])
    sig("distfind_version", c_char_p, [])
    sig("distfind_last_error", c_char_p, [])
    sig("distfind_last_trace", c_int, [ctypes.POINTER(DistfindLevel), c_int])
    sig("distfind_batch_distance", c_int,
        [ctypes.POINTER(DistfindBatchItem), c_int, ctypes.c_int64, c_int, c_int, rp])
    sig("distfind_last_batch_stats", None, [ctypes.POINTER(DistfindBatchStats)])
    hp = ctypes.c_void_p
    sig("distfind_session_new", hp, [ctypes.c_int64, c_int])
    sig("distfind_session_add", c_int, [hp, ctypes.POINTER(DistfindBatchItem), c_int, ctypes.POINTER(ctypes.c_int64)])
    sig("distfind_session_step", c_int, [hp, ctypes.POINTER(c_int)])
    sig("distfind_session_poll", c_int, [hp, ctypes.POINTER(ctypes.c_int64), rp, c_int])
    sig("distfind_session_active", c_int, [hp])
    sig("distfind_session_device_bytes", ctypes.c_int64, [hp])
    sig("distfind_set_sound_bounds", c_int, [c_int])
    sig("qsf_hip_solve_bytes", ctypes.c_uint64, [ctypes.c_uint64])
    sig("qsf_hip_mem_info", c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)])
    sig("distfind_session_stats", None, [hp, ctypes.POINTER(DistfindBatchStats)])
    sig("distfind_session_free", None, [hp])

    vp = ctypes.c_void_p
    sig("codeaut_bz_collect", vp,
        [u8p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32])
    sig("codeaut_bz_ok", ctypes.c_int32, [vp])
    sig("codeaut_bz_count", ctypes.c_int64, [vp])
    sig("codeaut_bz_combos", ctypes.c_int64, [vp])
    sig("codeaut_bz_copy_rows", None, [vp, u8p])
    sig("codeaut_bz_backend", c_char_p, [vp])
    sig("codeaut_bz_free", None, [vp])
    sig("codeaut_backend_available", ctypes.c_int32, [c_char_p])
    sig("codeaut_version", c_char_p, [])

    pp = ctypes.POINTER(EnumPlan)
    sig("qsf_hip_available", c_int, [])
    sig("qsf_hip_enumerate", c_int, [pp, ctypes.POINTER(c_int)])
    sig("qsf_hip_enumerate_floor", c_int, [pp, c_int, ctypes.POINTER(c_int)])
    sig("qsf_hip_enumerate_floor_noalloc", c_int, [pp, c_int, ctypes.POINTER(c_int)])
    sig("qsf_hip_noalloc_scratch_bytes", ctypes.c_size_t, [])
    sig("qsf_hip_plan_variant", c_int, [pp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(c_int)])
    sig("qsf_hip_collect", c_int, [pp, c_int, ctypes.POINTER(u64p), ctypes.POINTER(ctypes.c_int64)])
    sig("qsf_hip_collect_upto", c_int, [pp, c_int, ctypes.POINTER(u64p), ctypes.POINTER(ctypes.c_int64)])
    sig("qsf_hip_last_collect_stats", None, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)])
    sig("qsf_hip_free", None, [vp])
    sig("qsf_hip_last_stats", None, [ctypes.POINTER(HipStats)])
    sig("qsf_hip_last_error", c_char_p, [])
    sig("qsf_hip_set_kernel_policy", c_int, [c_int])
    sig("qsf_hip_is_bound", c_int, [u64p, c_int, c_int, u64p, c_int, c_int, c_int, ctypes.c_uint64, c_int,
                                    ctypes.POINTER(c_int)])
    sig("qsf_hip_eliminate", c_int, [ctypes.POINTER(ElimJob), c_int, c_int, ctypes.c_uint64, c_int])
    sig("qsf_hip_last_elim_stats", None, [ctypes.POINTER(ElimStats)])
    _lib = L
    return L


def last_error() -> str:
    return (load().distfind_last_error() or b"").decode()
