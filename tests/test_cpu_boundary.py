"""The drop-in boundary, on the CPU.

* Every symbol the reference's ctypes layer attaches without a hasattr guard is exported
  by libqubitserf_amd.so (otherwise the replacement-library route fails at import).
* The HIP plugin (qubitserf_amd/plugin/backend_hip.cpp) compiles against the reference's
  own qsf/backend.hpp, with static_asserts that qsf::EnumPlan and qsf_enum_plan agree field
  for field (offset and size of all 14 fields).
* Without a device, the reference built with that plugin in its accelerator slot
  (oracle/_ref/libqubitserf_ref_hip.so) keeps the reference's contract: the plugin's
  enumerate/collect fall back to the reference's CPU backend and match the oracle,
  select_backend falls back to "cpu", and distfind_backend_available answers as
  cpp/distfind/capi.cpp:290-299 does.
* Malformed inputs give a return code, never an exception across the C ABI.
"""
import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from qubitserf_amd import _native
from tests import oracle as O

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")

# (symbol, where the reference's Python binds it without a guard)
UNGUARDED = [
    ("distfind_css_distance", "python/qubitserf/distfind/_native.py:49"),
    ("distfind_classical_distance", "python/qubitserf/distfind/_native.py:57"),
    ("distfind_operator_weight", "python/qubitserf/distfind/_native.py:64"),
    ("distfind_subsystem_distance", "python/qubitserf/distfind/_native.py:73"),
    ("distfind_stabilizer_distance", "python/qubitserf/distfind/_native.py:81"),
    ("distfind_subsystem_stabilizer_distance", "python/qubitserf/distfind/_native.py:88"),
    ("distfind_stabilizer_operator_weight", "python/qubitserf/distfind/_native.py:95"),
    ("distfind_backend_available", "python/qubitserf/distfind/_native.py:103"),
    ("distfind_version", "python/qubitserf/distfind/_native.py:105"),
    ("qaut_leon_run", "python/qubitserf/codeaut/_native.py:31"),
    ("qaut_leon_ok", "python/qubitserf/codeaut/_native.py:39"),
    ("qaut_leon_dim", "python/qubitserf/codeaut/_native.py:39"),
    ("qaut_leon_n", "python/qubitserf/codeaut/_native.py:39"),
    ("qaut_leon_num_codewords", "python/qubitserf/codeaut/_native.py:39"),
    ("qaut_leon_num_classes", "python/qubitserf/codeaut/_native.py:40"),
    ("qaut_leon_num_gens", "python/qubitserf/codeaut/_native.py:40"),
    ("qaut_leon_num_factors", "python/qubitserf/codeaut/_native.py:40"),
    ("qaut_leon_copy_gens", "python/qubitserf/codeaut/_native.py:54"),
    ("qaut_leon_copy_factors", "python/qubitserf/codeaut/_native.py:56"),
    ("qaut_leon_copy_weights", "python/qubitserf/codeaut/_native.py:58"),
    ("qaut_leon_free", "python/qubitserf/codeaut/_native.py:60"),
]
# attached only under hasattr(): the collect ABI is provided, Leon's optional extras are not
GUARDED_PROVIDED = [
    ("codeaut_bz_collect", "python/qubitserf/codeaut/_native.py:64"),
    ("codeaut_bz_ok", "python/qubitserf/codeaut/_native.py:72"),
    ("codeaut_bz_count", "python/qubitserf/codeaut/_native.py:74"),
    ("codeaut_bz_combos", "python/qubitserf/codeaut/_native.py:76"),
    ("codeaut_bz_copy_rows", "python/qubitserf/codeaut/_native.py:78"),
    ("codeaut_bz_backend", "python/qubitserf/codeaut/_native.py:80"),
    ("codeaut_bz_free", "python/qubitserf/codeaut/_native.py:82"),
    ("codeaut_backend_available", "python/qubitserf/codeaut/_native.py:85"),
]
GUARDED_ABSENT = ["qaut_leon_run_ex", "qaut_leon_selector", "qaut_leon_modulus", "qaut_leon_residue",
                  "qaut_leon_enumeration_ns", "qaut_leon_search_ns", "qaut_leon_num_incidences"]


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("sym,where", UNGUARDED + GUARDED_PROVIDED, ids=[s for s, _ in UNGUARDED + GUARDED_PROVIDED])
def test_reference_bound_symbol_is_exported(sym, where):
    assert sym in _exported(), f"{sym} (bound at {where}) is not exported"


def test_symbol_list_matches_reference_source():
    """The lists above are complete: every distfind_/codeaut_/qaut_ name the reference's two
    binding modules mention is in one of them, at the cited line."""
    if not (REF / "python/qubitserf/distfind/_native.py").exists():
        pytest.skip("reference tree not present")
    cited = {}
    for sym, where in UNGUARDED + GUARDED_PROVIDED:
        cited[sym] = where
    listed = set(cited) | set(GUARDED_ABSENT)
    for mod in ("python/qubitserf/distfind/_native.py", "python/qubitserf/codeaut/_native.py"):
        lines = (REF / mod).read_text().splitlines()
        names = {m for line in lines for m in re.findall(r"\b((?:distfind|codeaut|qaut_leon)_\w*[a-z])\b", line)}
        names -= {"distfind_result", "codeaut_lib_path"}
        assert names <= listed, sorted(names - listed)
    for sym, where in cited.items():
        f, line = where.split(":")
        assert sym in (REF / f).read_text().splitlines()[int(line) - 1], (sym, where)


def test_plugin_compiles_against_reference_headers():
    """backend_hip.cpp against the reference's qsf/backend.hpp: its static_asserts pin the
    EnumPlan layout (cpp/qsf/backend.hpp:21-42) field by field."""
    if not (REF / "cpp/qsf/backend.hpp").exists():
        pytest.skip("reference tree not present")
    gxx = shutil.which("g++")
    assert gxx
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{REF / 'cpp'}",
                        f"-I{ROOT / 'include'}", str(ROOT / "qubitserf_amd/plugin/backend_hip.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _no_device():
    return not _native.load().qsf_hip_available()


needs_plugin_build = pytest.mark.skipif(not O.ref_hip_available(), reason="oracle/_ref/libqubitserf_ref_hip.so not built")


@needs_plugin_build
def test_plugin_falls_back_to_reference_cpu_without_device():
    if not _no_device():
        pytest.skip("a device is present: the GPU suite covers the plugin there")
    L = O.ref_hip()
    assert L.ref_hip_plugin_name() == b"hip"
    rng = np.random.default_rng(17)
    for _ in range(12):
        K = int(rng.integers(6, 20)); stride = int(rng.integers(1, 3)); d = int(rng.integers(1, 5))
        ng = int(rng.integers(1, 3)); kcheck = int(rng.integers(0, 3))
        best = [O.WEIGHT_NONE, 40, 12][int(rng.integers(0, 3))]
        gamma = rng.integers(0, 2**64, size=(ng, K, stride), dtype=np.uint64)
        check = rng.integers(0, 2**64, size=(max(kcheck, 1), stride), dtype=np.uint64)
        got = L.ref_hip_plugin_enumerate(stride * 64, stride, K, d, ng, gamma.ctypes.data_as(O.u64p), kcheck,
                                         check.ctypes.data_as(O.u64p), best)
        want = O.enumerate_plan(gamma, stride * 64, d, check=check[:kcheck] if kcheck else None, current_best=best)
        assert got == want
    # collect: complete (ok) and equal to the oracle's multiset
    gamma = rng.integers(0, 2**64, size=(1, 14, 1), dtype=np.uint64) & np.uint64(0xFFFFFF)
    cap = 1 << 14
    out = np.zeros((cap, 1), np.uint64)
    ok = ctypes.c_int(0)
    rows = L.ref_hip_plugin_collect(24, 1, 14, 3, gamma.ctypes.data_as(O.u64p), 9, out.ctypes.data_as(O.u64p), cap,
                                     ctypes.byref(ok))
    want = O.collect_plan(gamma, 24, 3, 9)
    assert ok.value == 1 and rows == want.shape[0]
    assert sorted(out[:rows, 0].tolist()) == sorted(np.asarray(want).reshape(-1).tolist())


@needs_plugin_build
def test_reference_selection_semantics_with_plugin_without_device():
    """cpp/distfind/capi.cpp:290-299 and backend_cpu.cpp:84-93, unchanged, with the plugin in
    the accelerator slot: "auto" and "cpu" are always available, "gpu" only with a device, and
    a "gpu" solve without one runs on the reference's CPU backend."""
    if not _no_device():
        pytest.skip("a device is present")
    L = O.ref_hip()
    assert (L.distfind_backend_available(b"auto"), L.distfind_backend_available(b"cpu"),
            L.distfind_backend_available(b"gpu")) == (1, 1, 0)
    assert L.codeaut_backend_available(b"cpu") == 1
    from tests.golden.fixtures import dec, load

    case = load()["classical"][0]
    r = O.ref_classical_distance(dec(case["H"]), L=L, backend="gpu")
    assert r.backend == b"cpu" and (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"])


def test_malformed_batch_item_returns_code():
    """A negative size in a batch item is rejected with return code 1 (null/invalid), not
    an exception across the C ABI."""
    L = _native.load()
    H = np.ones((2, 4), np.uint8)
    item = _native.DistfindBatchItem()
    item.kind = 0
    item.A = H.ctypes.data_as(_native.u8p)
    item.ar, item.ac = -3, 4
    out = (_native.DistfindResult * 1)()
    assert L.distfind_batch_distance(ctypes.byref(item), 1, 0, 0, 0, out) == 1
    assert b"shape" in L.distfind_last_error()
    r = _native.DistfindResult()
    assert L.distfind_classical_distance(H.ctypes.data_as(_native.u8p), 2, -4, b"bz", b"gpu", 0, 0,
                                         ctypes.byref(r)) == 1
    assert L.distfind_stabilizer_distance(H.ctypes.data_as(_native.u8p), -2, 4, b"bz", b"M", b"gpu", 0, 0,
                                          ctypes.byref(r)) == 1


def test_leon_handle_reports_not_ok():
    L = ctypes.CDLL(str(_native.LIB_PATH))
    L.qaut_leon_run.restype = ctypes.c_void_p
    L.qaut_leon_run.argtypes = [_native.u8p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    for f in ("qaut_leon_ok", "qaut_leon_n", "qaut_leon_num_gens"):
        getattr(L, f).argtypes = [ctypes.c_void_p]
    L.qaut_leon_free.argtypes = [ctypes.c_void_p]
    G = np.eye(3, 6, dtype=np.uint8)
    h = L.qaut_leon_run(G.ctypes.data_as(_native.u8p), 3, 6, 30, 1)
    assert h
    try:
        assert (L.qaut_leon_ok(h), L.qaut_leon_n(h), L.qaut_leon_num_gens(h)) == (0, 6, 0)
    finally:
        L.qaut_leon_free(h)
