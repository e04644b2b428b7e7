"""Host-side checks that need no GPU: the C ABI library loads and exports every function
declared in include/*.h, the Python mirror validates arguments like the reference
(python/qubitserf/distfind/api.py:66-72; tests/distfind/test_distance.py:210-226), and the
code constructions have their textbook parameters."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from qubitserf_amd import _native, codes, engine
from qubitserf_amd import distfind as df

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:qsf_hip|distfind|codeaut)_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for must in ("distfind_css_distance", "distfind_classical_distance", "distfind_stabilizer_distance",
                 "distfind_operator_weight", "distfind_subsystem_distance", "distfind_subsystem_stabilizer_distance",
                 "distfind_stabilizer_operator_weight", "distfind_backend_available", "distfind_version",
                 "codeaut_bz_collect", "codeaut_bz_ok", "codeaut_bz_count", "codeaut_bz_combos",
                 "codeaut_bz_copy_rows", "codeaut_bz_backend", "codeaut_bz_free", "codeaut_backend_available",
                 "qsf_hip_enumerate", "qsf_hip_collect", "qsf_hip_available"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_symbol(name):
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    assert hasattr(lib, name), name


def test_struct_layouts():
    # DistfindResult as cpp/distfind/capi.h:12-19; qsf_enum_plan as qsf::EnumPlan
    assert ctypes.sizeof(_native.DistfindResult) == 40
    assert _native.EnumPlan.gamma.offset == 24 and _native.EnumPlan.check.offset == 40
    assert _native.EnumPlan.current_best.offset == 64 and _native.EnumPlan.buffers_key.offset == 72
    assert ctypes.sizeof(_native.EnumPlan) == 80


@pytest.mark.parametrize("backend", ["metal", "cuda", "tpu"])
def test_backend_names_validated(backend):
    with pytest.raises(ValueError, match="auto/cpu/gpu"):
        df.css_distance(*codes.steane(), backend=backend)


def test_which_validated():
    with pytest.raises(ValueError, match="min/z/x"):
        df.css_distance(*codes.steane(), which="y")


def test_cpu_backend_is_not_served_here():
    with pytest.raises(RuntimeError):
        df.classical_distance(codes.hamming_parity(3), backend="cpu")


def test_gpu_request_without_device_raises():
    if engine.available():
        pytest.skip("a gfx950 device is present")
    with pytest.raises(RuntimeError, match="no GPU backend"):
        df.css_distance(*codes.steane(), backend="gpu")
    assert df.available_backends() == []


def test_symplectic_shape_validated():
    with pytest.raises(ValueError):
        df.stabilizer_distance(np.zeros((2, 5), np.uint8))


def test_c_abi_error_codes_without_device():
    L = _native.load()
    out = _native.DistfindResult()
    H = np.ascontiguousarray(codes.hamming_parity(3))
    p = H.ctypes.data_as(_native.u8p)
    assert L.distfind_classical_distance(None, 3, 7, b"bz", b"gpu", 0, 0, ctypes.byref(out)) == 1
    assert L.distfind_classical_distance(p, 3, 7, b"cc", b"gpu", 0, 0, ctypes.byref(out)) == 4
    assert L.distfind_classical_distance(p, 3, 7, b"mitm", b"gpu", 0, 0, ctypes.byref(out)) == 5   # no device here
    S6 = np.array([[1, 0, 0, 0, 1, 0], [0, 1, 0, 0, 0, 1]], np.uint8)   # non-CSS: symplectic mitm, no device here
    assert L.distfind_stabilizer_distance(S6.ctypes.data_as(_native.u8p), 2, 6, b"mitm", b"M", b"gpu", 0, 0,
                                          ctypes.byref(out)) == 5
    assert L.distfind_stabilizer_distance(S6.ctypes.data_as(_native.u8p), 2, 6, b"cc", b"M", b"gpu", 0, 0,
                                          ctypes.byref(out)) == 3   # connected cluster has no non-CSS form
    assert L.distfind_classical_distance(p, 3, 7, b"bz", b"cpu", 0, 0, ctypes.byref(out)) == 5
    S = np.zeros((2, 5), np.uint8)
    assert L.distfind_stabilizer_distance(S.ctypes.data_as(_native.u8p), 2, 5, b"bz", b"M", b"gpu", 0, 0,
                                          ctypes.byref(out)) == 2
    assert L.distfind_operator_weight(p, 3, 7, p, 3, 7, p, p, 7, b"bz", b"gpu", 0, 0, None) == 4


def test_pack_unpack_roundtrip():
    rng = np.random.default_rng(0)
    for n in (1, 63, 64, 65, 130, 200):
        M = rng.integers(0, 2, (7, n), dtype=np.uint8)
        P = engine.pack_rows(M)
        assert P.shape == (7, (n + 63) // 64)
        assert np.array_equal(engine.unpack_rows(P, n), M)


def test_constructions():
    Hx, Hz = codes.toric(5)
    assert Hx.shape == (25, 50) and not ((Hx @ Hz.T) % 2).any()
    Hx, Hz = codes.surface(4)
    assert Hx.shape[1] == 16 + 9 and not ((Hx @ Hz.T) % 2).any()
    Hx, Hz = codes.gross()
    n = Hx.shape[1]
    assert n == 144 and not ((Hx @ Hz.T) % 2).any()
    k = n - codes.gf2_rank(Hx) - codes.gf2_rank(Hz)
    assert k == 12
    for ell, kk in ((6, 12), (9, 8), (12, 12)):
        Hx, Hz = codes.bb_family(ell)
        assert 2 * ell * 6 - codes.gf2_rank(Hx) - codes.gf2_rank(Hz) == kk
    H = codes.config1_parity(0)
    assert H.shape == (32, 64) and codes.gf2_rank(H) == 32
    Gx, Gz = codes.quantum_reed_muller(1, 5)
    assert Gx.shape == (6, 32)
    qc = codes.quasi_cyclic(np.random.default_rng(1), 20)
    assert qc.shape == (60, 120) and (qc.sum(axis=1) == 18).all() and (qc.sum(axis=0) == 9).all()


def test_config3_law():
    batch = codes.config3_batch(0, 16)
    for H in batch:
        m, n = H.shape
        assert 64 <= n <= 160
        assert 0.29 <= (n - m) / n <= 0.61
        assert codes.gf2_rank(H) == m


def test_matroid_partition_matches_oracle():
    """The host matroid partition (product) equals the oracle's restatement of the
    reference's (cpp/distfind/css.cpp:274-328) class for class, on generators where the
    greedy peel is short of full-rank sets and the partition decides the lower bounds."""
    from tests import oracle as O
    L = _native.load()
    L.distfind_matroid_parts.argtypes = [_native.u8p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    ip = ctypes.POINTER(ctypes.c_int)
    O.lib().oracle_matroid_parts.argtypes = [O.u64p, ctypes.c_int, ctypes.c_int, ip]
    rng = np.random.default_rng(12)
    cases = [codes.gf2_kernel(codes.random_rate_code(rng, 40, 120, 0.3, 0.45)) for _ in range(25)]
    cases += [codes.gf2_kernel(codes.quasi_cyclic(rng, L)) for L in (5, 7, 9)]
    cases += [codes.gf2_kernel(codes.hamming_parity(4)), codes.reed_muller_generator(1, 5)]
    for G in cases:
        G = np.ascontiguousarray(G, np.uint8)
        K, n = G.shape
        a = np.zeros(n, np.int32)
        assert L.distfind_matroid_parts(G.ctypes.data_as(_native.u8p), K, n, a.ctypes.data_as(ip)) == 0
        b = np.zeros(n, np.int32)
        packed = engine.pack_rows(G)
        assert O.lib().oracle_matroid_parts(packed.ctypes.data_as(O.u64p), K, n, b.ctypes.data_as(ip)) >= 0
        assert np.array_equal(a, b), (K, n)
