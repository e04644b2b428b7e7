"""The gfx950 engine as the reference's own GPU plugin, on the device.

oracle/_ref/libqubitserf_ref_hip.so is the reference's unchanged CPU sources (its drivers
bz.cpp / css.cpp / stab.cpp, its C ABI capi.cpp, its select_backend) with the HIP plugin
qubitserf_amd/plugin/backend_hip.cpp in the accelerator slot the CUDA translation unit
fills (oracle/ref.mk). Every "gpu" level the reference's level loop hands to
Backend::enumerate / Backend::collect then runs through qsf_hip_enumerate /
qsf_hip_collect on the device. Results must equal the reference CPU library's answers in
the golden fixtures; the result's backend field must name the plugin."""
import hashlib

import numpy as np
import pytest

from tests import oracle as O
from tests.golden.fixtures import dec, load

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.ref_hip_available(), reason="oracle/_ref/libqubitserf_ref_hip.so not built")]
GOLD = load()
HEAVY = load("reference_heavy.json")


@pytest.fixture(scope="module")
def plugged(hip):
    L = O.ref_hip()
    assert L.ref_hip_plugin_name() == b"hip"
    # the reference's own availability rules (cpp/distfind/capi.cpp:290-299) see the plugin
    assert (L.distfind_backend_available(b"gpu"), L.distfind_backend_available(b"auto"),
            L.distfind_backend_available(b"cpu")) == (1, 1, 1)
    assert L.codeaut_backend_available(b"gpu") == 1
    return L


def _device_ran(hip):
    return hip.last_stats()["codewords"] > 0


@pytest.mark.parametrize("case", GOLD["classical"], ids=[c["name"] for c in GOLD["classical"]])
def test_classical_through_reference_driver(plugged, hip, case):
    print(f"plugin classical {case['name']}", flush=True)
    r = O.ref_classical_distance(dec(case["H"]), L=plugged, backend="gpu")
    assert r.backend == b"hip"
    assert (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"])


@pytest.mark.parametrize("case", GOLD["css"], ids=[c["name"] for c in GOLD["css"]])
def test_css_through_reference_driver(plugged, hip, case):
    print(f"plugin css {case['name']}", flush=True)
    Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
    for w in ("M", "Z", "X"):
        r = O.ref_css_distance(Hx, Hz, w, L=plugged, backend="gpu")
        assert r.backend == b"hip"
        assert (r.distance, r.proven) == (case["ref"][w]["distance"], case["ref"][w]["proven"]), w


@pytest.mark.parametrize("case", GOLD["stabilizer"], ids=[c["name"] for c in GOLD["stabilizer"]])
def test_stabilizer_through_reference_driver(plugged, hip, case):
    print(f"plugin stabilizer {case['name']}", flush=True)
    r = O.ref_stabilizer_distance(dec(case["S"]), L=plugged, backend="gpu")
    assert r.backend == b"hip"
    assert (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"])


@pytest.mark.parametrize("case", GOLD["collect"], ids=[c["name"] for c in GOLD["collect"]])
def test_collect_through_reference_driver(plugged, hip, case):
    """cpp/codeaut/bz.cpp with backend 1 ("auto") selects the plugin for every subset size."""
    print(f"plugin collect {case['name']}", flush=True)
    rows, combos, ok, backend = O.ref_collect(dec(case["G"]), case["p"], case["keep"], L=plugged, backend=1,
                                              with_meta=True)
    assert (ok, backend) == (1, "hip")
    assert rows.shape[0] == case["ref"]["count"] and combos == case["ref"]["combos"]
    assert hashlib.sha256(np.packbits(rows, axis=1).tobytes()).hexdigest() == case["ref"]["sha256"]


def test_gross_code_through_reference_driver(plugged, hip):
    """Config 2 end to end through the reference's own interleaved driver: its information
    sets, bounds and level loop, every level on the device (the reference enumerates each
    level to the end; no early-stop floor through this interface)."""
    case = next(c for c in HEAVY["css"] if c["name"] == "bb_l12")
    Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
    print("plugin gross code", flush=True)
    r = O.ref_css_distance(Hx, Hz, "M", L=plugged, backend="gpu")
    assert r.backend == b"hip"
    assert (r.distance, r.proven) == (case["ref"]["M"]["distance"], case["ref"]["M"]["proven"])
    assert _device_ran(hip)
    print(f"plugin gross code: d={r.distance} proven={r.proven} in {r.seconds:.3f} s", flush=True)
