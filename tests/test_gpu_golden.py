"""The device engine, through the C ABI, against the reference's own answers
(tests/golden/reference_cases.json, written by tests/golden/make_golden.py from the
reference CPU library). Distances and the proven flag must be identical; levels and lower
bounds may differ only through the random information-set upper bound, which the
reference seeds per hardware thread (cpp/distfind/bz.cpp:30-63)."""
import hashlib

import numpy as np
import pytest

from tests.golden.fixtures import dec, load

pytestmark = pytest.mark.gpu
GOLD = load()


@pytest.mark.parametrize("case", GOLD["classical"], ids=[c["name"] for c in GOLD["classical"]])
def test_classical(hip, case):
    from qubitserf_amd import distfind as df

    r = df.api.classical_distance_raw(dec(case["H"]), "bz", "gpu")
    assert (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"])
    assert r.backend.decode() == "hip"


@pytest.mark.parametrize("case", GOLD["css"], ids=[c["name"] for c in GOLD["css"]])
def test_css(hip, case):
    from qubitserf_amd import distfind as df

    Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
    for w in ("M", "Z", "X"):
        r = df.api.css_distance_raw(Hx, Hz, "bz", w, "gpu")
        assert (r.distance, r.proven) == (case["ref"][w]["distance"], case["ref"][w]["proven"]), w


@pytest.mark.parametrize("case", GOLD["stabilizer"], ids=[c["name"] for c in GOLD["stabilizer"]])
def test_stabilizer(hip, case):
    from qubitserf_amd import distfind as df

    r = df.api.stabilizer_distance_raw(dec(case["S"]), "bz", "M", "gpu")
    assert (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"])


@pytest.mark.parametrize("case", GOLD["collect"], ids=[c["name"] for c in GOLD["collect"]])
def test_collect(hip, case):
    from qubitserf_amd.codeaut import bz_collect

    rows, combos = bz_collect(dec(case["G"]), case["p"], case["keep"])
    assert rows.shape[0] == case["ref"]["count"]
    assert combos == case["ref"]["combos"]
    assert hashlib.sha256(np.packbits(rows, axis=1).tobytes()).hexdigest() == case["ref"]["sha256"]
