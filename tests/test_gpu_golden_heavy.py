"""The device engine against the reference CPU library's certified distances for the heavy
bivariate-bicycle instances (tests/golden/reference_heavy.json, written by
tests/golden/make_golden_heavy.py; the gross code [[144, 12, 12]] took the reference
about 80 minutes on 6 cores). Distance and the proven flag must be identical."""
import pytest

from tests.golden.fixtures import dec, load

pytestmark = pytest.mark.gpu
HEAVY = load("reference_heavy.json")


def test_fixture_holds_the_gross_code():
    names = [c["name"] for c in HEAVY["css"]]
    assert "bb_l12" in names
    gross = HEAVY["css"][names.index("bb_l12")]
    assert gross["Hx"]["shape"] == [72, 144] and gross["ref"]["M"]["distance"] == 12


@pytest.mark.parametrize("case", HEAVY["css"], ids=[c["name"] for c in HEAVY["css"]])
def test_heavy_css(hip, case):
    from qubitserf_amd import codes
    from qubitserf_amd import distfind as df

    Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
    ell = int(case["name"].split("_l")[1])
    bx, bz = codes.bb_family(ell)
    assert (Hx == bx).all() and (Hz == bz).all(), "fixture and code builder disagree"
    r = df.api.css_distance_raw(Hx, Hz, "bz", "M", "gpu")
    want = case["ref"]["M"]
    assert (r.distance, r.proven) == (want["distance"], bool(want["proven"]))
    assert r.backend.decode() == "hip"
