"""Meet-in-the-middle on the device (method="mitm"): an independent certifier, checked
against the reference's own BZ answers (golden cases small enough for the method: at most
63 parity checks, n <= 128, a weight split of at most 2^27 parts), the oracle's MITM, and
known distances; single CSS sides and the interleaved minimum."""
import numpy as np
import pytest

from qubitserf_amd import codes
from tests import oracle as O
from tests.golden.fixtures import dec, load

pytestmark = pytest.mark.gpu
GOLD = load()


def _small(H):
    return H.shape[1] <= 64 and codes.gf2_rank(H) <= 63


CLASSICAL = [c for c in GOLD["classical"] if _small(dec(c["H"]))]


def test_enough_golden_cases():
    assert len(CLASSICAL) >= 20


@pytest.mark.parametrize("case", CLASSICAL, ids=[c["name"] for c in CLASSICAL])
def test_classical_golden(hip, case):
    from qubitserf_amd import distfind as df
    H = dec(case["H"])
    r = df.api.classical_distance_raw(H, "mitm", "gpu")
    assert r.backend.decode() == "hip"
    assert (r.distance, bool(r.proven)) == (case["ref"]["distance"], bool(case["ref"]["proven"]))
    if r.distance > 0:
        assert r.lower_bound == r.distance


@pytest.mark.parametrize("seed", range(6))
def test_random_vs_oracle_mitm(hip, seed):
    from qubitserf_amd import distfind as df
    rng = np.random.default_rng(100 + seed)
    H = codes.random_rate_code(rng, 30, 60)
    got = df.classical_distance(H, method="mitm", backend="gpu")
    assert got == O.mitm_distance(0, H) == O.classical_distance(H).distance


@pytest.mark.parametrize("name,HxHz,d", [("steane", codes.steane(), 3), ("shor", codes.shor(), 3),
                                         ("toric4", codes.toric(4), 4), ("surface5", codes.surface(5), 5),
                                         ("toric5", codes.toric(5), 5)])
def test_css_known(hip, name, HxHz, d):
    from qubitserf_amd import distfind as df
    Hx, Hz = HxHz
    for which in ("min", "z", "x"):
        assert df.css_distance(Hx, Hz, method="mitm", which=which, backend="gpu") == d, which
    raw = df.api.css_distance_raw(Hx, Hz, "mitm", "M", "gpu")
    assert raw.proven and raw.lower_bound == d


def test_css_sides_vs_oracle(hip):
    from qubitserf_amd import distfind as df
    Hx, Hz = codes.bb_family(6)
    for kind, which in ((1, "z"), (2, "x")):
        assert df.css_distance(Hx, Hz, method="mitm", which=which, backend="gpu") == O.mitm_distance(kind, Hx, Hz)


def test_limits_are_errors(hip):
    from qubitserf_amd import distfind as df
    H = codes.random_full_rank(np.random.default_rng(1), 70, 140)   # 70 parity checks
    with pytest.raises(ValueError, match="parity checks"):
        df.classical_distance(H, method="mitm", backend="gpu")
