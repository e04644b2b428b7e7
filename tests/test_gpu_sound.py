"""The sound level schedule on the device against the oracle's sound schedule: every
golden classical and CSS case, and the batched workload's code 32, where the reference's
schedule certifies 12 above the true distance 11 (tests/test_cpu_soundness.py)."""
import pytest

from tests import oracle as O
from tests.golden.fixtures import dec, load

pytestmark = pytest.mark.gpu
GOLD = load()


@pytest.fixture
def sound(hip):
    from qubitserf_amd import distfind as df

    with df.sound_bounds(), O.sound():
        yield df


def test_code_32_certified_at_11(sound):
    from qubitserf_amd import codes

    H = codes.config3_batch(0, 256)[32]
    r = sound.api.classical_distance_raw(H, "bz", "gpu")
    assert (r.distance, r.proven) == (11, 1)
    assert r.lower_bound >= 11
    b = sound.batch_distance([("classical", H)])[0]
    assert (b.distance, b.proven) == (11, 1)


def test_default_schedule_is_the_reference_schedule(hip):
    """Off (the default), the device follows the reference's schedule: the same levels and
    the same information sets per level as the oracle's reference-schedule restatement."""
    from qubitserf_amd import distfind as df

    for case in GOLD["classical"][:16]:
        H = dec(case["H"])
        r = df.api.classical_distance_raw(H, "bz", "gpu")
        assert (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"]), case["name"]
        act = [t["active"] for t in df.last_trace()]
        gamma, ranks, _, _ = O.side_plan(0, H)
        K = gamma.shape[1]
        assert act == [sum(1 for rk in ranks if (d + 1) - (K - rk) > 0) for d in range(1, len(act) + 1)], case["name"]


@pytest.mark.parametrize("case", GOLD["classical"], ids=[c["name"] for c in GOLD["classical"]])
def test_classical_sound(sound, case):
    H = dec(case["H"])
    r = sound.api.classical_distance_raw(H, "bz", "gpu")
    o = O.classical_distance(H)
    assert (r.distance, r.proven) == (o.distance, o.proven) == (case["ref"]["distance"], case["ref"]["proven"])


@pytest.mark.parametrize("case", GOLD["css"], ids=[c["name"] for c in GOLD["css"]])
def test_css_sound(sound, case):
    Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
    for w in ("M", "Z", "X"):
        r = sound.api.css_distance_raw(Hx, Hz, "bz", w, "gpu")
        o = O.css_distance(Hx, Hz, w)
        assert (r.distance, r.proven) == (o.distance, o.proven), (case["name"], w)
