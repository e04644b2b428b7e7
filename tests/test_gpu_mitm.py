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


# ---- non-CSS stabilizer codes: the symplectic meet in the middle -----------------------

STAB = [c for c in GOLD["stabilizer"]]


@pytest.mark.parametrize("case", STAB, ids=[c["name"] for c in STAB])
def test_symplectic_golden(hip, case):
    """Symplectic MITM against the reference's certified distances of its non-CSS fixtures."""
    from qubitserf_amd import distfind as df
    S = dec(case["S"])
    r = df.api.stabilizer_distance_raw(S, "mitm", "M", "gpu")
    assert r.backend.decode() == "hip"
    assert (r.distance, bool(r.proven)) == (case["ref"]["distance"], bool(case["ref"]["proven"]))
    if O.ref_available():   # the reference's own symplectic meet in the middle agrees
        assert O.ref_stabilizer_distance(S, method="mitm").distance == r.distance


def _css_symplectic(Hx, Hz):
    """[z | x] stabilizer matrix of a CSS code: Z-type rows of Hz, X-type rows of Hx."""
    q = Hx.shape[1]
    return np.vstack([np.hstack([Hz, np.zeros_like(Hz)]), np.hstack([np.zeros_like(Hx), Hx])]).astype(np.uint8)


def _local_cliffords(S, rng):
    """Each qubit gets a random single-qubit Clifford (H: swap its z and x columns; S:
    z += x, X -> Y). Every operator keeps the set of qubits it touches, so distances are
    unchanged, while the code stops being CSS."""
    S = S.copy()
    q = S.shape[1] // 2
    for j in range(q):
        for op in rng.choice(["H", "S"], size=int(rng.integers(1, 4))):
            if op == "H":
                S[:, [j, q + j]] = S[:, [q + j, j]]
            else:
                S[:, j] ^= S[:, q + j]
    return S


@pytest.mark.parametrize("name,seed", [("toric3", 0), ("steane", 1), ("shor", 2), ("surface3", 3), ("toric3", 4)])
def test_symplectic_local_clifford(hip, name, seed):
    """Symplectic MITM on CSS codes scrambled by local Cliffords: the known distance, equal to
    the device BZ (via the (a|b)->(a|b|a^b) isometry) and the oracle's BZ."""
    from qubitserf_amd import distfind as df
    Hx, Hz = {"toric3": lambda: codes.toric(3), "steane": codes.steane, "shor": codes.shor,
              "surface3": lambda: codes.surface(3)}[name]()
    want = 3
    S = _local_cliffords(_css_symplectic(Hx, Hz), np.random.default_rng(seed))
    got = df.stabilizer_distance(S, method="mitm", backend="gpu")
    assert got == want
    assert df.stabilizer_distance(S, method="bz", backend="gpu") == want
    assert O.stabilizer_distance(S).distance == want
    if O.ref_available():
        assert O.ref_stabilizer_distance(S, method="mitm").distance == want


def test_symplectic_five_qubit(hip):
    from qubitserf_amd import distfind as df
    S = np.array([[0, 1, 1, 0, 0, 1, 0, 0, 1, 0], [0, 0, 1, 1, 0, 0, 1, 0, 0, 1],
                  [0, 0, 0, 1, 1, 1, 0, 1, 0, 0], [1, 0, 0, 0, 1, 0, 1, 0, 1, 0]], np.uint8)
    r = df.api.stabilizer_distance_raw(S, "mitm", "M", "gpu")
    assert (r.distance, bool(r.proven), r.lower_bound) == (3, True, 3)
