"""Pins the oracle (the CPU restatement in oracle/bz_oracle.cpp) before it is trusted as
the checker of the device engine:

* against the reference's own answers (tests/golden/reference_cases.json, produced by the
  reference CPU library compiled from its sources),
* against brute force over all 2^K messages and an independent meet-in-the-middle search,
* against textbook distances and the closed form of the BZ lower bound,
* its level enumerator against a plain per-combination enumerator and a numpy one.
"""
import itertools
import math

import numpy as np
import pytest

from qubitserf_amd import codes
from tests import oracle as O
from tests.golden.fixtures import dec, load

GOLD = load()
HW = GOLD["meta"]["hw_threads"]
NONE = 1 << 30


def rec(r):
    return dict(distance=r.distance, lower_bound=r.lower_bound, proven=r.proven, levels=r.levels)


# ---------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("case", GOLD["classical"], ids=[c["name"] for c in GOLD["classical"]])
def test_classical_matches_reference(case):
    # same random-information-set worker count as the reference run => same level trace
    got = rec(O.classical_distance(dec(case["H"]), is_threads=HW))
    assert got == case["ref"]


@pytest.mark.parametrize("case", GOLD["css"], ids=[c["name"] for c in GOLD["css"]])
def test_css_matches_reference(case):
    Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
    for w in ("M", "Z", "X"):
        assert rec(O.css_distance(Hx, Hz, which=w, is_threads=HW)) == case["ref"][w], w


@pytest.mark.parametrize("case", GOLD["stabilizer"], ids=[c["name"] for c in GOLD["stabilizer"]])
def test_stabilizer_matches_reference(case):
    r = O.stabilizer_distance(dec(case["S"]), is_threads=HW)
    assert (r.distance, r.proven) == (case["ref"]["distance"], case["ref"]["proven"])


@pytest.mark.parametrize("case", GOLD["collect"], ids=[c["name"] for c in GOLD["collect"]])
def test_collect_matches_reference(case):
    """The reference dedups and sorts collected codewords; the oracle's level collector
    plus the same dedup must give the same set."""
    G = dec(case["G"])
    m, n = G.shape
    gamma = pack(G)[None]
    seen = set()
    for sw in range(1, min(case["p"], m) + 1):
        for row in O.collect_plan(gamma, n, sw, case["keep"]):
            seen.add(tuple(int(x) for x in row))
    assert len(seen) == case["ref"]["count"]
    if case["ref"]["rows"] is not None:
        want = {tuple(int(x) for x in r) for r in pack(dec(case["ref"]["rows"]))}
        assert seen == want


# ---------------------------------------------------------------- brute force / MITM

def pack(M):
    from qubitserf_amd.engine import pack_rows
    return pack_rows(M)


def numpy_min_weight(gen, det=None):
    """Min weight over nonzero messages (logical when det is given), by direct expansion."""
    K, n = gen.shape
    best = NONE
    for bits in range(1, 1 << K):
        msg = np.array([(bits >> i) & 1 for i in range(K)], dtype=np.uint8)
        cw = (msg @ gen) % 2
        if det is not None and not ((det @ cw) % 2).any():
            continue
        best = min(best, int(cw.sum()))
    return best


@pytest.mark.parametrize("seed", range(30))
def test_bz_equals_bruteforce_and_mitm_classical(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(12, 26))
    k = int(rng.integers(3, min(12, n - 2)))
    H = codes.random_full_rank(rng, n - k, n)
    bz = O.classical_distance(H).distance
    assert bz == O.bruteforce_distance(0, H) == O.mitm_distance(0, H)


@pytest.mark.parametrize("name,HxHz", [("steane", codes.steane()), ("shor", codes.shor()),
                                       ("toric3", codes.toric(3)), ("surface3", codes.surface(3)),
                                       ("surface4", codes.surface(4)), ("bb72", codes.bb_family(6))])
def test_bz_equals_bruteforce_and_mitm_css(name, HxHz):
    Hx, Hz = HxHz
    for kind, w in ((1, "Z"), (2, "X")):
        bz = O.css_distance(Hx, Hz, which=w).distance
        assert bz == O.mitm_distance(kind, Hx, Hz, 16)
        if O.problem_info(kind, Hx, Hz)["K"] <= 22:
            assert bz == O.bruteforce_distance(kind, Hx, Hz)


def test_bruteforce_is_independent_of_bz():
    """The brute-force checker itself, on a code computed by hand: the [7,4] Hamming code
    has 7 codewords of weight 3, so min weight 3; numpy expansion agrees."""
    H = codes.hamming_parity(3)
    assert O.bruteforce_distance(0, H) == 3
    G = np.array([[1, 0, 0, 0, 0, 1, 1], [0, 1, 0, 0, 1, 0, 1], [0, 0, 1, 0, 1, 1, 0], [0, 0, 0, 1, 1, 1, 1]],
                 dtype=np.uint8)
    assert not ((H @ G.T) % 2).any()
    assert numpy_min_weight(G) == 3


# ---------------------------------------------------------------- textbook distances

@pytest.mark.parametrize("L", range(3, 8))
def test_toric_and_surface(L):
    assert O.css_distance(*codes.toric(L)).distance == L
    assert O.css_distance(*codes.surface(L)).distance == L


@pytest.mark.parametrize("n", range(3, 10))
def test_repetition(n):
    assert O.classical_distance(codes.repetition_parity(n)).distance == n


def test_small_known_codes():
    assert O.classical_distance(codes.hamming_parity(3)).distance == 3
    assert O.classical_distance(codes.hamming_parity(4)).distance == 3
    ext = np.vstack([np.hstack([codes.hamming_parity(3), np.zeros((3, 1), np.uint8)]), np.ones((1, 8), np.uint8)])
    assert O.classical_distance(ext).distance == 4           # extended Hamming [8,4,4]
    assert O.css_distance(*codes.steane()).distance == 3
    assert O.css_distance(*codes.shor()).distance == 3
    assert O.css_distance(*codes.bb_family(6)).distance == 6  # [[72,12,6]]
    for m in (5, 6, 7):
        assert O.css_distance(*codes.quantum_reed_muller(1, m)).distance == 4


def test_five_qubit_code():
    z = np.array([0, 1, 1, 0, 0]); x = np.array([1, 0, 0, 1, 0])
    S = np.array([np.concatenate([np.roll(z, s), np.roll(x, s)]) for s in range(4)], dtype=np.uint8)
    assert O.stabilizer_distance(S).distance == 3


def test_min_is_min_of_sides():
    for Hx, Hz in (codes.toric(4), codes.shor(),
                   codes.hypergraph_product(codes.repetition_parity(3), codes.repetition_parity(4))):
        m = O.css_distance(Hx, Hz, "M").distance
        z = O.css_distance(Hx, Hz, "Z").distance
        x = O.css_distance(Hx, Hz, "X").distance
        assert m == min(z, x)


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("name,kind,mats", [
    ("config1", 0, (codes.config1_parity(0), None)),
    ("toric6Z", 1, codes.toric(6)),
    ("bb72Z", 1, codes.bb_family(6)),
    ("surface5X", 2, codes.surface(5)),
])
def test_lower_bound_closed_form(name, kind, mats):
    """After the last level d the reported lower bound is sum_g max(0, d+1-(K-r_g)),
    rounded up to even for an even code (cpp/distfind/bz.cpp:159-162)."""
    A, B = mats
    info = O.problem_info(kind, A, B)
    if kind == 0:
        r = O.classical_distance(A)
    else:
        r = O.css_distance(A, B, which="Z" if kind == 1 else "X")
    d = r.levels
    lb = sum(max(0, d + 1 - (info["K"] - rk)) for rk in info["ranks"])
    if info["even"]:
        lb += lb & 1
    assert r.lower_bound == lb
    assert r.proven and r.distance <= lb


def test_work_closed_form():
    """Level work is C(K, d) x active information sets: a systematic generator's
    enumeration at level d visits C(K, d) codewords, each once."""
    rng = np.random.default_rng(2)
    K, n, d = 9, 40, 3
    G = np.concatenate([np.eye(K, dtype=np.uint8), rng.integers(0, 2, (K, n - K), dtype=np.uint8)], axis=1)
    rows = O.collect_plan(pack(G)[None], n, d, n)   # keep = n: every nonzero codeword
    assert rows.shape[0] == math.comb(K, d)
    assert len({tuple(r) for r in rows.tolist()}) == math.comb(K, d)


# ---------------------------------------------------------------- the level enumerator

def rand_plan(rng, K, stride, ng, kcheck):
    gamma = rng.integers(0, 2**64, size=(ng, K, stride), dtype=np.uint64)
    check = rng.integers(0, 2**64, size=(kcheck, stride), dtype=np.uint64) if kcheck else None
    return gamma, check


def numpy_level(gamma, d, check, best, num_gamma):
    ng, K, stride = gamma.shape
    out = best
    for g in range(num_gamma):
        for comb in itertools.combinations(range(K), d):
            cw = np.bitwise_xor.reduce(gamma[g, list(comb)], axis=0)
            if check is not None:
                logical = any(bin(int(np.bitwise_xor.reduce(np.bitwise_and(c, cw)))).count("1") & 1 for c in check)
                if not logical:
                    continue
            w = sum(bin(int(x)).count("1") for x in cw)
            out = min(out, w)
    return out


def test_enumerator_vs_plain_and_numpy():
    rng = np.random.default_rng(0xC0FFEE)
    for i in range(60):
        K = int(rng.integers(4, 15)); stride = int(rng.integers(1, 4)); d = min(int(rng.integers(1, 6)), K)
        ng = int(rng.integers(1, 4)); kcheck = int(rng.integers(0, 4))
        best = NONE if i % 4 == 0 else int(rng.integers(0, stride * 64 + 1))
        gamma, check = rand_plan(rng, K, stride, ng, kcheck)
        n = stride * 64
        a = O.enumerate_plan(gamma, n, d, check=check, current_best=best)
        b = O.enumerate_brute(gamma, n, d, check=check, current_best=best)
        assert a == b, (K, stride, d, ng, kcheck, best)
        if i < 15:
            assert a == numpy_level(gamma, d, check, best, ng)


@pytest.mark.parametrize("threads", [1, 2, 3, 8])
def test_enumerator_thread_split(threads):
    rng = np.random.default_rng(threads)
    gamma, check = rand_plan(rng, 18, 2, 2, 2)
    a = O.enumerate_plan(gamma, 128, 5, check=check, current_best=NONE, threads=threads)
    b = O.enumerate_brute(gamma, 128, 5, check=check, current_best=NONE)
    assert a == b


def test_collect_vs_numpy():
    rng = np.random.default_rng(4)
    for K, stride, d, keep in [(8, 1, 2, 40), (10, 2, 3, 70), (7, 1, 4, 34)]:
        gamma, _ = rand_plan(rng, K, stride, 1, 0)
        got = sorted(tuple(int(x) for x in r) for r in O.collect_plan(gamma, stride * 64, d, keep))
        want = []
        for comb in itertools.combinations(range(K), d):
            cw = np.bitwise_xor.reduce(gamma[0, list(comb)], axis=0)
            w = sum(bin(int(x)).count("1") for x in cw)
            if 1 <= w <= keep:
                want.append(tuple(int(x) for x in cw))
        assert got == sorted(want)


def test_reference_enumerator_agrees_with_oracle():
    """Where the reference library is built, its own CPU enumerator on single levels."""
    if not O.ref_available():
        pytest.skip("reference library not built here")
    rng = np.random.default_rng(8)
    for i in range(20):
        K = int(rng.integers(6, 20)); stride = int(rng.integers(1, 4)); d = min(int(rng.integers(1, 6)), K)
        gamma, check = rand_plan(rng, K, stride, int(rng.integers(1, 3)), int(rng.integers(0, 3)))
        best = NONE if i % 3 == 0 else int(rng.integers(10, 60))
        assert O.ref_enumerate_plan(gamma, stride * 64, d, check=check, current_best=best) == \
            O.enumerate_plan(gamma, stride * 64, d, check=check, current_best=best)


# ---------------------------------------------------------------- random information sets

def _gf2_inv(A):
    """Inverse over GF(2) of a square 0/1 matrix, by row reduction of [A | I]."""
    k = A.shape[0]
    M = np.concatenate([A.astype(np.uint8) & 1, np.eye(k, dtype=np.uint8)], axis=1)
    for c in range(k):
        r = c + int(np.flatnonzero(M[c:, c])[0])
        M[[c, r]] = M[[r, c]]
        for rr in np.flatnonzero(M[:, c]):
            if rr != c:
                M[rr] ^= M[c]
    return M[:, k:]


def _nullspace(H):
    """Basis of {x : H x = 0} over GF(2), by row reduction of H."""
    H = (np.asarray(H, np.uint8) & 1).copy()
    r, n = H.shape
    piv, row = [], 0
    for c in range(n):
        nz = [i for i in range(row, r) if H[i, c]]
        if not nz:
            continue
        H[[row, nz[0]]] = H[[nz[0], row]]
        for i in range(r):
            if i != row and H[i, c]:
                H[i] ^= H[row]
        piv.append(c)
        row += 1
    free = [c for c in range(n) if c not in piv]
    B = np.zeros((len(free), n), np.uint8)
    for j, f in enumerate(free):
        B[j, f] = 1
        for i, p in enumerate(piv):
            B[j, p] = H[i, f]
    return B


def _golay23_generator():
    """Cyclic [23,12,7] Golay code from g(x) = 1 + x^2 + x^4 + x^5 + x^6 + x^10 + x^11."""
    g = np.zeros(23, np.uint8)
    g[[0, 2, 4, 5, 6, 10, 11]] = 1
    return np.stack([np.roll(g, i) for i in range(12)])


def _random_basis(rng, K, n):
    while True:
        G = rng.integers(0, 2, size=(K, n), dtype=np.uint8)
        if codes.gf2_rank(G) == K:
            return G


@pytest.mark.parametrize("K,n", [(1, 5), (6, 13), (20, 70), (33, 130)])
def test_is_trial_is_the_unique_rref(K, n):
    """Each trial's systematic form is THE reduced row echelon form of the row space over
    the trial's column order, computed here another way: the pivots are the columns that
    raise the column rank in that order, and R = G[:, pivots]^-1 G."""
    from qubitserf_amd.engine import pack_rows, unpack_rows
    rng = np.random.default_rng(K * 1000 + n)
    G = _random_basis(rng, K, n)
    for t in (0, 1, 17):
        rows, perm = O.is_trial(pack_rows(G), n, t)
        assert sorted(perm.tolist()) == list(range(n))
        piv, rank = [], 0
        for c in perm:
            cand = piv + [int(c)]
            if codes.gf2_rank(G[:, cand]) > rank:
                piv, rank = cand, rank + 1
            if rank == K:
                break
        R = (_gf2_inv(G[:, piv]).astype(np.int64) @ G.astype(np.int64)) % 2
        assert np.array_equal(unpack_rows(rows, n), R.astype(np.uint8))


def test_is_trial_orders_differ_and_depend_on_seed():
    from qubitserf_amd.engine import pack_rows
    G = pack_rows(_random_basis(np.random.default_rng(5), 8, 40))
    p = [O.is_trial(G, 40, t)[1].tolist() for t in range(4)]
    assert len({tuple(x) for x in p}) == 4
    assert O.is_trial(G, 40, 0, seed=1)[1].tolist() != p[0]


def test_is_bound_is_min_over_trials_and_bounds_distance():
    from qubitserf_amd.engine import pack_rows, unpack_rows
    rng = np.random.default_rng(11)
    for trial in range(6):
        K, n = int(rng.integers(3, 12)), int(rng.integers(14, 40))
        G = _random_basis(rng, K, n)
        H = _nullspace(G)
        d = O.bruteforce_distance(0, H)
        Gp = pack_rows(G)
        want = min(int(unpack_rows(O.is_trial(Gp, n, t)[0], n).sum(axis=1).min()) for t in range(9))
        assert O.is_bound_trials(Gp, n, trials=9) == want
        assert O.is_bound_trials(Gp, n, trials=256) >= d
        assert O.is_bound_trials(Gp, n, trials=9, init=1) == 1
    # trials split across calls fold to the same bound
    assert min(O.is_bound_trials(Gp, n, trials=5), O.is_bound_trials(Gp, n, trials=4, first=5)) == \
        O.is_bound_trials(Gp, n, trials=9)


def test_is_bound_exact_on_small_codes():
    """Hamming [7,4,3] and the [23,12,7] Golay code: 256 systematic forms reach d."""
    from qubitserf_amd.engine import pack_rows
    for G, d in ((_nullspace(codes.hamming_parity(3)), 3), (_golay23_generator(), 7)):
        assert codes.gf2_rank(G) == G.shape[0]
        assert O.is_bound_trials(pack_rows(G), G.shape[1], trials=256) == d


def test_is_bound_logical_test():
    """With detector rows only odd-overlap rows count: the 6x6 toric code's searched space
    holds its weight-4 plaquette stabilizers, yet the bound with detectors is >= d = 6."""
    gamma, ranks, check, even = O.side_plan(1, *codes.toric(6))
    G = gamma[0]
    n = 2 * 36
    assert check.shape[0] > 0
    assert O.is_bound_trials(G, n, check=check, trials=256) >= 6
    assert O.is_bound_trials(G, n, trials=256) < 6


@pytest.mark.parametrize("name,kind,mats", [("toric5", 1, codes.toric(5)), ("bb72", 2, codes.bb_family(6)),
                                            ("hamming4", 0, (codes.hamming_parity(4), None))])
def test_greedy_sets_are_information_sets(name, kind, mats):
    """Every set of the peel is an information set: its generator spans the code, is the
    identity on the set's pivot columns (the columns left over that raise the rank, in
    order), and the sets are disjoint and cover the nonzero columns."""
    from qubitserf_amd.engine import unpack_rows
    gen, _ = O.problem_gen(kind, *mats)
    n = mats[0].shape[1]
    G = unpack_rows(gen, n)
    K = G.shape[0]
    sets, ranks = O.greedy_sets(gen, n)
    seen = np.zeros(n, bool)
    nonzero = G.any(axis=0)
    for S, r in zip(sets, ranks):
        M = unpack_rows(S, n)
        assert codes.gf2_rank(np.vstack([G, M])) == K == codes.gf2_rank(M)
        # pivots: the columns left over, in increasing order, that raise the column rank
        piv, rank = [], 0
        for c in np.flatnonzero(~seen & nonzero):
            if codes.gf2_rank(G[:, piv + [int(c)]]) > rank:
                piv, rank = piv + [int(c)], rank + 1
        assert rank == r
        for i, c in enumerate(piv):
            col = M[:, c]
            assert col[i] == 1 and col.sum() == 1
            seen[c] = True
    assert np.array_equal(seen, nonzero)
