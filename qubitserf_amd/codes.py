"""Code constructions used by the tests and by the benchmark workloads.

Every builder returns plain ``numpy.uint8`` 0/1 matrices: a parity-check ``H`` for a
classical code, ``(Hx, Hz)`` for a CSS code. The families are the ones the reference
tests and benchmarks on (``python/qubitserf/_constructions.py``, ``distfind/codes.py``):
repetition, Hamming, Steane, Shor, toric and surface codes as hypergraph products,
bivariate-bicycle codes (the gross code among them), quantum Reed–Muller codes, and the
seeded random / quasi-cyclic ensembles the benchmark configurations name.

Random ensembles use ``numpy.random.default_rng(seed)`` (PCG64) so every instance is
reproducible from its seed.
"""
from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------- GF(2) helpers


def gf2_rank(M: np.ndarray) -> int:
    """Rank over GF(2) of a 0/1 matrix (rows packed into Python ints)."""
    rows = [int("".join("1" if b else "0" for b in r[::-1]) or "0", 2) for r in np.asarray(M) & 1]
    rank = 0
    pivots: dict[int, int] = {}
    for v in rows:
        while v:
            h = v.bit_length() - 1
            if h in pivots:
                v ^= pivots[h]
            else:
                pivots[h] = v
                rank += 1
                break
    return rank


def gf2_rref(M: np.ndarray) -> np.ndarray:
    """Reduced row echelon form over GF(2) of the rows of M, zero rows dropped: a generator
    of rowspace(M) that is the identity on its pivot columns (a systematic generator)."""
    A = (np.asarray(M, np.uint8) & 1).copy()
    m, n = A.shape
    r = 0
    for c in range(n):
        if r == m:
            break
        nz = np.flatnonzero(A[r:, c])
        if nz.size == 0:
            continue
        k = r + int(nz[0])
        A[[r, k]] = A[[k, r]]
        hit = np.flatnonzero(A[:, c])
        hit = hit[hit != r]
        A[hit] ^= A[r]
        r += 1
    return A[:r].copy()


def gf2_kernel(H: np.ndarray) -> np.ndarray:
    """A basis (rows) of {x : H x = 0} over GF(2): reduced row echelon form of H, then one
    basis vector per free column (the free column set to 1, pivots solved)."""
    A = (np.asarray(H, np.uint8) & 1).copy()
    m, n = A.shape
    piv = []
    r = 0
    for c in range(n):
        if r == m:
            break
        nz = np.flatnonzero(A[r:, c])
        if nz.size == 0:
            continue
        k = r + int(nz[0])
        A[[r, k]] = A[[k, r]]
        hit = np.flatnonzero(A[:, c])
        hit = hit[hit != r]
        A[hit] ^= A[r]
        piv.append(c)
        r += 1
    free = np.setdiff1d(np.arange(n), piv)
    B = np.zeros((free.size, n), np.uint8)
    B[np.arange(free.size), free] = 1
    if piv:
        B[:, piv] = A[: len(piv)][:, free].T
    return B


# ---------------------------------------------------------------- classical families


def repetition_parity(n: int) -> np.ndarray:
    """(n-1) x n checks x_i + x_{i+1} of the [n, 1, n] repetition code."""
    H = np.zeros((n - 1, n), dtype=np.uint8)
    idx = np.arange(n - 1)
    H[idx, idx] = 1
    H[idx, idx + 1] = 1
    return H


def cyclic_repetition_parity(n: int) -> np.ndarray:
    """n x n closed-loop checks x_i + x_{i+1 mod n} (rank n-1)."""
    H = np.zeros((n, n), dtype=np.uint8)
    idx = np.arange(n)
    H[idx, idx] = 1
    H[idx, (idx + 1) % n] = 1
    return H


def hamming_parity(r: int) -> np.ndarray:
    """r x (2^r - 1) check of the Hamming code: column j is the binary expansion of j+1."""
    n = (1 << r) - 1
    j = np.arange(1, n + 1)
    return ((j[None, :] >> np.arange(r)[:, None]) & 1).astype(np.uint8)


def reed_muller_generator(r: int, m: int) -> np.ndarray:
    """Generator of RM(r, m): evaluations of all monomials of degree <= r on F_2^m."""
    pts = ((np.arange(1 << m)[None, :] >> np.arange(m)[:, None]) & 1).astype(np.uint8)
    rows = [np.ones(1 << m, dtype=np.uint8)]
    from itertools import combinations

    for deg in range(1, r + 1):
        for S in combinations(range(m), deg):
            rows.append(np.prod(pts[list(S)], axis=0).astype(np.uint8))
    return np.array(rows, dtype=np.uint8)


# ---------------------------------------------------------------- CSS families


def steane():
    H = hamming_parity(3)
    return H.copy(), H.copy()


def shor():
    """Shor [[9,1,3]]: Z-checks on neighbouring pairs inside each block of three, X-checks
    on pairs of blocks."""
    Hz = np.zeros((6, 9), dtype=np.uint8)
    for b in range(3):
        for t in range(2):
            Hz[2 * b + t, 3 * b + t] = 1
            Hz[2 * b + t, 3 * b + t + 1] = 1
    Hx = np.zeros((2, 9), dtype=np.uint8)
    Hx[0, 0:6] = 1
    Hx[1, 3:9] = 1
    return Hx, Hz


def hypergraph_product(H1: np.ndarray, H2: np.ndarray):
    """Hx = [H1 (x) I_n2 | I_m1 (x) H2^T],  Hz = [I_n1 (x) H2 | H1^T (x) I_m2]."""
    m1, n1 = H1.shape
    m2, n2 = H2.shape
    Hx = np.hstack([np.kron(H1, np.eye(n2, dtype=np.uint8)), np.kron(np.eye(m1, dtype=np.uint8), H2.T)]) % 2
    Hz = np.hstack([np.kron(np.eye(n1, dtype=np.uint8), H2), np.kron(H1.T, np.eye(m2, dtype=np.uint8))]) % 2
    return Hx.astype(np.uint8), Hz.astype(np.uint8)


def toric(L: int):
    """[[2L^2, 2, L]] toric code: hypergraph product of two length-L cyclic repetition codes."""
    H = cyclic_repetition_parity(L)
    return hypergraph_product(H, H)


def surface(L: int):
    """[[L^2 + (L-1)^2, 1, L]] planar surface code: product of two repetition codes."""
    H = repetition_parity(L)
    return hypergraph_product(H, H)


def quantum_reed_muller(r: int, m: int):
    """CSS code with Hx = Hz = generator of RM(r, m); for r = 1 it is [[2^m, 2^m-2(m+1), 4]]."""
    G = reed_muller_generator(r, m)
    return G.copy(), G.copy()


def _shift(power: int, ell: int) -> np.ndarray:
    S = np.zeros((ell, ell), dtype=np.uint8)
    i = np.arange(ell)
    S[i, (i + power) % ell] = 1
    return S


def bivariate_bicycle(ell: int, m: int, a_terms, b_terms):
    """Bivariate-bicycle code: x = S_ell (x) I_m, y = I_ell (x) S_m, A and B sums of
    monomials x^i y^j given as (i, j) pairs; Hx = [A | B], Hz = [B^T | A^T], n = 2 ell m."""

    def poly(terms):
        P = np.zeros((ell * m, ell * m), dtype=np.uint8)
        for i, j in terms:
            P ^= np.kron(_shift(i, ell), _shift(j, m))
        return P

    A, B = poly(a_terms), poly(b_terms)
    return np.hstack([A, B]).astype(np.uint8), np.hstack([B.T, A.T]).astype(np.uint8)


BB_A = [(3, 0), (0, 1), (0, 2)]   # A = x^3 + y + y^2
BB_B = [(0, 3), (1, 0), (2, 0)]   # B = y^3 + x + x^2


def gross():
    """[[144, 12, 12]] gross code: ell = 12, m = 6 with the polynomials above."""
    return bivariate_bicycle(12, 6, BB_A, BB_B)


def bb_family(ell: int, m: int = 6):
    """The bivariate-bicycle family sharing the gross code's polynomials."""
    return bivariate_bicycle(ell, m, BB_A, BB_B)


# ---------------------------------------------------------------- seeded ensembles


def random_full_rank(rng: np.random.Generator, m: int, n: int) -> np.ndarray:
    """m x n i.i.d. Bernoulli(1/2) matrix, redrawn until its GF(2) rank is m."""
    while True:
        H = rng.integers(0, 2, size=(m, n), dtype=np.uint8)
        if gf2_rank(H) == m:
            return H


def config1_parity(seed: int = 0) -> np.ndarray:
    """32 x 64 full-rank Bernoulli(1/2) parity check (a [64, 32] random code)."""
    return random_full_rank(np.random.default_rng(seed), 32, 64)


def random_rate_code(rng: np.random.Generator, n_lo: int, n_hi: int, r_lo=0.3, r_hi=0.6) -> np.ndarray:
    """n ~ U{n_lo..n_hi}, rate ~ U[r_lo, r_hi], k = round(rate n); H is (n-k) x n full rank."""
    n = int(rng.integers(n_lo, n_hi + 1))
    rate = float(rng.uniform(r_lo, r_hi))
    k = int(round(rate * n))
    k = min(max(k, 1), n - 1)
    return random_full_rank(rng, n - k, n)


def config3_batch(seed: int = 0, count: int = 256):
    """The batched workload: `count` random codes, n ~ U{64..160}, rate ~ U[0.3, 0.6]."""
    rng = np.random.default_rng(seed)
    return [random_rate_code(rng, 64, 160) for _ in range(count)]


def quasi_cyclic(rng: np.random.Generator, L: int, rows: int = 3, cols: int = 6, w: int = 3) -> np.ndarray:
    """rows x cols array of L x L circulants, each with a first row of weight w at
    distinct uniform positions; H is (rows L) x (cols L)."""
    H = np.zeros((rows * L, cols * L), dtype=np.uint8)
    for bi in range(rows):
        for bj in range(cols):
            first = np.zeros(L, dtype=np.uint8)
            first[rng.choice(L, size=w, replace=False)] = 1
            for s in range(L):
                H[bi * L + s, bj * L:(bj + 1) * L] = np.roll(first, s)
    return H


def config4_eval(seed: int = 0, count: int = 16, L_lo: int = 20, L_hi: int = 40):
    """Evaluation set of quasi-cyclic codes, circulant size L ~ U{L_lo..L_hi}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        L = int(rng.integers(L_lo, L_hi + 1))
        out.append(quasi_cyclic(rng, L))
    return out
