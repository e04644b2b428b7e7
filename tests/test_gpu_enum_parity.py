"""Device enumerator vs the oracle on randomized EnumPlans (the design of the reference's
cpp/distfind/tests/cpu_enum_ref.cpp:27-127, with the device as the system under test and
the oracle's CPU enumerator as the checker). Integer results: exact equality."""
import numpy as np
import pytest

from tests import oracle as O

pytestmark = pytest.mark.gpu
NONE = 1 << 30


def rand_plan(rng, K, stride, ng, kcheck, systematic=False):
    gamma = rng.integers(0, 2**64, size=(ng, K, stride), dtype=np.uint64)
    if systematic:   # identity on K columns spread over the row, like real information sets
        for g in range(ng):
            cols = rng.choice(stride * 64, size=K, replace=False)
            for r, c in enumerate(cols):
                w, b = divmod(int(c), 64)
                gamma[g, :, w] &= ~np.uint64(1 << b)
                gamma[g, r, w] |= np.uint64(1 << b)
    check = rng.integers(0, 2**64, size=(kcheck, stride), dtype=np.uint64) if kcheck else None
    return gamma, check


def run_case(hip, rng, K, stride, d, ng, kcheck, best, systematic=False, num_gamma=None):
    n = stride * 64
    gamma, check = rand_plan(rng, K, stride, ng, kcheck, systematic)
    got = hip.enumerate_level(gamma, n, d, check=check, current_best=best, num_gamma=num_gamma)
    kern = hip.last_stats()["kernel"]
    want = O.enumerate_plan(gamma, n, d, check=check, current_best=best, num_gamma=num_gamma)
    assert got == want, f"K={K} stride={stride} d={d} ng={ng} kcheck={kcheck} best={best}: gpu={got} oracle={want} kernel={kern}"
    return kern


@pytest.mark.parametrize("policy", ["auto", "sweep", "tiled"])
def test_small_random_plans(hip, policy):
    rng = np.random.default_rng(0xC0FFEE)
    with hip.kernel_policy(policy):
        _small_random_plans(hip, rng)


def _small_random_plans(hip, rng):
    for _ in range(60):
        K = int(rng.integers(4, 15)); stride = int(rng.integers(1, 4)); d = int(rng.integers(1, 6))
        d = min(d, K)
        ng = int(rng.integers(1, 4)); kcheck = int(rng.integers(0, 4))
        roll = int(rng.integers(0, 4))
        best = NONE if roll == 0 else int(rng.integers(0, stride * 64 + 1))
        run_case(hip, rng, K, stride, d, ng, kcheck, best)


@pytest.mark.parametrize("policy", ["auto", "sweep", "tiled"])
def test_edge_plans(hip, policy):
    with hip.kernel_policy(policy):
        _edge_plans(hip)


def _edge_plans(hip):
    rng = np.random.default_rng(1)
    run_case(hip, rng, 8, 1, 1, 1, 0, NONE)     # d == 1
    run_case(hip, rng, 6, 2, 1, 2, 2, NONE)
    run_case(hip, rng, 5, 1, 5, 1, 1, NONE)     # d == K: one combination
    run_case(hip, rng, 7, 2, 4, 2, 0, NONE)     # every codeword logical
    run_case(hip, rng, 10, 3, 3, 1, 2, 0)       # current_best 0: nothing lighter
    run_case(hip, rng, 12, 1, 4, 3, 1, 1)       # very tight cap


@pytest.mark.parametrize("seed", range(12))
def test_tiled_kernel_plans(hip, seed):
    """Levels big enough for the tiled kernel, random and systematic generators."""
    rng = np.random.default_rng(100 + seed)
    K = int(rng.integers(24, 60)); stride = int(rng.integers(1, 4)); d = int(rng.integers(3, 6))
    ng = int(rng.integers(1, 3)); kcheck = int(rng.integers(0, 4))
    best = [NONE, 40, 25, int(stride * 64 // 3)][seed % 4]
    with hip.kernel_policy("tiled"):
        kern = run_case(hip, rng, K, stride, d, ng, kcheck, best, systematic=bool(seed % 2))
    assert kern == 2, "expected the tiled kernel on this level"


def test_tiled_masked_active_prefix(hip):
    """num_gamma < num_gamma_total: only the active prefix is enumerated."""
    rng = np.random.default_rng(5)
    run_case(hip, rng, 40, 2, 4, 3, 2, 60, systematic=True, num_gamma=1)
    run_case(hip, rng, 40, 2, 4, 3, 2, 60, systematic=True, num_gamma=2)


def test_key_reuse_across_levels(hip):
    """One buffers_key per solve: device tables cached across levels stay correct, and a
    new solve with a new key (same host buffer) is never served stale data."""
    rng = np.random.default_rng(9)
    n = 128
    for trial in range(3):
        gamma, check = rand_plan(rng, 36, 2, 2, 2, systematic=True)
        key = 0xABC000 + trial
        for d in (1, 2, 3, 4):
            got = hip.enumerate_level(gamma, n, d, check=check, current_best=NONE, key=key)
            want = O.enumerate_plan(gamma, n, d, check=check, current_best=NONE)
            assert got == want, (trial, d, got, want)


def _sorted_rows(a, stride):
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, stride)
    if a.shape[0] == 0:
        return a
    return a[np.lexsort(a.T[::-1])]


def _check_collect(hip, gamma, n, d, keep):
    got = hip.collect_level(gamma, n, d, keep)
    want = O.collect_plan(gamma, n, d, keep)
    stride = gamma.shape[2]
    g, w = _sorted_rows(got, stride), _sorted_rows(want, stride)
    assert g.shape == w.shape and np.array_equal(g, w), (gamma.shape, d, keep, g.shape, w.shape)
    return g.shape[0]


@pytest.mark.parametrize("policy", ["sweep", "tiled"])
def test_collect_matches_oracle(hip, policy):
    """Collect mode: the multiset of emitted codewords equals the oracle's (each
    combination exactly once, duplicates across generators kept)."""
    rng = np.random.default_rng(11)
    with hip.kernel_policy(policy):
        for K, stride, d, keep, ng in [(12, 1, 3, 30, 1), (20, 2, 3, 50, 2), (30, 1, 2, 28, 1), (16, 3, 4, 80, 1),
                                       (40, 2, 4, 50, 2)]:
            gamma, _ = rand_plan(rng, K, stride, ng, 0, systematic=True)
            _check_collect(hip, gamma, stride * 64, d, keep)


@pytest.mark.parametrize("K,d,stride,keep", [(64, 6, 2, 26), (90, 5, 3, 41), (120, 5, 3, 27)])
def test_tiled_split_levels(hip, K, d, stride, keep):
    """Levels whose skewed products are re-split with fixed rows: min mode against the
    oracle, and collect mode as an exact-cover check (every light codeword of the level
    appears exactly as often as the oracle emits it)."""
    rng = np.random.default_rng(K * 100 + d)
    gamma, check = rand_plan(rng, K, stride, 1, 2, systematic=True)
    n = stride * 64
    best = keep + 4
    got = hip.enumerate_level(gamma, n, d, check=check, current_best=best)
    assert hip.last_stats()["kernel"] == 2
    want = O.enumerate_plan(gamma, n, d, check=check, current_best=best)
    assert got == want, (K, d, got, want)
    rows = _check_collect(hip, gamma, n, d, keep)
    assert rows > 0


@pytest.mark.parametrize("fd", [1, 2, 3, 4])
def test_tiled_fast_width(hip, fd, monkeypatch):
    """Every fast-test width of the tiled kernel (dwords tested before the exact slow path)."""
    monkeypatch.setenv("QSA_TILED_FD", str(fd))
    rng = np.random.default_rng(500 + fd)
    with hip.kernel_policy("tiled"):
        for i in range(6):
            # stride 3 and K <= 60 leave >= 132 residual bits, so every width is available
            K = int(rng.integers(30, 61)); stride = 3; d = int(rng.integers(3, 6))
            kcheck = int(rng.integers(0, 3))
            best = [NONE, 60, 35, 20, 12, 90][i]
            run_case(hip, rng, K, stride, d, 1 + i % 2, kcheck, best, systematic=True)
            assert hip.last_stats()["fd"] == fd


def test_tiled_large_levels_with_groups(hip):
    """Levels with many tiles, so several stream chunks share one register tile."""
    rng = np.random.default_rng(77)
    for K, d, best in [(64, 5, 30), (80, 5, 40), (100, 4, 50), (128, 4, 60)]:
        stride = 3
        gamma, check = rand_plan(rng, K, stride, 1, 2, systematic=True)
        got = hip.enumerate_level(gamma, stride * 64, d, check=check, current_best=best)
        st = hip.last_stats()
        want = O.enumerate_plan(gamma, stride * 64, d, check=check, current_best=best)
        assert got == want, (K, d, got, want, st)
        assert st["kernel"] == 2


@pytest.mark.parametrize("K,stride,d,kern", [(128, 3, 3, 2), (129, 4, 3, 2), (256, 5, 3, 2), (257, 5, 3, 1),
                                              (34, 1, 32, 2), (36, 1, 33, 1)])
def test_kernel_boundaries(hip, K, stride, d, kern):
    """Either side of the kernels' limits, tiled policy, against the oracle: K = 128 / 129
    (one- and two-word list masks), K = 256 / 257 (tiled kernel / row sweep), d = 32 / 33
    (tiled kernel / row sweep)."""
    rng = np.random.default_rng(K * 7 + d)
    with hip.kernel_policy("tiled"):
        for best in (NONE, 30, 8):
            run_case(hip, rng, K, stride, d, 2, 1, best, systematic=True)
            assert hip.last_stats()["kernel"] == kern, (K, d, hip.last_stats())


def test_small_level_sweep(hip):
    """Levels under 2^16 codewords in the automatic policy run the small-level sweep (rows
    and binomials staged in LDS when they fit, read from global memory when not: K x stride
    > 2048 words, K x d > 1024): results equal the oracle's, loose and tight bounds."""
    rng = np.random.default_rng(2100)
    cases = [(32, 1, 3, 2, 0), (12, 2, 6, 3, 2), (64, 2, 2, 1, 1), (200, 4, 2, 2, 0), (300, 16, 1, 1, 3),
             (40, 16, 2, 2, 2), (20, 1, 19, 1, 0), (66, 3, 64, 1, 1), (150, 3, 2, 3, 2)]
    for i, (K, stride, d, ng, kcheck) in enumerate(cases):
        for best in (NONE, stride * 64 // 3, 3):
            run_case(hip, rng, K, stride, d, ng, kcheck, best, systematic=bool(i % 2))
            assert hip.last_stats()["kernel"] == 3, (K, d)


@pytest.mark.parametrize("S", [64, 128, 256, 512])
def test_tiled_stream_lengths(hip, S, monkeypatch):
    """Every stream chunk length the launch sizing can pick (the automatic choice halves it
    from 512 for launches with few tiles), forced, against the oracle: one- and two-word
    masks, random and systematic generators, loose and tight bounds."""
    monkeypatch.setenv("QSA_TILED_S", str(S))
    rng = np.random.default_rng(2300 + S)
    with hip.kernel_policy("tiled"):
        for i, (K, stride, d) in enumerate([(60, 2, 4), (96, 3, 4), (140, 4, 3), (40, 1, 5)]):
            best = [NONE, 25, 40, 12][i]
            run_case(hip, rng, K, stride, d, 1 + i % 2, i % 3, best, systematic=i % 2 == 0)
            assert hip.last_stats()["kernel"] == 2


@pytest.mark.parametrize("fd", [1, 2, 3, 4])
def test_tiled_wide_rows(hip, fd, monkeypatch):
    """K > 128: the VALU kernel with two-word list masks (each half of the rows up to 128
    rows), every fast-test width, random and systematic generators, against the oracle."""
    monkeypatch.setenv("QSA_TILED_FD", str(fd))
    rng = np.random.default_rng(1700 + fd)
    cases = [(129, 5, 3), (150, 5, 3), (176, 5, 3), (200, 6, 3), (256, 6, 2), (140, 5, 4)]
    with hip.kernel_policy("tiled"):
        for i, (K, stride, d) in enumerate(cases):
            best = [NONE, 60, 40, 90, NONE, 30][i]
            run_case(hip, rng, K, stride, d, 1 + i % 2, i % 3, best, systematic=i % 3 != 1)
            st = hip.last_stats()
            assert st["kernel"] == 2 and st["fd"] == fd, st


def test_tiled_wide_collect(hip):
    """K > 128 in collect mode: the multiset of emitted codewords equals the oracle's, an
    exact-cover check of the products over rows 0..255 (fixed row sets included)."""
    rng = np.random.default_rng(1800)
    with hip.kernel_policy("tiled"):
        for K, stride, d in [(150, 4, 3), (200, 5, 2), (240, 5, 3)]:
            gamma, _ = rand_plan(rng, K, stride, 1, 0, systematic=True)
            n = stride * 64
            low = O.enumerate_plan(gamma, n, d)
            rows = _check_collect(hip, gamma, n, d, low + 8)
            assert rows > 0
            assert hip.last_stats()["kernel"] == 2


@pytest.mark.parametrize("xd", [1, 2, 4])
def test_mfma_hot_test(hip, xd, monkeypatch):
    """The int8-MFMA hot test (bz_mfma) with each medium-test width, on random and
    systematic generators, loose and tight bounds, against the oracle."""
    monkeypatch.setenv("QSA_TILED_VARIANT", f"0,{xd},4")
    rng = np.random.default_rng(900 + xd)
    with hip.kernel_policy("tiled"):
        for i in range(10):
            K = int(rng.integers(24, 121)); stride = 3 if xd > 1 else int(rng.integers(1, 4))
            d = int(rng.integers(3, 6)) if K < 90 else 4
            kcheck = int(rng.integers(0, 4))
            best = [100, 70, 45, 30, 22, 14, 9, 60, 35, 25][i]
            run_case(hip, rng, K, stride, d, 1 + i % 2, kcheck, best, systematic=bool(i % 3))
            assert hip.last_stats()["fd"] == 0
        # above 100 (products are compared with min(bound, 127)) the VALU kernel takes over
        run_case(hip, rng, 60, 3, 4, 1, 1, NONE, systematic=True)
        assert hip.last_stats()["fd"] > 0


def test_mfma_chosen_on_large_levels(hip, monkeypatch):
    """Automatic choice: a level with a tight bound runs a 30-bit matrix test (int8, or the
    packed fp4 test); results equal the oracle's. (The matrix tests are taken only on levels of >= 2^33 codewords; the size
    rule is lowered here so that the oracle can check the level.)"""
    monkeypatch.setenv("QSA_MATRIX_MIN_WORK", "0")
    rng = np.random.default_rng(3)
    gamma, check = rand_plan(rng, 80, 3, 1, 2, systematic=True)
    n = 192
    want = O.enumerate_plan(gamma, n, 5, check=check, current_best=7)
    got = hip.enumerate_level(gamma, n, 5, check=check, current_best=7)
    st = hip.last_stats()
    assert st["fd"] == 0 and st["mfma"] in (1, 5) and got == want, st
    monkeypatch.delenv("QSA_MATRIX_MIN_WORK")
    got = hip.enumerate_level(gamma, n, 5, check=check, current_best=7)
    assert hip.last_stats()["mfma"] == 0 and got == want   # C(80, 5) is below the size rule



@pytest.mark.parametrize("mm,xd,stride", [(2, 2, 2), (2, 4, 3), (3, 4, 3), (4, 4, 3), (5, 4, 2), (5, 4, 3)])
def test_f4_hot_test(hip, mm, xd, stride, monkeypatch):
    """The fp4 matrix-core hot tests (bz_f4: 32x32x64 products with 60 test bits, 16x16x128
    with 124, two chained 32x32x64 with 124, packed 32x32x64 with 30 bits on two lane
    elements per row, bounds <= 30) on random and systematic generators, loose and tight
    bounds, against the oracle. Systematic generators flag every row (the lane and stream
    constants are the subset sizes); random ones flag few (the constants vary per
    element)."""
    rpl = 2 if mm == 3 else 4
    monkeypatch.setenv("QSA_TILED_VARIANT", f"0,{xd},{rpl},{mm}")
    rng = np.random.default_rng(1300 + 10 * mm + xd)
    kmax = 64 if stride == 2 else 120
    with hip.kernel_policy("tiled"):
        for i in range(12):
            K = int(rng.integers(24, kmax + 1))
            d = int(rng.integers(3, 6)) if K < 90 else 4
            kcheck = int(rng.integers(0, 4))
            best = [100, 70, 45, 30, 22, 14, 9, 60, 35, 25, 18, 5][i]
            if mm == 5:   # the packed test runs at bounds <= 30
                best = [30, 29, 27, 24, 22, 14, 9, 30, 26, 20, 12, 5][i]
            run_case(hip, rng, K, stride, d, 1 + i % 2, kcheck, best, systematic=bool(i % 3))
            st = hip.last_stats()
            assert st["fd"] == 0 and st["mfma"] == mm, st
        # above 100 the VALU kernel takes over
        run_case(hip, rng, 60, 3, 4, 1, 1, NONE, systematic=True)
        assert hip.last_stats()["mfma"] == 0


def test_f4_chosen_on_high_thresholds(hip, monkeypatch):
    """Automatic choice: levels whose bound leaves too many light codewords on 30 test bits
    run the fp4 32x32x64 test (60 bits), looser ones the two chained products (124 bits);
    each level runs the cost model's pick and its result equals the oracle's. The size rule
    (>= 2^33 codewords) is lowered so that the oracle can check the levels."""
    monkeypatch.setenv("QSA_MATRIX_MIN_WORK", "0")
    rng = np.random.default_rng(31)
    seen = set()
    for K, d, best in [(80, 5, 14), (72, 4, 16), (64, 4, 30), (90, 4, 40), (80, 4, 34), (72, 5, 36)]:
        gamma, check = rand_plan(rng, K, 3, 1, 2, systematic=True)
        n = 192
        want = O.enumerate_plan(gamma, n, d, check=check, current_best=best)
        pick = hip.plan_variant(gamma, n, d, check=check, current_best=best)
        got = hip.enumerate_level(gamma, n, d, check=check, current_best=best)
        st = hip.last_stats()
        assert got == want, (K, d, best, got, want, st)
        assert (st["fd"], st["xd"], st["rpl"], st["mfma"]) == pick, (K, d, best, pick, st)
        seen.add(st["mfma"])
    assert 4 in seen and seen & {2, 5}, seen


def test_automatic_choice_runs_the_cost_models_pick(hip, monkeypatch):
    """The engine runs, on a fresh solve, exactly the variant its measured cost model picks
    (qsf_hip_plan_variant, host only), across tight and loose bounds, with results equal to
    the oracle's. The size rule is lowered so that the oracle can check the levels; a bound
    above 100 leaves the VALU kernel only."""
    monkeypatch.setenv("QSA_MATRIX_MIN_WORK", "0")
    rng = np.random.default_rng(41)
    kinds = set()
    for K, d, best in [(80, 5, 7), (80, 5, 13), (72, 4, 16), (64, 4, 30), (96, 4, 9), (60, 5, 22), (70, 4, 120)]:
        gamma, check = rand_plan(rng, K, 3, 1, 2, systematic=True)
        n = 192
        pick = hip.plan_variant(gamma, n, d, check=check, current_best=best)
        got = hip.enumerate_level(gamma, n, d, check=check, current_best=best)
        st = hip.last_stats()
        ran = (st["fd"], st["xd"], st["rpl"], st["mfma"])
        assert ran == pick, (K, d, best, ran, pick)
        assert got == O.enumerate_plan(gamma, n, d, check=check, current_best=best), (K, d, best)
        kinds.add(pick[3])
    assert 0 in kinds and len(kinds) >= 3, kinds


@pytest.mark.parametrize("thin", ["0", "64", "256", "4096", "100000000"])
def test_matrix_levels_with_thin_products_on_the_valu_kernel(hip, thin, monkeypatch):
    """A matrix-test level's products with a list shorter than QSA_THIN_BELOW run on the
    VALU kernel in a second launch with the same job and bound (0: none; 1e8: every
    product). Results equal the oracle's for the int8 and fp4 tests, with tight and loose
    bounds; the run reports the matrix variant while its launch has products left."""
    monkeypatch.setenv("QSA_MATRIX_MIN_WORK", "0")
    monkeypatch.setenv("QSA_THIN_BELOW", thin)
    rng = np.random.default_rng(77)
    seen = set()
    for K, d, best in [(80, 5, 7), (80, 5, 13), (72, 4, 16), (96, 4, 9), (60, 5, 9), (110, 4, 12)]:
        gamma, check = rand_plan(rng, K, 3, 1 + K % 2, 2, systematic=K % 3 != 1)
        n = 192
        want = O.enumerate_plan(gamma, n, d, check=check, current_best=best)
        got = hip.enumerate_level(gamma, n, d, check=check, current_best=best)
        st = hip.last_stats()
        assert got == want, (K, d, best, thin, got, want, st)
        seen.add(st["mfma"])
    if int(thin) <= 256:
        assert seen & {1, 2}, seen


def sparse_plan(rng, K, n, ng, p, kcheck):
    """Systematic generators with sparse redundancy (density p), so that light codewords
    exist and the matrix tests are the cost model's pick at a bound just above them."""
    stride = (n + 63) // 64
    gamma = np.zeros((ng, K, stride), np.uint64)
    for g in range(ng):
        cols = rng.permutation(n)
        M = np.zeros((K, n), np.uint8)
        M[np.arange(K), cols[:K]] = 1
        M[:, cols[K:]] = rng.random((K, n - K)) < p
        for w in range(stride):
            blk = M[:, 64 * w:64 * w + 64].astype(np.uint64)
            gamma[g, :, w] = (blk << np.arange(blk.shape[1], dtype=np.uint64)).sum(1)
    check = rng.integers(0, 2**64, size=(kcheck, stride), dtype=np.uint64) if kcheck else None
    return gamma, check


@pytest.mark.parametrize("thin", ["256", "100000000"])
def test_fixed_set_lists_on_expanded_levels(hip, thin, monkeypatch):
    """Levels large enough that a short list (fewer than 64 subsets) is expanded into fixed
    row sets, on a matrix test (sparse generators: the cost model picks the int8 or fp4
    test at a bound just above the level's minimum): the children with a short list run
    as fixed-set-list rectangles on the VALU kernel (every (fixed set, subset) pair on the
    stream). With the bound one above the minimum, the minimum itself must be found: a
    missed or doubly counted pair, a wrong fixed set or its syndrome, shows as a wrong
    result."""
    monkeypatch.setenv("QSA_MATRIX_MIN_WORK", "0")
    monkeypatch.setenv("QSA_THIN_BELOW", thin)
    rng = np.random.default_rng(4242)
    n = 192
    seen_side = seen_matrix = False
    for K, d, ng, p, kc in [(80, 6, 1, 0.03, 0), (96, 5, 1, 0.03, 2), (72, 6, 2, 0.04, 0), (88, 5, 2, 0.03, 1),
                            (100, 5, 1, 0.03, 0)]:
        gamma, check = sparse_plan(rng, K, n, ng, p, kc)
        low = O.enumerate_plan(gamma, n, d, check=check)
        for best in (low + 1, low + 4, NONE):
            want = O.enumerate_plan(gamma, n, d, check=check, current_best=best)
            pick = hip.plan_variant(gamma, n, d, check=check, current_best=best)
            got = hip.enumerate_level(gamma, n, d, check=check, current_best=best)
            st = hip.last_stats()
            assert got == want, (K, d, ng, best, thin, got, want, st)
            assert st["side_codewords"] <= st["codewords"], st
            if pick[3] > 0:
                seen_matrix = True
                seen_side = seen_side or st["side_codewords"] > 0
    assert seen_matrix
    assert seen_side, "no matrix-test level ran products on the VALU kernel"
