// Host-side check of the row-mask helpers the enumeration kernels use (masks.hpp):
// colex unrank and successor on one- and two-word masks against brute force and against
// each other, the 32-bit small-table unrank against the global-table one, bit ranges of 256-bit masks against bit-by-bit extraction, and the set-bit
// walk. Built and run by tests/test_cpu_masks.py; prints the first mismatch and exits 1.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "masks.hpp"

using qsa::M2;
using qsa::u64;

namespace {

constexpr int kBk = 65;   // binomial columns, as the device table
std::vector<u64> binom_table(int nmax) {
    std::vector<u64> B((size_t)(nmax + 1) * kBk, 0);
    for (int i = 0; i <= nmax; ++i) {
        B[(size_t)i * kBk] = 1;
        for (int j = 1; j < kBk && j <= i; ++j) {
            const u64 a = B[(size_t)(i - 1) * kBk + j - 1], b = j <= i - 1 ? B[(size_t)(i - 1) * kBk + j] : 0;
            B[(size_t)i * kBk + j] = (a + b < a) ? ~0ull : a + b;   // saturating
        }
    }
    return B;
}

int fails = 0;
void fail(const char* what, int m, int k, u64 i) {
    if (fails++ < 5) std::printf("FAIL %s m=%d k=%d i=%llu\n", what, m, k, (unsigned long long)i);
}

bool less128(M2 a, M2 b) { return a.hi != b.hi ? a.hi < b.hi : a.lo < b.lo; }
int pop(M2 a) { return __builtin_popcountll(a.lo) + __builtin_popcountll(a.hi); }
bool eq(M2 a, M2 b) { return a.lo == b.lo && a.hi == b.hi; }

// Every k-subset of [0, m) in colex order: successor chain == unrank, strictly increasing,
// popcount k, no bit at or above m, C(m, k) elements.
void check_chain_m2(const std::vector<u64>& B, int m, int k) {
    const u64 total = B[(size_t)m * kBk + k];
    M2 x = qsa::colex_unrank<M2>(0, k, m, B.data(), kBk);
    for (u64 i = 0; i < total; ++i) {
        const M2 u = qsa::colex_unrank<M2>(i, k, m, B.data(), kBk);
        if (!eq(u, x)) { fail("m2 successor != unrank", m, k, i); return; }
        if (pop(x) != k) { fail("m2 popcount", m, k, i); return; }
        const M2 lim = m >= 128 ? M2{0, 0} : (m >= 64 ? M2{0, ~0ull << (m - 64)} : M2{~0ull << m, ~0ull});
        if ((x.lo & lim.lo) | (x.hi & lim.hi)) { fail("m2 bit past range", m, k, i); return; }
        if (i + 1 < total) {
            const M2 nx = qsa::colex_next(x);
            if (!less128(x, nx)) { fail("m2 not increasing", m, k, i); return; }
            x = nx;
        }
    }
}

void check_chain_u64(const std::vector<u64>& B, int m, int k) {
    const u64 total = B[(size_t)m * kBk + k];
    u64 x = qsa::colex_mask(0, k, m, B.data(), kBk);
    for (u64 i = 0; i < total; ++i) {
        if (qsa::colex_mask(i, k, m, B.data(), kBk) != x) { fail("u64 successor != unrank", m, k, i); return; }
        const M2 w = qsa::colex_unrank<M2>(i, k, m, B.data(), kBk);
        if (w.lo != x || w.hi != 0) { fail("u64 != m2 below 64 rows", m, k, i); return; }
        if (i + 1 < total) {
            const u64 nx = qsa::colex_next(x);
            const M2 nw = qsa::colex_next(M2{x, 0});
            if (nw.lo != nx || nw.hi != 0) { fail("u64 next != m2 next", m, k, i); return; }
            x = nx;
        }
    }
}

// Brute force: the i-th integer below 2^m with popcount k is unrank(i).
void check_brute(const std::vector<u64>& B, int m, int k) {
    u64 i = 0;
    for (u64 v = 0; v < (1ull << m); ++v) {
        if (__builtin_popcountll(v) != k) continue;
        const M2 u = qsa::colex_unrank<M2>(i, k, m, B.data(), kBk);
        if (u.lo != v || u.hi != 0) { fail("brute force", m, k, i); return; }
        ++i;
    }
    if (i != B[(size_t)m * kBk + k]) fail("brute force count", m, k, i);
}

// Subsets that cross the word boundary: rows {a .. a + k - 1} shifted across bit 64.
void check_boundary(const std::vector<u64>& B) {
    for (int m : {65, 70, 96, 127, 128})
        for (int k = 1; k <= 6; ++k) {
            const u64 total = B[(size_t)m * kBk + k];
            std::mt19937_64 rng(1000 * m + k);
            for (int t = 0; t < 20000; ++t) {
                const u64 i = rng() % (total - 1);
                const M2 a = qsa::colex_unrank<M2>(i, k, m, B.data(), kBk);
                const M2 b = qsa::colex_unrank<M2>(i + 1, k, m, B.data(), kBk);
                if (!eq(qsa::colex_next(a), b)) { fail("random successor", m, k, i); return; }
            }
            const M2 last = qsa::colex_unrank<M2>(total - 1, k, m, B.data(), kBk);
            M2 want{0, 0};
            for (int r = m - k; r < m; ++r) qsa::set_bit(want, r);
            if (!eq(last, want)) fail("last subset", m, k, total - 1);
        }
}

// The 32-bit LDS table: exact copies of the global binomials it holds, and the same
// unrank as the global table for every list it serves (lists of up to 64 rows, up to 8
// elements), at the first, last and random indices.
void check_small_table(const std::vector<u64>& B) {
    std::vector<std::uint32_t> S((size_t)qsa::kSmallBinRows * qsa::kSmallBinCols);
    for (int n = 0; n < qsa::kSmallBinRows; ++n)
        for (int i = 0; i < qsa::kSmallBinCols; ++i) {
            const u64 v = B[(size_t)n * kBk + i];
            if (v > 0xffffffffull) { fail("small table overflow", n, i, v); return; }
            S[(size_t)n * qsa::kSmallBinCols + i] = (std::uint32_t)v;
        }
    std::mt19937_64 rng(11);
    for (int m = 1; m <= 64; ++m)
        for (int k = 0; k <= qsa::kSmallBinK && k <= m; ++k) {
            const u64 total = B[(size_t)m * kBk + k];
            for (int t = 0; t < 400; ++t) {
                const u64 i = t == 0 ? 0 : t == 1 ? total - 1 : rng() % total;
                const u64 a = qsa::colex_unrank<u64>(i, k, m, B.data(), kBk);
                const u64 b = qsa::colex_unrank_small<u64>(i, k, m, S.data());
                const u64 c = qsa::colex_unrank_any<u64>(i, k, m, S.data(), B.data(), kBk);
                const M2 d = qsa::colex_unrank_small<M2>(i, k, m, S.data());
                if (a != b || a != c || d.lo != a || d.hi != 0) { fail("small-table unrank", m, k, i); return; }
            }
        }
    // the first subset of each size: rank 0 of the unrank
    for (int k = 0; k <= 64; ++k) {
        const u64 want = k == 64 ? ~0ull : ((1ull << k) - 1);
        if (qsa::colex_first<u64>(k) != want) fail("colex_first<u64>", 64, k, 0);
        if (k <= 8 && qsa::colex_unrank_small<u64>(0, k, 64, S.data()) != want) fail("first vs unrank", 64, k, 0);
    }
    for (int k = 0; k <= 128; ++k) {
        const M2 f = qsa::colex_first<M2>(k);
        M2 want{k >= 64 ? ~0ull : ((1ull << k) - 1), k >= 128 ? ~0ull : k > 64 ? ((1ull << (k - 64)) - 1) : 0ull};
        if (!eq(f, want)) fail("colex_first<M2>", 128, k, 0);
    }
    // outside the table: the global path
    for (int m : {65, 100, 128})
        for (int k : {2, 9}) {
            const u64 total = B[(size_t)m * kBk + k];
            const u64 i = total / 3;
            const M2 a = qsa::colex_unrank<M2>(i, k, m, B.data(), kBk);
            const M2 c = qsa::colex_unrank_any<M2>(i, k, m, S.data(), B.data(), kBk);
            if (!eq(a, c)) fail("unrank_any fallback", m, k, i);
        }
}

void check_ranges() {
    std::mt19937_64 rng(7);
    for (int t = 0; t < 200000; ++t) {
        u64 f[4] = {rng(), rng(), rng(), rng()};
        if (t % 3 == 0) f[rng() % 4] = 0;
        const int base = (int)(rng() % 256);
        const int m2 = 1 + (int)(rng() % 128), m1 = 1 + (int)(rng() % 64);
        const M2 r2 = qsa::range_mask<M2>(f, base, m2);
        const u64 r1 = qsa::range_mask<u64>(f, base, m1);
        for (int j = 0; j < 128; ++j) {
            const int src = base + j;
            const int want = (j < m2 && src < 256) ? (int)((f[src >> 6] >> (src & 63)) & 1) : 0;
            const int got = j < 64 ? (int)((r2.lo >> j) & 1) : (int)((r2.hi >> (j - 64)) & 1);
            if (got != want) { fail("range_mask<M2>", m2, base, (u64)j); return; }
            if (j < 64) {
                const int w1 = (j < m1 && src < 256) ? (int)((f[src >> 6] >> (src & 63)) & 1) : 0;
                if ((int)((r1 >> j) & 1) != w1) { fail("range_mask<u64>", m1, base, (u64)j); return; }
            }
        }
        const M2 a{rng(), rng()}, b{rng(), rng()};
        int want_pop = 0;
        for (int j = 0; j < 128; ++j)
            want_pop += j < 64 ? (int)(((a.lo & b.lo) >> j) & 1) : (int)(((a.hi & b.hi) >> (j - 64)) & 1);
        if (qsa::pop_and(a, b) != want_pop) { fail("pop_and", 0, 0, (u64)t); return; }
        M2 w = a;
        int prev = -1, n = 0;
        while (qsa::nonzero(w)) {
            const int bit = qsa::take_low(w);
            const bool set = bit < 64 ? (a.lo >> bit) & 1 : (a.hi >> (bit - 64)) & 1;
            if (bit <= prev || !set) { fail("take_low order", bit, prev, (u64)t); return; }
            prev = bit;
            ++n;
        }
        if (n != pop(a)) { fail("take_low count", n, pop(a), (u64)t); return; }
    }
}

}  // namespace

int main() {
    const std::vector<u64> B = binom_table(256);
    for (int m = 1; m <= 18; ++m)
        for (int k = 0; k <= m && k <= 6; ++k) check_brute(B, m, k);
    for (int m : {1, 5, 33, 64})
        for (int k = 0; k <= 4 && k <= m; ++k) check_chain_u64(B, m, k);
    for (int m : {65, 66, 80, 100, 128})
        for (int k = 0; k <= 3; ++k) check_chain_m2(B, m, k);
    check_chain_m2(B, 70, 4);
    check_chain_m2(B, 128, 4);
    check_boundary(B);
    check_small_table(B);
    check_ranges();
    if (fails) {
        std::printf("%d failures\n", fails);
        return 1;
    }
    std::printf("masks ok\n");
    return 0;
}
