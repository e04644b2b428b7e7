// Row masks of list subsets: colex unrank and successor, flagged-row counts and bit ranges,
// for lists of up to 64 rows (one u64 word) and up to 128 rows (M2, two words). Shared by
// the enumeration kernels (device) and their host-side unit test, which compiles this
// header with a plain C++ compiler.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define QSA_HD __host__ __device__ __forceinline__
#else
#define QSA_HD inline
#endif

namespace qsa {

using u64 = std::uint64_t;

// Row masks of a list's subsets: u64 for lists of up to 64 rows, M2 for up to 128.
struct M2 {
    u64 lo, hi;
};
QSA_HD u64 mask_xor(u64 a, u64 b) { return a ^ b; }
QSA_HD M2 mask_xor(M2 a, M2 b) { return M2{a.lo ^ b.lo, a.hi ^ b.hi}; }
QSA_HD bool nonzero(u64 m) { return m != 0; }
QSA_HD bool nonzero(M2 m) { return (m.lo | m.hi) != 0; }
QSA_HD int pop_and(u64 m, u64 f) { return __builtin_popcountll(m & f); }
QSA_HD int pop_and(M2 m, M2 f) { return __builtin_popcountll(m.lo & f.lo) + __builtin_popcountll(m.hi & f.hi); }
QSA_HD void set_bit(u64& m, int b) { m |= 1ull << b; }
QSA_HD void set_bit(M2& m, int b) {
    if (b < 64) m.lo |= 1ull << b; else m.hi |= 1ull << (b - 64);
}
// Index of the lowest set bit, which is cleared (m != 0).
QSA_HD int take_low(u64& m) {
    const int b = __builtin_ctzll(m);
    m &= m - 1;
    return b;
}
QSA_HD int take_low(M2& m) {
    if (m.lo) {
        const int b = __builtin_ctzll(m.lo);
        m.lo &= m.lo - 1;
        return b;
    }
    const int b = 64 + __builtin_ctzll(m.hi);
    m.hi &= m.hi - 1;
    return b;
}

// Colex unrank: the idx-th k-subset of [0, m) in colex order (= increasing order of the
// subset's bitmask), as a bitmask. Element c_i (i = k..1, decreasing) is the largest c
// with C(c, i) <= remaining index (combinatorial number system).
template <class Mask>
QSA_HD Mask colex_unrank(u64 idx, int k, int m, const u64* B, int bk) {
    Mask mask{};
    int cap = m;
    for (int i = k; i >= 1; --i) {
        int lo = i - 1, hi = cap - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (B[mid * bk + i] <= idx) lo = mid; else hi = mid - 1;
        }
        set_bit(mask, lo);
        idx -= (lo >= i) ? B[lo * bk + i] : 0ull;
        cap = lo;
    }
    return mask;
}
QSA_HD u64 colex_mask(u64 idx, int k, int m, const u64* B, int bk) {
    return colex_unrank<u64>(idx, k, m, B, bk);
}

// The unrank's binomials for the common case, small enough for LDS: C(n, i) for n < 64
// and i <= 8 all fit 32 bits (the largest, C(63, 8) = 3 872 894 697), stored row-major with
// kSmallBinCols columns. The kernels copy them once per workgroup, so an unrank of a list of
// at most 64 rows and 8 elements costs LDS reads instead of a chain of dependent global
// loads; other lists use the global table.
constexpr int kSmallBinRows = 64, kSmallBinK = 8, kSmallBinCols = kSmallBinK + 1;
template <class Mask>
QSA_HD Mask colex_unrank_small(u64 idx, int k, int m, const std::uint32_t* S) {
    Mask mask{};
    int cap = m;
    for (int i = k; i >= 1; --i) {
        int lo = i - 1, hi = cap - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if ((u64)S[mid * kSmallBinCols + i] <= idx) lo = mid; else hi = mid - 1;
        }
        set_bit(mask, lo);
        idx -= (lo >= i) ? (u64)S[lo * kSmallBinCols + i] : 0ull;
        cap = lo;
    }
    return mask;
}
// The small table when the list fits it, else the global one.
template <class Mask>
QSA_HD Mask colex_unrank_any(u64 idx, int k, int m, const std::uint32_t* S, const u64* B, int bk) {
    return (m <= kSmallBinRows && k <= kSmallBinK) ? colex_unrank_small<Mask>(idx, k, m, S)
                                                   : colex_unrank<Mask>(idx, k, m, B, bk);
}

// The first k-subset in colex order: rows 0 .. k - 1.
template <class Mask>
QSA_HD Mask colex_first(int k) {
    Mask m{};
    for (int b = 0; b < k; ++b) set_bit(m, b);
    return m;
}

// Next larger integer with the same popcount (colex successor); x != 0.
QSA_HD u64 colex_next(u64 x) {
    u64 c = x & (~x + 1ull);
    u64 r = x + c;
    return (((r ^ x) >> 2) >> __builtin_ctzll(c)) | r;
}
// The same on 128 bits: r = x + lowest set bit; the bits that the carry cleared, less two,
// refill the bottom. The successor must exist (x is not the last subset of 128 rows).
QSA_HD M2 colex_next(M2 x) {
    M2 r;
    if (x.lo) {
        const u64 c = x.lo & (~x.lo + 1ull);
        r.lo = x.lo + c;
        r.hi = x.hi + (r.lo < x.lo ? 1ull : 0ull);
    } else {
        r.lo = 0;
        r.hi = x.hi + (x.hi & (~x.hi + 1ull));
    }
    const int ones = __builtin_popcountll(r.lo ^ x.lo) + __builtin_popcountll(r.hi ^ x.hi) - 2;
    r.lo |= ones >= 64 ? ~0ull : ((1ull << ones) - 1ull);
    if (ones > 64) r.hi |= (1ull << (ones - 64)) - 1ull;
    return r;
}

// 64 bits of the 256-bit mask f[0..3] from bit off on (bits past 255 read as 0).
QSA_HD u64 word_at(const u64* f, int off) {
    const int w = off >> 6, sh = off & 63;
    const u64 a = w < 4 ? f[w] : 0ull;
    if (sh == 0) return a;
    const u64 b = w + 1 < 4 ? f[w + 1] : 0ull;
    return (a >> sh) | (b << (64 - sh));
}
QSA_HD u64 low_bits(u64 v, int m) { return m >= 64 ? v : v & ((1ull << m) - 1ull); }
// Bits [base, base + m) of the 256-bit mask f, shifted down to bit 0 (m <= 64, or <= 128 for M2).
template <class Mask>
QSA_HD Mask range_mask(const u64* f, int base, int m);
template <>
QSA_HD u64 range_mask<u64>(const u64* f, int base, int m) {
    return low_bits(word_at(f, base), m);
}
template <>
QSA_HD M2 range_mask<M2>(const u64* f, int base, int m) {
    return M2{low_bits(word_at(f, base), m), m > 64 ? low_bits(word_at(f, base + 64), m - 64) : 0ull};
}

}  // namespace qsa
