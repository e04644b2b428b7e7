// Bit-packed GF(2) rows and dense matrices for the host side of the engine.
//
// Layout (shared with the plan handed to the device): a row of `cols` bits is
// `words(cols)` little-endian u64 words, bit j in word j>>6 at offset j&63
// (cpp/qsf/bits.hpp:1-3 uses the same layout, so plans cross the boundary unchanged).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

namespace qsa {

using u64 = std::uint64_t;
using u32 = std::uint32_t;
using u8 = std::uint8_t;

constexpr int kNoWeight = 1 << 30;   // "no codeword found" sentinel (cpp/qsf/backend.hpp:19)

inline int words(int bits) { return (bits + 63) >> 6; }
inline int popc(u64 x) { return __builtin_popcountll(x); }

struct BitMat {
    int rows = 0, cols = 0, ld = 0;   // ld = u64 words per row
    std::vector<u64> data;

    BitMat() = default;
    BitMat(int r, int c) : rows(r), cols(c), ld(words(c)), data((size_t)r * words(c), 0) {}

    u64* row(int i) { return data.data() + (size_t)i * ld; }
    const u64* row(int i) const { return data.data() + (size_t)i * ld; }
    bool bit(int i, int j) const { return (row(i)[j >> 6] >> (j & 63)) & 1; }
    void set(int i, int j) { row(i)[j >> 6] |= 1ull << (j & 63); }
    void xor_into(int dst, const u64* src) { u64* d = row(dst); for (int k = 0; k < ld; ++k) d[k] ^= src[k]; }
    void swap(int a, int b) {
        if (a == b) return;
        u64 *x = row(a), *y = row(b);
        for (int k = 0; k < ld; ++k) { u64 t = x[k]; x[k] = y[k]; y[k] = t; }
    }
    int weight(int i) const { int s = 0; const u64* r = row(i); for (int k = 0; k < ld; ++k) s += popc(r[k]); return s; }
    void append(const u64* src) { data.insert(data.end(), src, src + ld); ++rows; }

    static BitMat from_bytes(const u8* a, int r, int c) {
        BitMat m(r, c);
        for (int i = 0; i < r; ++i) {
            const u8* s = a + (size_t)i * c;
            u64* d = m.row(i);
            for (int j = 0; j < c; ++j) d[j >> 6] |= (u64)(s[j] & 1) << (j & 63);
        }
        return m;
    }
};

inline int parity_and(const u64* a, const u64* b, int ld) {
    u64 acc = 0;
    for (int k = 0; k < ld; ++k) acc ^= a[k] & b[k];
    return popc(acc) & 1;
}

// Gauss–Jordan elimination visiting columns in the given order. A column gets a pivot
// when some row at or below the current rank has a 1 there; the FIRST such row is
// swapped up and cleared out of every other row. Returns the pivot columns in row order.
// Same pivot rule as cpp/qsf/gf2.cpp:42-58, so information sets and their generators
// agree with the reference's.
std::vector<int> eliminate(BitMat& m, const int* order, int count);
std::vector<int> eliminate_all(BitMat& m);

// Basis of {x : m x^T = 0}, one row per free column in increasing column order: the free
// column's own bit plus the pivot columns whose rows carry it (cpp/qsf/gf2.cpp:72-94).
BitMat kernel_basis(const BitMat& m);

// Rows of kernel_basis(other) that are independent of rowspan(self), chosen greedily in
// kernel order (cpp/distfind/css.cpp:94-105).
BitMat quotient(const BitMat& self, const BitMat& other);

bool in_rowspan(const BitMat& G, const u64* v);

}  // namespace qsa
