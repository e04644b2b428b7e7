#include "problem.hpp"

#include <algorithm>
#include <functional>
#include <numeric>

namespace qsa {

// ------------------------------------------------------------------ GF(2) elimination

std::vector<int> eliminate(BitMat& m, const int* order, int count) {
    std::vector<int> piv;
    int rank = 0;
    for (int t = 0; t < count && rank < m.rows; ++t) {
        const int c = order[t];
        const int w = c >> 6;
        const u64 bitm = 1ull << (c & 63);
        int sel = -1;
        for (int r = rank; r < m.rows; ++r)
            if (m.row(r)[w] & bitm) { sel = r; break; }
        if (sel < 0) continue;
        m.swap(rank, sel);
        const u64* p = m.row(rank);
        for (int r = 0; r < m.rows; ++r)
            if (r != rank && (m.row(r)[w] & bitm)) m.xor_into(r, p);
        piv.push_back(c);
        ++rank;
    }
    return piv;
}

std::vector<int> eliminate_all(BitMat& m) {
    std::vector<int> order(m.cols);
    std::iota(order.begin(), order.end(), 0);
    return eliminate(m, order.data(), m.cols);
}

BitMat kernel_basis(const BitMat& in) {
    BitMat m = in;
    std::vector<int> piv = eliminate_all(m);
    std::vector<char> is_piv(m.cols, 0);
    for (int p : piv) is_piv[p] = 1;
    BitMat ker(m.cols - (int)piv.size(), m.cols);
    int out = 0;
    for (int f = 0; f < m.cols; ++f) {
        if (is_piv[f]) continue;
        ker.set(out, f);
        for (size_t r = 0; r < piv.size(); ++r)
            if (m.bit((int)r, f)) ker.set(out, piv[r]);
        ++out;
    }
    return ker;
}

namespace {

// Reducing basis kept in insertion order, each vector tagged with its lowest set bit.
struct Reducer {
    int ld;
    std::vector<std::vector<u64>> vecs;
    std::vector<int> low;
    explicit Reducer(int cols) : ld(words(cols)) {}
    bool reduce(std::vector<u64>& v) const {
        for (size_t i = 0; i < vecs.size(); ++i)
            if ((v[low[i] >> 6] >> (low[i] & 63)) & 1)
                for (int k = 0; k < ld; ++k) v[k] ^= vecs[i][k];
        for (u64 x : v) if (x) return true;
        return false;
    }
    void add(const std::vector<u64>& v) {
        int l = -1;
        for (int k = 0; k < ld && l < 0; ++k) if (v[k]) l = k * 64 + __builtin_ctzll(v[k]);
        vecs.push_back(v);
        low.push_back(l);
    }
};

bool every_row_even(const BitMat& m) {
    for (int i = 0; i < m.rows; ++i) if (m.weight(i) & 1) return false;
    return true;
}

int lightest_row(const BitMat& m) {
    int b = kNoWeight;
    for (int i = 0; i < m.rows; ++i) b = std::min(b, m.weight(i));
    return b;
}

}  // namespace

BitMat quotient(const BitMat& self, const BitMat& other) {
    BitMat ko = kernel_basis(other);
    Reducer red(self.cols);
    for (int i = 0; i < self.rows; ++i) {
        std::vector<u64> v(self.row(i), self.row(i) + self.ld);
        if (red.reduce(v)) red.add(v);
    }
    BitMat out(0, self.cols);
    for (int i = 0; i < ko.rows; ++i) {
        std::vector<u64> v(ko.row(i), ko.row(i) + ko.ld);
        if (!red.reduce(v)) continue;
        red.add(v);
        out.append(ko.row(i));
    }
    return out;
}

bool in_rowspan(const BitMat& G, const u64* v) {
    BitMat R = G;
    std::vector<int> piv = eliminate_all(R);
    std::vector<u64> t(v, v + R.ld);
    for (size_t r = 0; r < piv.size(); ++r)
        if ((t[piv[r] >> 6] >> (piv[r] & 63)) & 1)
            for (int k = 0; k < R.ld; ++k) t[k] ^= R.row((int)r)[k];
    for (u64 x : t) if (x) return false;
    return true;
}

// ------------------------------------------------------------------ problem builders

Problem css_side(const BitMat& Hx, const BitMat& Hz, char side) {
    const BitMat& self = side == 'X' ? Hz : Hx;
    const BitMat& other = side == 'X' ? Hx : Hz;
    Problem p;
    p.n = self.cols;
    p.gen = kernel_basis(self);
    p.detector = quotient(self, other);
    BitMat logicals = quotient(other, self);
    p.even = every_row_even(p.gen);
    p.seed_upper = logicals.rows ? lightest_row(logicals) : 0;
    return p;
}

Problem classical(const BitMat& H) {
    Problem p;
    p.n = H.cols;
    p.gen = kernel_basis(H);
    p.detector = BitMat(0, H.cols);
    p.every_nonzero = true;
    p.even = every_row_even(p.gen);
    p.seed_upper = p.gen.rows ? lightest_row(p.gen) : 0;
    return p;
}

static BitMat swap_halves(const BitMat& m) {
    const int q = m.cols / 2;
    BitMat o(m.rows, m.cols);
    for (int i = 0; i < m.rows; ++i)
        for (int j = 0; j < q; ++j) {
            if (m.bit(i, j)) o.set(i, q + j);
            if (m.bit(i, q + j)) o.set(i, j);
        }
    return o;
}

bool is_css_symplectic(const BitMat& S) {
    const int q = S.cols / 2;
    for (int i = 0; i < S.rows; ++i) {
        bool z = false, x = false;
        for (int j = 0; j < q; ++j) { z = z || S.bit(i, j); x = x || S.bit(i, q + j); }
        if (z && x) return false;
    }
    return true;
}

void split_css(const BitMat& S, BitMat& Hx, BitMat& Hz) {
    const int q = S.cols / 2;
    Hx = BitMat(0, q);
    Hz = BitMat(0, q);
    BitMat zr(1, q), xr(1, q);
    for (int i = 0; i < S.rows; ++i) {
        std::fill(zr.data.begin(), zr.data.end(), 0);
        std::fill(xr.data.begin(), xr.data.end(), 0);
        bool z = false, x = false;
        for (int j = 0; j < q; ++j) {
            if (S.bit(i, j)) { zr.set(0, j); z = true; }
            if (S.bit(i, q + j)) { xr.set(0, j); x = true; }
        }
        if (x) Hx.append(xr.row(0));
        else if (z) Hz.append(zr.row(0));
    }
}

Problem symplectic_plain(const BitMat& S) {
    const int q = S.cols / 2;
    BitMat sw = swap_halves(S);
    Problem p;
    p.n = 2 * q;
    p.qubits = q;
    p.gen = kernel_basis(sw);                      // centralizer of S
    p.detector = swap_halves(quotient(S, sw));     // pre-swapped so <det, c> is a dot
    int best = 0;
    for (int i = 0; i < p.gen.rows; ++i) {
        if (in_rowspan(S, p.gen.row(i))) continue; // a stabilizer, not a logical
        int w = 0;
        for (int j = 0; j < q; ++j) w += (p.gen.bit(i, j) || p.gen.bit(i, q + j)) ? 1 : 0;
        if (best == 0 || w < best) best = w;
    }
    p.seed_upper = best;
    return p;
}

Problem symplectic_isometric(const BitMat& S) {
    const Problem s = symplectic_plain(S);
    const int q = s.qubits;
    Problem e;
    e.n = 3 * q;
    e.even = true;                  // every image (a|b|a^b) has even weight
    e.seed_upper = 2 * s.seed_upper;
    e.gen = BitMat(s.gen.rows, e.n);
    for (int i = 0; i < s.gen.rows; ++i)
        for (int j = 0; j < q; ++j) {
            bool a = s.gen.bit(i, j), b = s.gen.bit(i, q + j);
            if (a) e.gen.set(i, j);
            if (b) e.gen.set(i, q + j);
            if (a != b) e.gen.set(i, 2 * q + j);
        }
    e.detector = BitMat(s.detector.rows, e.n);
    for (int i = 0; i < s.detector.rows; ++i)
        for (int j = 0; j < 2 * q; ++j)
            if (s.detector.bit(i, j)) e.detector.set(i, j);
    return e;
}

// ------------------------------------------------------------------ information sets

namespace {

// Edmonds matroid partition of the column matroid of gen (columns as K-bit vectors),
// by augmenting-path search: a column joins the first class it is independent of; if it
// is dependent on every class, it displaces a member of a circuit, which is then placed
// recursively (never back into the class it left). The span test and circuit bookkeeping
// follow cpp/distfind/css.cpp:231-268 step for step, so the classes (and hence the
// information sets and the BZ lower bounds) match the reference's. Each class keeps its
// reduced basis (with the origin of every basis vector) between span tests: a column that
// joins a class extends it; a swap on an augmenting path (z in, y out, z the XOR of the
// circuit C, y in C) keeps the span, so only the origins that hold y change, to
// origin ^ C ^ {z}, and undoing the swap changes those that hold z by the same mask. The
// span and the circuit of a column do not depend on how the basis was built, since a
// class is independent.
class Partitioner {
  public:
    explicit Partitioner(const BitMat& gen)
        : n_(gen.cols), kw_(words(gen.rows)), ow_(words(gen.cols)), col_((size_t)gen.cols * words(gen.rows), 0) {
        for (int j = 0; j < n_; ++j) {
            bool any = false;
            for (int i = 0; i < gen.rows; ++i)
                if (gen.bit(i, j)) { col_[(size_t)j * kw_ + (i >> 6)] |= 1ull << (i & 63); any = true; }
            (any ? nonzero_ : zero_).push_back(j);
        }
    }

    bool run(std::vector<std::vector<int>>& out) {
        seen_.assign(n_, 0);
        for (int z : nonzero_) {
            std::fill(seen_.begin(), seen_.end(), 0);
            if (!place(z, -1)) {
                classes_.push_back({z});
                basis_.emplace_back();
            }
        }
        if (classes_.empty()) classes_.push_back({});
        for (int j : zero_) classes_[0].push_back(j);
        std::vector<char> hit(n_, 0);
        int total = 0;
        for (auto& c : classes_)
            for (int j : c) { if (hit[j]) return false; hit[j] = 1; ++total; }
        if (total != n_) return false;
        out = classes_;
        return true;
    }

  private:
    const u64* colv(int j) const { return col_.data() + (size_t)j * kw_; }
    // A class's reduced basis: vector b (kw words) with pivot lead[b] (its lowest set bit
    // when it was added; later vectors are reduced on it) and origin o (ow words: the class
    // columns whose XOR it is).
    struct Basis {
        std::vector<u64> v, o;
        std::vector<int> lead;
        bool valid = false;
    };
    // tv_ = column col reduced by basis B, to_ the class columns it was reduced by
    void reduce(const Basis& B, int col) {
        const int kw = kw_, ow = ow_;
        tv_.assign(colv(col), colv(col) + kw);
        to_.assign(ow, 0ull);
        for (size_t b = 0; b < B.lead.size(); ++b) {
            const int l = B.lead[b];
            if ((tv_[l >> 6] >> (l & 63)) & 1) {
                const u64* v = B.v.data() + b * kw;
                const u64* o = B.o.data() + b * ow;
                for (int k = 0; k < kw; ++k) tv_[k] ^= v[k];
                for (int k = 0; k < ow; ++k) to_[k] ^= o[k];
            }
        }
    }
    // add column col (reduced into tv_ / to_ by reduce) to B if it is independent of it
    void extend(Basis& B, int col) {
        int l = -1;
        for (int k = 0; k < kw_ && l < 0; ++k)
            if (tv_[k]) l = k * 64 + __builtin_ctzll(tv_[k]);
        if (l < 0) return;
        to_[col >> 6] ^= 1ull << (col & 63);
        B.v.insert(B.v.end(), tv_.begin(), tv_.end());
        B.o.insert(B.o.end(), to_.begin(), to_.end());
        B.lead.push_back(l);
    }
    // class c's member `out` swapped for `in` on a circuit: the origins that hold `out`
    // take delta = (circuit mask) ^ {z}
    void swap_member(int c, const std::vector<u64>& delta, int out) {
        Basis& B = basis_[c];
        if (!B.valid) return;
        for (size_t b = 0; b < B.lead.size(); ++b) {
            u64* o = B.o.data() + b * ow_;
            if ((o[out >> 6] >> (out & 63)) & 1)
                for (int k = 0; k < ow_; ++k) o[k] ^= delta[k];
        }
    }
    Basis& basis(int c) {
        Basis& B = basis_[c];
        if (!B.valid) {
            B.v.clear(); B.o.clear(); B.lead.clear();
            for (int idx : classes_[c]) {
                reduce(B, idx);
                extend(B, idx);
            }
            B.valid = true;
        }
        return B;
    }
    // true iff column z lies in the span of class c; then *circuit = the class columns
    // whose XOR is z, ascending. A class is always independent, so that set is unique and
    // equals the one the reference's origin-tracking reduction finds (css.cpp:231-268).
    bool spans(int c, int z, std::vector<int>* circuit) {
        reduce(basis(c), z);
        for (u64 x : tv_)
            if (x) return false;
        if (circuit) {
            circuit->clear();
            for (int k = 0; k < ow_; ++k)
                for (u64 m = to_[k]; m; m &= m - 1) circuit->push_back(k * 64 + __builtin_ctzll(m));
        }
        return true;
    }
    bool place(int z, int forbidden) {
        for (int c = 0; c < (int)classes_.size(); ++c)
            if (c != forbidden && !spans(c, z, nullptr)) {   // tv_ / to_ hold z reduced by c
                classes_[c].push_back(z);
                extend(basis_[c], z);
                return true;
            }
        for (int c = 0; c < (int)classes_.size(); ++c) {
            if (c == forbidden) continue;
            std::vector<int> circ;
            if (!spans(c, z, &circ)) continue;
            std::vector<u64> delta(ow_, 0ull);   // the circuit's mask, and z
            for (int y : circ) delta[y >> 6] ^= 1ull << (y & 63);
            delta[z >> 6] ^= 1ull << (z & 63);
            for (int y : circ) {
                if (seen_[y]) continue;
                seen_[y] = 1;
                auto& cls = classes_[c];
                cls.erase(std::find(cls.begin(), cls.end(), y));
                cls.push_back(z);
                swap_member(c, delta, y);
                if (place(y, c)) return true;
                cls.erase(std::find(cls.begin(), cls.end(), z));
                cls.push_back(y);
                swap_member(c, delta, z);
            }
        }
        return false;
    }

    int n_, kw_, ow_;
    std::vector<u64> col_;
    std::vector<u64> tv_, to_;   // span-test scratch
    std::vector<int> nonzero_, zero_;
    std::vector<std::vector<int>> classes_;
    std::vector<Basis> basis_;   // of each class
    std::vector<char> seen_;
};

}  // namespace

int nonzero_columns(const BitMat& gen) {
    std::vector<u64> any(gen.ld, 0);
    for (int i = 0; i < gen.rows; ++i)
        for (int k = 0; k < gen.ld; ++k) any[k] |= gen.row(i)[k];
    int c = 0;
    for (u64 x : any) c += __builtin_popcountll(x);
    return c;
}

bool matroid_parts(const BitMat& gen, std::vector<std::vector<int>>& parts) {
    std::vector<std::vector<int>> all;
    Partitioner P(gen);
    if (!P.run(all)) return false;
    // plain (unstable) sort by size, as the reference orders its classes (css.cpp:354-355)
    std::sort(all.begin(), all.end(),
              [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() > b.size(); });
    parts.clear();
    for (auto& c : all)
        if (!c.empty()) parts.push_back(c);
    return true;
}

}  // namespace qsa
