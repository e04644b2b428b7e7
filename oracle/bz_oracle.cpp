// ORACLE — test infrastructure only. A CPU restatement of the reference's
// Brouwer–Zimmermann path, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the CHECKER. Nothing in the product links, loads or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference repository root). tests/test_cpu_oracle.py pins the restatement: golden
// vectors from the reference's own CPU library, brute force over all 2^K messages, a
// meet-in-the-middle search, closed forms and textbook distances.
//
// Build:  g++ -O3 -std=c++17 -fPIC -shared -pthread oracle/bz_oracle.cpp -o oracle/_build/liboracle.so
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace orc {

using u64 = std::uint64_t;
using u8 = std::uint8_t;
constexpr int NONE = 1 << 30;                 // cpp/qsf/backend.hpp:19
constexpr u64 BINF = ~0ull;                   // cpp/qsf/combinatorics.hpp:14

inline int wfor(int bits) { return (bits + 63) >> 6; }            // bits.hpp:15
inline int pc64(u64 x) { return __builtin_popcountll(x); }        // bits.hpp:17-26

// ---- dense GF(2) matrix (cpp/qsf/gf2.hpp:9-31) -------------------------------------
struct Mat {
    int rows = 0, cols = 0, stride = 0;
    std::vector<u64> w;
    Mat() = default;
    Mat(int r, int c) : rows(r), cols(c), stride(wfor(c)), w((size_t)r * wfor(c), 0) {}
    u64* row(int i) { return w.data() + (size_t)i * stride; }
    const u64* row(int i) const { return w.data() + (size_t)i * stride; }
    int get(int i, int j) const { return (int)((row(i)[j >> 6] >> (j & 63)) & 1); }
    void set1(int i, int j) { row(i)[j >> 6] |= 1ull << (j & 63); }
    int weight(int i) const { int s = 0; for (int k = 0; k < stride; ++k) s += pc64(row(i)[k]); return s; }
    void addrow(int src, int dst) { for (int k = 0; k < stride; ++k) row(dst)[k] ^= row(src)[k]; }
    void swaprows(int a, int b) { if (a != b) for (int k = 0; k < stride; ++k) std::swap(row(a)[k], row(b)[k]); }
    void push(const u64* v) { w.insert(w.end(), v, v + stride); ++rows; }
};

inline int dot(const u64* a, const u64* b, int stride) {          // bits.hpp:41-45
    u64 acc = 0;
    for (int k = 0; k < stride; ++k) acc ^= a[k] & b[k];
    return pc64(acc) & 1;
}

Mat dense(const u8* a, int r, int c) {                              // gf2.cpp:17-23
    Mat m(r, c);
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < c; ++j)
            if (a[(size_t)i * c + j] & 1) m.set1(i, j);
    return m;
}

// Gauss–Jordan over the given column order; pivot = first row at/below the rank with a 1
// in that column (gf2.cpp:25-40 full order, :42-58 restricted order).
std::vector<int> rref_cols(Mat& m, const std::vector<int>& cols) {
    std::vector<int> piv;
    int rank = 0;
    for (size_t ci = 0; ci < cols.size() && rank < m.rows; ++ci) {
        int c = cols[ci], sel = -1;
        for (int r = rank; r < m.rows; ++r) if (m.get(r, c)) { sel = r; break; }
        if (sel < 0) continue;
        m.swaprows(rank, sel);
        for (int r = 0; r < m.rows; ++r) if (r != rank && m.get(r, c)) m.addrow(rank, r);
        piv.push_back(c);
        ++rank;
    }
    return piv;
}
std::vector<int> rref(Mat& m) {
    std::vector<int> all(m.cols);
    std::iota(all.begin(), all.end(), 0);
    return rref_cols(m, all);
}

Mat nullspace(const Mat& in) {                                      // gf2.cpp:72-94
    Mat m = in;
    std::vector<int> piv = rref(m);
    std::vector<char> isp(m.cols, 0);
    for (int p : piv) isp[p] = 1;
    Mat ns(m.cols - (int)piv.size(), m.cols);
    int out = 0;
    for (int f = 0; f < m.cols; ++f) {
        if (isp[f]) continue;
        ns.set1(out, f);
        for (size_t r = 0; r < piv.size(); ++r) if (m.get((int)r, f)) ns.set1(out, piv[r]);
        ++out;
    }
    return ns;
}

// Rows of ker(other) independent of rowspan(self), greedily in ker order
// (cpp/distfind/css.cpp:94-105 with the reducing basis of :18-45).
Mat quotient_basis(const Mat& self, const Mat& other) {
    Mat ko = nullspace(other);
    const int st = wfor(self.cols);
    std::vector<std::vector<u64>> B;
    std::vector<int> P;
    auto reduce = [&](std::vector<u64>& v) {
        for (size_t i = 0; i < B.size(); ++i)
            if ((v[P[i] >> 6] >> (P[i] & 63)) & 1) for (int k = 0; k < st; ++k) v[k] ^= B[i][k];
        for (int k = 0; k < st; ++k) if (v[k]) return true;
        return false;
    };
    auto low = [&](const std::vector<u64>& v) {
        for (int k = 0; k < st; ++k) if (v[k]) return k * 64 + __builtin_ctzll(v[k]);
        return -1;
    };
    for (int i = 0; i < self.rows; ++i) {
        std::vector<u64> t(self.row(i), self.row(i) + st);
        if (reduce(t)) { B.push_back(t); P.push_back(low(t)); }
    }
    Mat out(0, self.cols);
    for (int i = 0; i < ko.rows; ++i) {
        std::vector<u64> t(ko.row(i), ko.row(i) + st);
        if (!reduce(t)) continue;
        B.push_back(t); P.push_back(low(t));
        out.push(ko.row(i));
    }
    return out;
}

// ---- distance subproblem (cpp/distfind/css.hpp:21-30) -----------------------------
struct Problem {
    int n = 0;
    Mat gen;      // K x n code space basis
    Mat check;    // logical detector
    bool even = false, count_all = false;
    int seed_upper = 0;
};

bool all_even(const Mat& m) { for (int i = 0; i < m.rows; ++i) if (m.weight(i) & 1) return false; return true; }
int min_weight(const Mat& m) { int b = NONE; for (int i = 0; i < m.rows; ++i) b = std::min(b, m.weight(i)); return b; }

Problem css_side(const Mat& Hself, const Mat& Hother) {             // css.cpp:78-87
    Problem p;
    p.n = Hself.cols;
    p.gen = nullspace(Hself);
    p.check = quotient_basis(Hself, Hother);
    Mat logicals = quotient_basis(Hother, Hself);
    p.even = all_even(p.gen);
    p.seed_upper = logicals.rows ? min_weight(logicals) : 0;
    return p;
}
Problem css_problem(const Mat& Hx, const Mat& Hz, char which) {     // css.cpp:107-109
    return which == 'X' ? css_side(Hz, Hx) : css_side(Hx, Hz);
}
Problem classical_problem(const Mat& H) {                           // css.cpp:111-120
    Problem p;
    p.n = H.cols;
    p.gen = nullspace(H);
    p.check = Mat(0, H.cols);
    p.count_all = true;
    p.even = all_even(p.gen);
    p.seed_upper = p.gen.rows ? min_weight(p.gen) : 0;
    return p;
}

// ---- information sets (css.cpp:177-203 greedy, :274-328 matroid partition, :336-366) ----
struct Gamma { Mat g; int rank; };

std::vector<Gamma> gammas_greedy(const Mat& gen) {
    std::vector<Gamma> seq;
    const int n = gen.cols;
    std::vector<int> act(n);
    std::iota(act.begin(), act.end(), 0);
    Mat W = gen;
    while (!act.empty()) {
        std::vector<int> piv = rref_cols(W, act);
        seq.push_back({W, (int)piv.size()});
        std::vector<char> isp(n, 0), zero(n, 1);
        for (int c : piv) isp[c] = 1;
        for (int r = 0; r < W.rows; ++r) for (int c : act) if (W.get(r, c)) zero[c] = 0;
        std::vector<int> nxt;
        for (int c : act) if (!isp[c] && !zero[c]) nxt.push_back(c);
        if (nxt.size() == act.size()) break;
        act.swap(nxt);
    }
    return seq;
}

static std::vector<int> sdiff(const std::vector<int>& a, const std::vector<int>& b) {
    std::vector<int> r;
    size_t i = 0, j = 0;
    while (i < a.size() && j < b.size()) {
        if (a[i] < b[j]) r.push_back(a[i++]);
        else if (b[j] < a[i]) r.push_back(b[j++]);
        else { ++i; ++j; }
    }
    while (i < a.size()) r.push_back(a[i++]);
    while (j < b.size()) r.push_back(b[j++]);
    return r;
}

// In-span test of column z over class `cls` with origin tracking (css.cpp:231-268).
static bool span_of(const std::vector<int>& cls, const std::vector<u64>& cv, int kw, int z,
                    std::vector<int>* circuit) {
    std::vector<std::vector<u64>> B; std::vector<int> P; std::vector<std::vector<int>> O;
    auto low = [&](const std::vector<u64>& v) {
        for (int k = 0; k < kw; ++k) if (v[k]) return k * 64 + __builtin_ctzll(v[k]);
        return -1;
    };
    for (int idx : cls) {
        std::vector<u64> v(cv.begin() + (size_t)idx * kw, cv.begin() + (size_t)idx * kw + kw);
        std::vector<int> o{idx};
        for (size_t b = 0; b < B.size(); ++b)
            if ((v[P[b] >> 6] >> (P[b] & 63)) & 1) { for (int k = 0; k < kw; ++k) v[k] ^= B[b][k]; o = sdiff(o, O[b]); }
        int l = low(v);
        if (l >= 0) { B.push_back(v); P.push_back(l); O.push_back(o); }
    }
    std::vector<u64> v(cv.begin() + (size_t)z * kw, cv.begin() + (size_t)z * kw + kw);
    std::vector<int> o;
    for (size_t b = 0; b < B.size(); ++b)
        if ((v[P[b] >> 6] >> (P[b] & 63)) & 1) { for (int k = 0; k < kw; ++k) v[k] ^= B[b][k]; o = sdiff(o, O[b]); }
    for (int k = 0; k < kw; ++k) if (v[k]) return false;
    if (circuit) *circuit = o;
    return true;
}

static bool matroid_partition(const Mat& gen, std::vector<std::vector<int>>& out) {
    const int n = gen.cols, K = gen.rows, kw = wfor(K);
    std::vector<u64> cv((size_t)n * kw, 0);
    std::vector<int> nz, zs;
    for (int j = 0; j < n; ++j) {
        bool any = false;
        for (int i = 0; i < K; ++i) if (gen.get(i, j)) { cv[(size_t)j * kw + (i >> 6)] |= 1ull << (i & 63); any = true; }
        (any ? nz : zs).push_back(j);
    }
    std::vector<std::vector<int>> cls;
    std::vector<char> vis(n, 0);
    std::function<bool(int, int)> place = [&](int z, int forb) -> bool {
        for (int c = 0; c < (int)cls.size(); ++c) {
            if (c == forb) continue;
            if (!span_of(cls[c], cv, kw, z, nullptr)) { cls[c].push_back(z); return true; }
        }
        for (int c = 0; c < (int)cls.size(); ++c) {
            if (c == forb) continue;
            std::vector<int> circ;
            if (!span_of(cls[c], cv, kw, z, &circ)) continue;
            for (int y : circ) {
                if (vis[y]) continue;
                vis[y] = 1;
                auto& v = cls[c];
                v.erase(std::find(v.begin(), v.end(), y));
                v.push_back(z);
                if (place(y, c)) return true;
                v.erase(std::find(v.begin(), v.end(), z));
                v.push_back(y);
            }
        }
        return false;
    };
    for (int z : nz) {
        std::fill(vis.begin(), vis.end(), 0);
        if (!place(z, -1)) cls.push_back({z});
    }
    if (cls.empty()) cls.push_back({});
    for (int j : zs) cls[0].push_back(j);
    std::vector<char> seen(n, 0);
    int tot = 0;
    for (auto& c : cls) for (int j : c) { if (seen[j]) return false; seen[j] = 1; ++tot; }
    if (tot != n) return false;
    out.swap(cls);
    return true;
}

bool matroid_partition_public(const Mat& gen, std::vector<std::vector<int>>& out) { return matroid_partition(gen, out); }

static int nfull(const std::vector<Gamma>& s, int K) { int f = 0; for (auto& g : s) if (g.rank == K) ++f; return f; }

std::vector<Gamma> gamma_sequence(const Mat& gen) {
    const int K = gen.rows, n = gen.cols;
    std::vector<Gamma> greedy = gammas_greedy(gen);
    if (K == 0) return greedy;
    int nnz = 0;
    for (int j = 0; j < n; ++j) for (int i = 0; i < K; ++i) if (gen.get(i, j)) { ++nnz; break; }
    const int maxfull = nnz / K, gfull = nfull(greedy, K);
    if (gfull >= maxfull || n > 200) return greedy;
    std::vector<std::vector<int>> part;
    if (!matroid_partition(gen, part)) return greedy;
    std::sort(part.begin(), part.end(), [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() > b.size(); });
    std::vector<Gamma> seq;
    for (auto& cols : part) {
        if (cols.empty()) continue;
        Mat W = gen;
        std::vector<int> piv = rref_cols(W, cols);
        seq.push_back({W, (int)piv.size()});
    }
    if (seq.empty() || nfull(seq, K) <= gfull) return greedy;
    return seq;
}

// ---- enumeration plan + CPU enumerator (cpp/qsf/backend.hpp:21-42, enum_core.hpp) -----
struct Plan {
    int n, stride, K, d, ng;
    const u64* gamma;          // ng x K x stride, row-major (g,row,word)
    int kcheck;
    const u64* check;          // kcheck x stride
    int best;
};

struct Binom {
    int N = 0, Kmax = 0;
    std::vector<u64> c;
    void build(int n, int k) {                                      // combinatorics.hpp:17-46
        N = n; Kmax = k;
        c.assign((size_t)(n + 1) * (k + 1), 0);
        for (int i = 0; i <= n; ++i) {
            c[(size_t)i * (k + 1)] = 1;
            for (int j = 1; j <= std::min(i, k); ++j) {
                u64 a = c[(size_t)(i - 1) * (k + 1) + j - 1];
                u64 b = c[(size_t)(i - 1) * (k + 1) + j];
                c[(size_t)i * (k + 1) + j] = (a == BINF || b == BINF || a > BINF - b) ? BINF : a + b;
            }
        }
    }
    u64 operator()(int n, int k) const {
        if (k < 0 || n < 0 || k > n) return 0;
        if (n > N || k > Kmax) return BINF;
        return c[(size_t)n * (Kmax + 1) + k];
    }
};

// lexicographic unrank of a d-subset of [0,K) (combinatorics.hpp:52-61)
static void unrank_lex(const Binom& B, int K, int d, u64 r, int* pos) {
    int x = 0;
    for (int i = 0; i < d; ++i)
        for (;;) {
            u64 cnt = B(K - 1 - x, d - 1 - i);
            if (r < cnt) { pos[i] = x++; break; }
            r -= cnt; ++x;
        }
}

static bool logical(const Plan& p, const u64* cw) {                 // enum_core.hpp:77-87
    if (p.kcheck == 0) return true;
    for (int c = 0; c < p.kcheck; ++c)
        if (dot(p.check + (size_t)c * p.stride, cw, p.stride)) return true;
    return false;
}

// Sink semantics: min-mode keeps the lightest logical lighter than the running bound;
// collect-mode appends every nonzero codeword with weight <= keep (enum_core.hpp:74-103).
struct Sink {
    bool collect = false;
    int best = NONE;           // min mode
    int keep = 0;              // collect mode
    std::vector<u64> hits;
    int bound() const { return collect ? keep + 1 : best; }
    void hit(const Plan& p, const u64* cw, int wt) {
        if (collect) { if (wt > 0) hits.insert(hits.end(), cw, cw + p.stride); return; }
        if (logical(p, cw)) best = wt;
    }
};

// Two-level enumeration of combinations [start,end) in lexicographic order: the outer d-1
// rows are XOR-maintained, the last row sweeps; weight is accumulated word by word with an
// early exit at the bound (enum_core.hpp:109-194).
static void enum_range(const Plan& p, const Binom& B, u64 start, u64 end, Sink& s) {
    const int K = p.K, st = p.stride, D = p.d, M = D - 1;
    std::vector<int> pos(D);
    std::vector<u64> base(st), cw(st);
    for (int g = 0; g < p.ng; ++g) {
        const u64* G = p.gamma + (size_t)g * K * st;
        unrank_lex(B, K, D, start, pos.data());
        std::fill(base.begin(), base.end(), 0);
        for (int i = 0; i < M; ++i) for (int w = 0; w < st; ++w) base[w] ^= G[(size_t)pos[i] * st + w];
        int last = pos[D - 1];
        u64 done = 0, need = end - start;
        while (done < need) {
            int hi = K;
            if ((u64)(K - last) > need - done) hi = last + (int)(need - done);
            const int first = last;
            for (; last < hi; ++last) {
                int bnd = s.bound(), wt = 0;
                bool light = true;
                for (int w = 0; w < st && light; ++w) {
                    cw[w] = base[w] ^ G[(size_t)last * st + w];
                    wt += pc64(cw[w]);
                    light = wt < bnd;
                }
                if (!light) continue;
                s.hit(p, cw.data(), wt);
            }
            done += (u64)(hi - first);
            if (done >= need) break;
            int j = M - 1;
            while (j >= 0 && pos[j] == K - D + j) --j;
            if (j < 0) break;
            for (int t = j; t < M; ++t) for (int w = 0; w < st; ++w) base[w] ^= G[(size_t)pos[t] * st + w];
            ++pos[j];
            for (int t = j + 1; t < M; ++t) pos[t] = pos[t - 1] + 1;
            for (int t = j; t < M; ++t) for (int w = 0; w < st; ++w) base[w] ^= G[(size_t)pos[t] * st + w];
            last = pos[M - 1] + 1;
        }
    }
}

// Threads split [0, C(K,d)) proportionally (enum_core.hpp:200-215); thread count as
// cpu_level_threads (backend_cpu.cpp:18-25).
static int level_threads(int T, u64 total) {
    if (T <= 0) T = std::max(1u, std::thread::hardware_concurrency());
    if (total < 4096) T = 1;
    if ((u64)T > total) T = (int)total;
    return T;
}

static void run_level(const Plan& p, const Binom& B, int threads, std::vector<Sink>& sinks, bool collect, int keep) {
    u64 total = B(p.K, p.d);
    if (total == 0) { sinks.assign(1, Sink{}); return; }
    int T = level_threads(threads, total);
    sinks.assign(T, Sink{});
    for (auto& s : sinks) { s.collect = collect; s.keep = keep; s.best = p.best; }
    if (T == 1) { enum_range(p, B, 0, total, sinks[0]); return; }
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
        u64 a = (u64)((__uint128_t)total * t / T), b = (u64)((__uint128_t)total * (t + 1) / T);
        pool.emplace_back([&, t, a, b] { enum_range(p, B, a, b, sinks[t]); });
    }
    for (auto& th : pool) th.join();
}

int enumerate(const Plan& p, const Binom& B, int threads) {         // backend_cpu.cpp:33-47
    std::vector<Sink> sinks;
    run_level(p, B, threads, sinks, false, 0);
    int best = p.best;
    for (auto& s : sinks) best = std::min(best, s.best);
    return best;
}

// ---- BZ driver (cpp/distfind/bz.cpp) ------------------------------------------------
struct Result { int distance = -1, lower_bound = 0, proven = 0, levels = 0; double seconds = 0; };

// Random information-set upper bound (bz.cpp:24-70): trials split over `nthreads`
// std::thread workers, worker t seeded seed + 0x9e3779b9*(t+1) (a single worker uses the
// bare seed), each trial a std::shuffle of the column order followed by a restricted RREF.
int is_seed(const Problem& pr, int init, int trials, unsigned seed, int nthreads) {
    const int n = pr.n, K = pr.gen.rows, st = pr.gen.stride;
    if (K == 0) return init;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > trials) nthreads = trials;
    auto run = [&](int lo, int hi, unsigned ts) {
        std::mt19937 rng(ts);
        std::vector<int> cols(n);
        int best = init;
        for (int it = lo; it < hi; ++it) {
            std::iota(cols.begin(), cols.end(), 0);
            std::shuffle(cols.begin(), cols.end(), rng);
            Mat M = pr.gen;
            rref_cols(M, cols);
            for (int r = 0; r < M.rows; ++r) {
                int w = M.weight(r);
                if (w == 0 || w >= best) continue;
                bool lg = pr.count_all;
                for (int c = 0; c < pr.check.rows && !lg; ++c) if (dot(pr.check.row(c), M.row(r), st)) lg = true;
                if (lg) best = w;
            }
        }
        return best;
    };
    if (nthreads == 1) return run(0, trials, seed);
    std::vector<int> res(nthreads, init);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) {
        int lo = (int)((long)trials * t / nthreads), hi = (int)((long)trials * (t + 1) / nthreads);
        pool.emplace_back([&, t, lo, hi] { res[t] = run(lo, hi, seed + 0x9e3779b9u * (unsigned)(t + 1)); });
    }
    for (auto& th : pool) th.join();
    int b = init;
    for (int r : res) b = std::min(b, r);
    return b;
}

// Sound level schedule (not in the reference). The reference enumerates a generator of
// rank r < K only at levels d with (d+1) - (K-r) > 0 (bz.cpp:142-144) yet adds its
// max(0, (d+1) - (K-r)) to the bound (bz.cpp:159-161); that contribution assumes every
// combination of at most d rows of that generator was enumerated, including levels
// 1..K-r-1, which were skipped, so a lighter codeword there can be missed. In sound mode
// the generators that contribute by the level D where the bound first reaches the upper
// bound (those with K - r <= D) are enumerated from level 1, and only they enter the bound.
static bool g_sound = false;
int sound_prefix(int K, const std::vector<int>& ranks, bool even, int upper) {
    int D = K;
    for (int d = 1; d <= K; ++d) {
        int o = 0;
        for (int r : ranks) o += std::max(0, (d + 1) - (K - r));
        if (even) o += o & 1;
        if (o >= upper) { D = d; break; }
    }
    int a = 0;
    while (a < (int)ranks.size() && K - ranks[a] <= D) ++a;   // ranks sorted descending
    return a;
}

// One BZ subproblem, stepped one weight level at a time (bz.cpp:196-272).
struct Side {
    bool has = false, even = false;
    int n = 0, K = 0, st = 0, ng = 0, kcheck = 0;
    std::vector<u64> gamma, check;
    std::vector<int> ranks;
    Binom B;
    int inner = NONE, outer = 0, level = 0;
    bool proven = false, resolved = false;
    int threads = 0;

    void setup(const Problem& pr, int thr, int is_threads) {
        threads = thr;
        n = pr.n; K = pr.gen.rows; st = wfor(n); even = pr.even;
        has = !(K == 0 || (!pr.count_all && pr.check.rows == 0));
        if (!has) { resolved = true; return; }
        std::vector<Gamma> seq = gamma_sequence(pr.gen);
        std::stable_sort(seq.begin(), seq.end(), [](const Gamma& a, const Gamma& b) { return a.rank > b.rank; });
        ng = (int)seq.size();
        gamma.assign((size_t)ng * K * st, 0);
        ranks.resize(ng);
        for (int g = 0; g < ng; ++g) {
            ranks[g] = seq[g].rank;
            for (int r = 0; r < K; ++r) for (int w = 0; w < st; ++w) gamma[((size_t)g * K + r) * st + w] = seq[g].g.row(r)[w];
        }
        kcheck = pr.count_all ? 0 : pr.check.rows;
        check.assign((size_t)kcheck * st, 0);
        for (int c = 0; c < kcheck; ++c) for (int w = 0; w < st; ++w) check[(size_t)c * st + w] = pr.check.row(c)[w];
        B.build(K, K);
        inner = pr.seed_upper > 0 ? pr.seed_upper : NONE;
        inner = is_seed(pr, inner, 256, 0x9e3779b9u, is_threads);
    }
    mutable int nsound = -1;   // sound mode: generators enumerated from level 1
    int active(int d) const {
        if (g_sound) {
            if (nsound < 0) nsound = sound_prefix(K, ranks, even, inner);
            return nsound;
        }
        int a = 0;
        while (a < ng && (d + 1) - (K - ranks[a]) > 0) ++a;
        return a;
    }
    void step(int d) {
        int a = active(d);
        int found = NONE;
        if (a > 0) {
            Plan p{n, st, K, d, a, gamma.data(), kcheck, check.data(), inner};
            found = enumerate(p, B, threads);
        }
        inner = std::min(inner, found);
        ++level;
        outer = 0;
        const int counted = g_sound ? nsound : ng;
        for (int g = 0; g < counted; ++g) outer += std::max(0, (d + 1) - (K - ranks[g]));
        if (even) outer += outer & 1;
    }
};

Result bz_single(const Problem& pr, int threads, int is_threads) {   // bz.cpp:72-182
    auto t0 = std::chrono::steady_clock::now();
    Result res;
    Side S;
    S.setup(pr, threads, is_threads);
    if (S.has) {
        for (int d = 1; d <= S.K; ++d) {
            S.step(d);
            if (S.inner <= S.outer) { S.proven = true; break; }
        }
        res.distance = S.inner >= NONE ? -1 : S.inner;
        res.levels = S.level;
        res.lower_bound = S.outer;
        res.proven = S.proven;
    }
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

Result bz_min_pair(const Problem& pz, const Problem& px, int threads, int is_threads) {  // bz.cpp:296-357
    auto t0 = std::chrono::steady_clock::now();
    Side Z, X;
    Z.setup(pz, threads, is_threads);
    X.setup(px, threads, is_threads);
    Side* sides[2] = {&Z, &X};
    int best = NONE;
    if (Z.has) best = std::min(best, Z.inner);
    if (X.has) best = std::min(best, X.inner);
    const int maxw = std::max(Z.has ? Z.K : 0, X.has ? X.K : 0);
    for (int d = 1; d <= maxw; ++d) {
        for (Side* S : sides) {
            if (S->resolved || d > S->K) continue;
            if (S->outer >= best) { S->resolved = true; continue; }
            S->step(d);
            best = std::min(best, S->inner);
            if (S->inner <= S->outer) { S->proven = true; S->resolved = true; }
            else if (S->outer >= best) S->resolved = true;
        }
        if (Z.resolved && X.resolved) break;
    }
    auto v = [](const Side& S) { return S.has ? S.inner : NONE; };
    int dist = std::min(v(Z), v(X));
    Result r;
    r.distance = dist >= NONE ? -1 : dist;
    int lb = std::min(Z.has ? Z.outer : NONE, X.has ? X.outer : NONE);
    r.lower_bound = lb >= NONE ? 0 : lb;
    r.levels = std::max(Z.level, X.level);
    r.proven = Z.resolved && X.resolved && dist < NONE;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

// ---- symplectic route (cpp/distfind/stab.cpp) ---------------------------------------
Mat swap_halves(const Mat& m) {                                     // stab.cpp:32-42
    const int q = m.cols / 2;
    Mat o(m.rows, m.cols);
    for (int i = 0; i < m.rows; ++i)
        for (int j = 0; j < q; ++j) {
            if (m.get(i, j)) o.set1(i, q + j);
            if (m.get(i, q + j)) o.set1(i, j);
        }
    return o;
}
bool is_css(const Mat& S) {                                         // stab.cpp:56-67
    const int q = S.cols / 2;
    for (int i = 0; i < S.rows; ++i) {
        bool z = false, x = false;
        for (int j = 0; j < q; ++j) { z |= S.get(i, j); x |= S.get(i, q + j); }
        if (z && x) return false;
    }
    return true;
}
void split(const Mat& S, Mat& Hx, Mat& Hz) {                        // stab.cpp:69-85
    const int q = S.cols / 2;
    Hx = Mat(0, q); Hz = Mat(0, q);
    for (int i = 0; i < S.rows; ++i) {
        Mat zr(1, q), xr(1, q);
        bool z = false, x = false;
        for (int j = 0; j < q; ++j) {
            if (S.get(i, j)) { zr.set1(0, j); z = true; }
            if (S.get(i, q + j)) { xr.set1(0, j); x = true; }
        }
        if (x) Hx.push(xr.row(0));
        else if (z) Hz.push(zr.row(0));
    }
}
bool in_rowspace(const Mat& G, const Mat& v1) {                     // css.cpp:139-146
    Mat R = G;
    std::vector<int> piv = rref(R);
    std::vector<u64> t(v1.row(0), v1.row(0) + v1.stride);
    for (size_t r = 0; r < piv.size(); ++r)
        if ((t[piv[r] >> 6] >> (piv[r] & 63)) & 1) for (int k = 0; k < R.stride; ++k) t[k] ^= R.row((int)r)[k];
    for (u64 x : t) if (x) return false;
    return true;
}
// Symplectic problem + weight-doubling isometry (stab.cpp:98-132, :147-178); returns
// the Hamming problem of length 3q whose distance is twice the symplectic distance.
Problem symplectic_isometric(const Mat& S) {
    const int q = S.cols / 2;
    Mat gen = nullspace(swap_halves(S));
    Mat L = quotient_basis(S, swap_halves(S));
    Mat chk = swap_halves(L);
    int best = 0;
    for (int i = 0; i < gen.rows; ++i) {
        Mat v(1, gen.cols);
        for (int k = 0; k < v.stride; ++k) v.row(0)[k] = gen.row(i)[k];
        if (in_rowspace(S, v)) continue;
        int w = 0;
        for (int j = 0; j < q; ++j) w += (gen.get(i, j) | gen.get(i, q + j));
        if (best == 0 || w < best) best = w;
    }
    Problem e;
    e.n = 3 * q;
    e.even = true;
    e.count_all = false;
    e.seed_upper = best > 0 ? 2 * best : 0;
    e.gen = Mat(gen.rows, e.n);
    for (int i = 0; i < gen.rows; ++i)
        for (int j = 0; j < q; ++j) {
            int a = gen.get(i, j), b = gen.get(i, q + j);
            if (a) e.gen.set1(i, j);
            if (b) e.gen.set1(i, q + j);
            if (a ^ b) e.gen.set1(i, 2 * q + j);
        }
    e.check = Mat(chk.rows, e.n);
    for (int i = 0; i < chk.rows; ++i) for (int j = 0; j < 2 * q; ++j) if (chk.get(i, j)) e.check.set1(i, j);
    return e;
}

// ---- brute force and meet-in-the-middle (independent of BZ) ----------------------
// Minimum weight over all nonzero messages of the K x n generator that are logical.
int brute_force(const Problem& pr) {
    const int K = pr.gen.rows, st = pr.gen.stride;
    if (K == 0 || (!pr.count_all && pr.check.rows == 0)) return -1;
    int best = NONE;
    std::vector<u64> cw(st, 0);
    // Gray-code walk over all 2^K messages: one row XOR per step.
    for (u64 i = 1; i < (1ull << K); ++i) {
        int r = __builtin_ctzll(i);
        for (int w = 0; w < st; ++w) cw[w] ^= pr.gen.row(r)[w];
        int wt = 0;
        for (int w = 0; w < st; ++w) wt += pc64(cw[w]);
        if (wt == 0 || wt >= best) continue;
        bool lg = pr.count_all;
        for (int c = 0; c < pr.check.rows && !lg; ++c) if (dot(pr.check.row(c), cw.data(), st)) lg = true;
        if (lg) best = wt;
    }
    return best >= NONE ? -1 : best;
}

// Meet in the middle over error supports (cpp/distfind/mitm.cpp:244-392 restated as the
// textbook syndrome join): a logical of weight w exists iff some e1 of weight ceil(w/2)
// over columns [0,n) and e2 of weight floor(w/2), disjoint, have equal syndromes under the
// parity check P = nullspace(gen) and a nonzero logical fingerprint. The search runs w =
// 1, 2, ... and returns the first w that has one.
int mitm(const Problem& pr, int wmax) {
    const int K = pr.gen.rows, n = pr.n;
    if (K == 0 || (!pr.count_all && pr.check.rows == 0)) return -1;
    Mat P = nullspace(pr.gen);            // (n-K) x n, rowspace = code's dual
    const int m = P.rows;
    const int fpk = pr.count_all ? 0 : pr.check.rows;
    // column syndromes (m bits) and fingerprints (fpk bits), packed in two u64 words each
    const int sw = wfor(m + 1), fw = wfor(fpk + 1);
    std::vector<u64> csyn((size_t)n * sw, 0), cfp((size_t)n * fw, 0);
    for (int j = 0; j < n; ++j) {
        for (int r = 0; r < m; ++r) if (P.get(r, j)) csyn[(size_t)j * sw + (r >> 6)] |= 1ull << (r & 63);
        for (int c = 0; c < fpk; ++c) if (pr.check.get(c, j)) cfp[(size_t)j * fw + (c >> 6)] |= 1ull << (c & 63);
    }
    struct Key { std::vector<u64> s; bool operator==(const Key& o) const { return s == o.s; } };
    struct H { size_t operator()(const Key& k) const { size_t h = 1469598103934665603ull; for (u64 x : k.s) { h ^= x; h *= 1099511628211ull; } return h; } };
    for (int w = 1; w <= std::min(wmax, n); ++w) {
        const int a = (w + 1) / 2, b = w / 2;
        // all a-subsets: syndrome -> list of (support mask as sorted columns, fingerprint)
        std::unordered_map<Key, std::vector<std::pair<std::vector<int>, std::vector<u64>>>, H> left;
        std::vector<int> pos(a);
        for (int i = 0; i < a; ++i) pos[i] = i;
        bool more = a <= n;
        while (more) {
            Key k{std::vector<u64>(sw, 0)};
            std::vector<u64> f(fw, 0);
            for (int c : pos) { for (int t = 0; t < sw; ++t) k.s[t] ^= csyn[(size_t)c * sw + t]; for (int t = 0; t < fw; ++t) f[t] ^= cfp[(size_t)c * fw + t]; }
            left[k].push_back({pos, f});
            int j = a - 1;
            while (j >= 0 && pos[j] == n - a + j) --j;
            if (j < 0) more = false;
            else { ++pos[j]; for (int t = j + 1; t < a; ++t) pos[t] = pos[t - 1] + 1; }
        }
        std::vector<int> q(b);
        for (int i = 0; i < b; ++i) q[i] = i;
        more = b <= n;
        while (more) {
            Key k{std::vector<u64>(sw, 0)};
            std::vector<u64> f(fw, 0);
            for (int c : q) { for (int t = 0; t < sw; ++t) k.s[t] ^= csyn[(size_t)c * sw + t]; for (int t = 0; t < fw; ++t) f[t] ^= cfp[(size_t)c * fw + t]; }
            auto it = left.find(k);
            if (it != left.end()) {
                for (auto& e : it->second) {
                    bool disjoint = true;
                    for (int c : q) if (std::binary_search(e.first.begin(), e.first.end(), c)) { disjoint = false; break; }
                    if (!disjoint) continue;
                    bool lg = pr.count_all;
                    for (int t = 0; t < fw && !lg; ++t) if (e.second[t] ^ f[t]) lg = true;
                    if (lg) return w;
                }
            }
            if (b == 0) break;
            int j = b - 1;
            while (j >= 0 && q[j] == n - b + j) --j;
            if (j < 0) more = false;
            else { ++q[j]; for (int t = j + 1; t < b; ++t) q[t] = q[t - 1] + 1; }
        }
    }
    return -1;
}

} // namespace orc

// ---------------------------------------------------------------------------------------
// C ABI for the tests (ctypes). `is_threads` is the random-information-set worker count
// (the reference uses std::thread::hardware_concurrency(); 0 here means the same).
extern "C" {

typedef struct { int distance, lower_bound, proven, levels; double seconds; } OracleResult;

static void fill(OracleResult* o, const orc::Result& r) {
    o->distance = r.distance; o->lower_bound = r.lower_bound; o->proven = r.proven;
    o->levels = r.levels; o->seconds = r.seconds;
}
static int hw_threads(int t) { return t > 0 ? t : (int)std::max(1u, std::thread::hardware_concurrency()); }

// Sound level schedule on (1) or off (0, the reference's); returns the previous setting.
int oracle_set_sound(int on) {
    const int old = orc::g_sound ? 1 : 0;
    orc::g_sound = on != 0;
    return old;
}

int oracle_enumerate(int n, int stride, int K, int d, int ng, const uint64_t* gamma,
                     int kcheck, const uint64_t* check, int current_best, int threads) {
    orc::Binom B; B.build(K, K);
    orc::Plan p{n, stride, K, d, ng, gamma, kcheck, check, current_best};
    return orc::enumerate(p, B, threads);
}

// Straight enumeration: for every d-subset (lexicographic) and generator, XOR the rows,
// test the detector, then take the weight (cpp/qsf/backend.hpp:84-100 + cpu_enum_ref.cpp:27-42).
int oracle_enumerate_brute(int n, int stride, int K, int d, int ng, const uint64_t* gamma,
                           int kcheck, const uint64_t* check, int current_best) {
    (void)n;
    orc::Binom B; B.build(K, K);
    orc::Plan p{n, stride, K, d, ng, gamma, kcheck, check, current_best};
    orc::u64 total = B(K, d);
    int best = current_best;
    std::vector<int> pos(d);
    std::vector<orc::u64> cw(stride);
    for (orc::u64 r = 0; r < total; ++r) {
        orc::unrank_lex(B, K, d, r, pos.data());
        for (int g = 0; g < ng; ++g) {
            std::fill(cw.begin(), cw.end(), 0);
            for (int i = 0; i < d; ++i) for (int w = 0; w < stride; ++w) cw[w] ^= gamma[((size_t)g * K + pos[i]) * stride + w];
            if (!orc::logical(p, cw.data())) continue;
            int wt = 0;
            for (int w = 0; w < stride; ++w) wt += orc::pc64(cw[w]);
            best = std::min(best, wt);
        }
    }
    return best;
}

// Collect mode: every codeword of weight in [1, keep] at level d, dups allowed
// (backend_cpu.cpp:49-66). Writes up to `cap` rows into `out`; returns the row count
// (which may exceed cap).
int64_t oracle_collect(int n, int stride, int K, int d, int ng, const uint64_t* gamma,
                       int keep, uint64_t* out, int64_t cap, int threads) {
    orc::Binom B; B.build(K, K);
    orc::Plan p{n, stride, K, d, ng, gamma, 0, nullptr, orc::NONE};
    std::vector<orc::Sink> sinks;
    orc::run_level(p, B, threads, sinks, true, keep);
    int64_t rows = 0;
    for (auto& s : sinks) {
        int64_t r = (int64_t)s.hits.size() / stride;
        for (int64_t i = 0; i < r; ++i, ++rows)
            if (rows < cap) std::memcpy(out + rows * stride, s.hits.data() + i * stride, (size_t)stride * 8);
    }
    return rows;
}

int oracle_classical_distance(const uint8_t* H, int rows, int cols, int threads, int is_threads,
                              OracleResult* out) {
    if (!H || !out) return 1;
    orc::Problem p = orc::classical_problem(orc::dense(H, rows, cols));
    fill(out, orc::bz_single(p, threads, hw_threads(is_threads)));
    return 0;
}

int oracle_css_distance(const uint8_t* Hx, int xr, int xc, const uint8_t* Hz, int zr, int zc,
                        char which, int threads, int is_threads, OracleResult* out) {
    if (!Hx || !Hz || !out) return 1;
    orc::Mat mx = orc::dense(Hx, xr, xc), mz = orc::dense(Hz, zr, zc);
    if (which == 'Z' || which == 'X')
        fill(out, orc::bz_single(orc::css_problem(mx, mz, which), threads, hw_threads(is_threads)));
    else
        fill(out, orc::bz_min_pair(orc::css_problem(mx, mz, 'Z'), orc::css_problem(mx, mz, 'X'),
                                   threads, hw_threads(is_threads)));
    return 0;
}

int oracle_stabilizer_distance(const uint8_t* S, int rows, int cols, char which, int threads,
                               int is_threads, OracleResult* out) {
    if (!S || !out) return 1;
    if (cols % 2) return 2;
    orc::Mat m = orc::dense(S, rows, cols);
    if (orc::is_css(m)) {
        orc::Mat Hx, Hz;
        orc::split(m, Hx, Hz);
        if (which == 'Z' || which == 'X')
            fill(out, orc::bz_single(orc::css_problem(Hx, Hz, which), threads, hw_threads(is_threads)));
        else
            fill(out, orc::bz_min_pair(orc::css_problem(Hx, Hz, 'Z'), orc::css_problem(Hx, Hz, 'X'),
                                       threads, hw_threads(is_threads)));
        return 0;
    }
    orc::Result r = orc::bz_single(orc::symplectic_isometric(m), threads, hw_threads(is_threads));
    if (r.distance > 0) r.distance /= 2;                            // stab.cpp:189-190
    if (r.lower_bound > 0) r.lower_bound /= 2;
    fill(out, r);
    return 0;
}

// Independent certifiers on the same subproblems (kind: 0 classical H; 1 CSS Z; 2 CSS X).
static orc::Problem make_problem(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc) {
    if (kind == 0) return orc::classical_problem(orc::dense(A, ar, ac));
    orc::Mat mx = orc::dense(A, ar, ac), mz = orc::dense(B2, br, bc);
    return orc::css_problem(mx, mz, kind == 1 ? 'Z' : 'X');
}
int oracle_bruteforce_distance(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc) {
    orc::Problem p = make_problem(kind, A, ar, ac, B2, br, bc);
    if (p.gen.rows > 26) return -2;
    return orc::brute_force(p);
}
int oracle_mitm_distance(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc, int wmax) {
    return orc::mitm(make_problem(kind, A, ar, ac, B2, br, bc), wmax);
}

// The information-set generators, ranks and detector of a subproblem, flattened as an
// EnumPlan would carry them (bz.cpp:95-128). Returns the number of generators, or -1 if a
// buffer is too small.
int oracle_side_plan(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc,
                     uint64_t* gamma, int64_t gamma_cap, int* ranks, int ranks_cap, uint64_t* check,
                     int64_t check_cap, int* K, int* kcheck, int* even) {
    orc::Problem p = make_problem(kind, A, ar, ac, B2, br, bc);
    std::vector<orc::Gamma> seq = orc::gamma_sequence(p.gen);
    std::stable_sort(seq.begin(), seq.end(), [](const orc::Gamma& a, const orc::Gamma& b) { return a.rank > b.rank; });
    const int k = p.gen.rows, st = orc::wfor(p.n), ng = (int)seq.size();
    *K = k; *kcheck = p.count_all ? 0 : p.check.rows; *even = p.even;
    if ((int64_t)ng * k * st > gamma_cap || ng > ranks_cap || (int64_t)*kcheck * st > check_cap) return -1;
    for (int g = 0; g < ng; ++g) {
        ranks[g] = seq[g].rank;
        std::memcpy(gamma + (size_t)g * k * st, seq[g].g.w.data(), (size_t)k * st * 8);
    }
    if (*kcheck) std::memcpy(check, p.check.w.data(), (size_t)*kcheck * st * 8);
    return ng;
}

// Upper bound from the problem's seed and 256 random information sets (bz.cpp:115-128).
int oracle_upper_seed(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc, int is_threads) {
    orc::Problem p = make_problem(kind, A, ar, ac, B2, br, bc);
    int inner = p.seed_upper > 0 ? p.seed_upper : orc::NONE;
    return orc::is_seed(p, inner, 256, 0x9e3779b9u, hw_threads(is_threads));
}

// The engine's random information-set trials. The per-trial work restates bz.cpp:41-51
// (systematic form over a random column order, lightest nonzero logical row); the column
// order comes from the engine's own stream in place of the reference's mt19937 +
// std::shuffle (bz.cpp:34-40): trial t shuffles 0..n-1 by Fisher–Yates, swapping i with
// floor(hi32(splitmix64()) (i + 1) / 2^32) for i = n-1 .. 1, the splitmix64 state seeded with
// seed ^ (0xd1b54a32d192ed03 * (t + 1)). When rows_out / perm_out are given (one trial,
// trials == 1 and first == t), they receive the systematic form (K x stride) and the order.
static uint64_t splitmix_step(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int oracle_is_bound_trials(const uint64_t* gen, int K, int n, const uint64_t* check, int kcheck, int every_nonzero,
                           int first, int trials, uint64_t seed, int init, uint64_t* rows_out, int* perm_out) {
    const int st = orc::wfor(n);
    if (K == 0 || (!every_nonzero && kcheck == 0)) return init;
    int best = init;
    std::vector<int> cols(n);
    for (int t = first; t < first + trials; ++t) {
        uint64_t s = seed ^ (0xd1b54a32d192ed03ull * (uint64_t)(t + 1));
        std::iota(cols.begin(), cols.end(), 0);
        for (int i = n - 1; i > 0; --i) std::swap(cols[i], cols[(size_t)(((splitmix_step(s) >> 32) * (uint64_t)(i + 1)) >> 32)]);
        orc::Mat M(K, n);
        std::memcpy(M.w.data(), gen, (size_t)K * st * 8);
        orc::rref_cols(M, cols);
        if (rows_out) std::memcpy(rows_out, M.w.data(), (size_t)K * st * 8);
        if (perm_out) std::memcpy(perm_out, cols.data(), (size_t)n * 4);
        for (int r = 0; r < K; ++r) {
            const int w = M.weight(r);
            if (w == 0 || w >= best) continue;
            bool lg = every_nonzero != 0;
            for (int c = 0; c < kcheck && !lg; ++c) lg = orc::dot(check + (size_t)c * st, M.row(r), st);
            if (lg) best = w;
        }
    }
    return best;
}

// The problem's generator (K x stride) and detector rows (kcheck x stride), css.cpp:78-120.
int oracle_problem_gen(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc,
                       uint64_t* gen, int64_t gen_cap, uint64_t* check, int64_t check_cap, int* K, int* kcheck) {
    orc::Problem p = make_problem(kind, A, ar, ac, B2, br, bc);
    *K = p.gen.rows;
    *kcheck = p.count_all ? 0 : p.check.rows;
    if ((int64_t)p.gen.w.size() > gen_cap || (int64_t)((size_t)*kcheck * p.gen.stride) > check_cap) return -1;
    std::memcpy(gen, p.gen.w.data(), p.gen.w.size() * 8);
    if (*kcheck) std::memcpy(check, p.check.w.data(), (size_t)*kcheck * p.gen.stride * 8);
    return 0;
}

// The greedy peel alone, in peel order (css.cpp:177-203): returns the number of sets;
// stores the first min(nsets, cap) generators (K x stride each) and ranks.
int oracle_greedy_sets(const uint64_t* gen, int K, int n, uint64_t* out, int cap, int* ranks) {
    orc::Mat G(K, n);
    std::memcpy(G.w.data(), gen, G.w.size() * 8);
    std::vector<orc::Gamma> seq = orc::gammas_greedy(G);
    for (int i = 0; i < (int)seq.size() && i < cap; ++i) {
        std::memcpy(out + (size_t)i * G.w.size(), seq[i].g.w.data(), G.w.size() * 8);
        ranks[i] = seq[i].rank;
    }
    return (int)seq.size();
}

// gen made systematic over each part's columns in order (css.cpp:358-361).
void oracle_part_sets(const uint64_t* gen, int K, int n, const int* cols, const int* off, int nparts,
                      uint64_t* out, int* ranks) {
    orc::Mat G(K, n);
    std::memcpy(G.w.data(), gen, G.w.size() * 8);
    for (int p = 0; p < nparts; ++p) {
        orc::Mat W = G;
        std::vector<int> c(cols + off[p], cols + off[p + 1]);
        ranks[p] = (int)orc::rref_cols(W, c).size();
        std::memcpy(out + (size_t)p * G.w.size(), W.w.data(), G.w.size() * 8);
    }
}

// The matroid partition of gen's columns (css.cpp:274-328), classes sorted by size as
// css.cpp:354-355 sorts them, empty classes dropped: part_of_col[c] = class of column c.
int oracle_matroid_parts(const uint64_t* gen, int K, int n, int* part_of_col) {
    orc::Mat G(K, n);
    std::memcpy(G.w.data(), gen, G.w.size() * 8);
    std::vector<std::vector<int>> part;
    if (!orc::matroid_partition_public(G, part)) return -1;
    std::sort(part.begin(), part.end(), [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() > b.size(); });
    std::fill(part_of_col, part_of_col + n, -1);
    int p = 0;
    for (auto& c : part) {
        if (c.empty()) continue;
        for (int j : c) part_of_col[j] = p;
        ++p;
    }
    return p;
}

// Problem dimensions and information-set ranks, for the host-logic tests.
int oracle_problem_info(int kind, const uint8_t* A, int ar, int ac, const uint8_t* B2, int br, int bc,
                        int* K, int* kcheck, int* even, int* seed_upper, int* ranks, int max_ranks) {
    orc::Problem p = make_problem(kind, A, ar, ac, B2, br, bc);
    *K = p.gen.rows; *kcheck = p.count_all ? 0 : p.check.rows; *even = p.even; *seed_upper = p.seed_upper;
    std::vector<orc::Gamma> seq = orc::gamma_sequence(p.gen);
    std::stable_sort(seq.begin(), seq.end(), [](const orc::Gamma& a, const orc::Gamma& b) { return a.rank > b.rank; });
    int ng = (int)seq.size();
    for (int g = 0; g < ng && g < max_ranks; ++g) ranks[g] = seq[g].rank;
    return ng;
}

} // extern "C"
