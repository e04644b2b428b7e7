#include "driver.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <cstdio>

#include "../../include/qsf_hip.h"

namespace qsa {

namespace {

using clk = std::chrono::steady_clock;

u64 fresh_key() {
    static std::atomic<u64> k{0x51f0000000000000ull};
    return ++k;
}

void print_elapsed(double secs) {
    if (secs < 1.0) std::fprintf(stderr, "Elapsed:[%dms]\n", (int)(secs * 1000.0 + 0.5));
    else std::fprintf(stderr, "Elapsed:[%.3fs]\n", secs);
}
void say_bound(int lb, double secs, const char* tag) {
    if (*tag) std::fprintf(stderr, "%s-distance bound: >%d\n", tag, lb);
    else std::fprintf(stderr, "Distance bound: >%d\n", lb);
    print_elapsed(secs);
}
void say_final(int d, double secs, const char* tag) {
    if (*tag) std::fprintf(stderr, "%s-distance: =%d\n", tag, d);
    else std::fprintf(stderr, "Distance: =%d\n", d);
    print_elapsed(secs);
}

// Information sets of gen as the level loop uses them: systematic generators computed on
// the device, ordered by rank, descending, stable (bz.cpp:96-97).
struct Family {
    int ng = 0;
    std::vector<u64> gamma;   // ng x K x stride
    std::vector<int> ranks;
    int full(int K) const { return (int)std::count(ranks.begin(), ranks.end(), K); }
};

Family device_sets(const BitMat& gen, const std::vector<std::vector<int>>* parts) {
    const int K = gen.rows, n = gen.cols, st = words(n);
    std::vector<int> cols, off;
    qsf_elim_job q{};
    q.gen = gen.data.data();
    q.K = K;
    q.n = n;
    q.init = kNoWeight;
    if (parts) {
        off.push_back(0);
        for (auto& p : *parts) {
            cols.insert(cols.end(), p.begin(), p.end());
            off.push_back((int)cols.size());
        }
        q.part_cols = cols.data();
        q.part_off = off.data();
        q.nparts = (int)parts->size();
    }
    int cap = parts ? std::max(1, (int)parts->size()) : std::min(n + 1, 16);
    std::vector<u64> sets;
    std::vector<int> ranks;
    for (;;) {   // a greedy peel longer than the first guess reruns with room for all sets
        sets.assign((size_t)cap * K * st + 1, 0);
        ranks.assign(cap, 0);
        q.sets = sets.data();
        q.set_cap = cap;
        q.ranks = ranks.data();
        const int rc = qsf_hip_eliminate(&q, 1, 0, 0, QSF_ELIM_SETS);
        if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
        if (q.nsets <= cap) break;
        cap = q.nsets;
    }
    std::vector<int> idx(q.nsets);
    for (int i = 0; i < q.nsets; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ranks[a] > ranks[b]; });
    Family f;
    f.ng = q.nsets;
    f.gamma.resize((size_t)f.ng * K * st);
    for (int i = 0; i < f.ng; ++i) {
        std::memcpy(f.gamma.data() + (size_t)i * K * st, sets.data() + (size_t)idx[i] * K * st, (size_t)K * st * 8);
        f.ranks.push_back(ranks[idx[i]]);
    }
    return f;
}

// Greedy peel, replaced by the matroid partition's sets when n <= 200 and the partition
// yields more full-rank sets (cpp/distfind/css.cpp:336-366).
Family information_sets(const BitMat& gen) {
    Family greedy = device_sets(gen, nullptr);
    const int K = gen.rows, n = gen.cols;
    if (K == 0) return greedy;
    const int gfull = greedy.full(K);
    if (gfull >= nonzero_columns(gen) / K || n > 200) return greedy;
    std::vector<std::vector<int>> parts;
    if (!matroid_parts(gen, parts) || parts.empty()) return greedy;
    Family seq = device_sets(gen, &parts);
    return seq.full(K) <= gfull ? greedy : seq;
}

// One subproblem's level-stepping state. Owns the flat generator / detector arrays the
// device caches under `key` for the whole solve.
struct Side {
    const char* tag = "";
    bool has = false, even = false;
    int n = 0, K = 0, stride = 0, ng = 0, kcheck = 0;
    std::vector<u64> gamma, check;
    std::vector<int> ranks;
    u64 key = 0;
    int inner = kNoWeight, outer = 0, level = 0;
    bool proven = false, resolved = false;
    const BzOptions* opt = nullptr;

    void setup(const Problem& p, const BzOptions& o, const char* t) {
        opt = &o;
        tag = t;
        n = p.n;
        K = p.gen.rows;
        stride = words(n);
        even = p.even;
        has = !(K == 0 || (!p.every_nonzero && p.detector.rows == 0));
        if (!has) { resolved = true; return; }
        Family fam;
        if (o.parts && !o.parts->empty()) {
            std::vector<std::vector<int>> parts;
            for (auto& c : *o.parts)
                if (!c.empty()) parts.push_back(c);
            fam = device_sets(p.gen, &parts);
        } else {
            fam = information_sets(p.gen);
        }
        ng = fam.ng;
        gamma = std::move(fam.gamma);
        ranks = std::move(fam.ranks);
        kcheck = p.every_nonzero ? 0 : p.detector.rows;
        check.assign((size_t)std::max(1, kcheck) * stride, 0);
        if (kcheck) std::memcpy(check.data(), p.detector.data.data(), (size_t)kcheck * stride * 8);
        key = fresh_key();
        inner = p.seed_upper > 0 ? p.seed_upper : kNoWeight;
        int rc = qsf_hip_is_bound(p.gen.data.data(), K, n, check.data(), kcheck, p.every_nonzero ? 1 : 0,
                                  o.is_trials, o.seed, inner, &inner);
        if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
    }

    int active(int d) const {
        int a = 0;
        while (a < ng && (d + 1) - (K - ranks[a]) > 0) ++a;
        return a;
    }

    void step(int d) {
        const int act = active(d);
        int found = kNoWeight;
        qsf_hip_stats st{};
        auto t0 = clk::now();
        if (act > 0) {
            qsf_enum_plan pl{};
            pl.n = n; pl.stride = stride; pl.K = K; pl.d = d;
            pl.num_gamma = act; pl.num_gamma_total = ng;
            pl.gamma = gamma.data(); pl.kcheck = kcheck; pl.check = check.data();
            pl.current_best = inner; pl.buffers_key = key;
            // every codeword not enumerated before this level weighs >= outer (the BZ bound
            // after level d-1), so a logical that light ends the level early
            int rc = qsf_hip_enumerate_floor(&pl, outer, &found);
            if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
            qsf_hip_last_stats(&st);
        }
        double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        inner = std::min(inner, found);
        ++level;
        outer = 0;
        for (int g = 0; g < ng; ++g) outer += std::max(0, (d + 1) - (K - ranks[g]));
        if (even) outer += outer & 1;
        if (opt->profile)
            std::fprintf(stderr, "[prof hip] d=%d K=%d ng=%d/%d stride=%d work=%llu kernel=%d fd=%d/%d medium=%llu slow=%llu -> %.3f ms (device %.3f ms)\n",
                         d, K, act, ng, stride, (unsigned long long)st.codewords, st.kernel, st.fd, st.xd,
                         (unsigned long long)st.medium_hits, (unsigned long long)st.slow_hits, ms, st.kernel_ms);
        if (opt->trace) opt->trace->push_back(LevelTrace{*tag ? tag[0] : 'S', d, act, ng, K, inner, outer, ms, st});
    }
};

}  // namespace

BzResult bz_distance(const Problem& p, const BzOptions& opt) {
    auto t0 = clk::now();
    BzResult r;
    Side S;
    S.setup(p, opt, "");
    if (S.has) {
        for (int d = 1; d <= S.K; ++d) {
            auto tl = clk::now();
            S.step(d);
            if (S.inner <= S.outer) {
                S.proven = true;
                if (opt.verbose) say_final(S.inner, std::chrono::duration<double>(clk::now() - t0).count(), "");
                break;
            }
            if (opt.verbose) say_bound(std::max(0, S.outer - 1), std::chrono::duration<double>(clk::now() - tl).count(), "");
        }
        r.distance = S.inner >= kNoWeight ? -1 : S.inner;
        r.levels = S.level;
        r.lower_bound = S.outer;
        r.proven = S.proven;
    }
    r.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return r;
}

BzResult bz_min_interleaved(const Problem& pz, const Problem& px, const BzOptions& opt) {
    auto t0 = clk::now();
    Side Z, X;
    Z.setup(pz, opt, "Z");
    X.setup(px, opt, "X");
    Side* sides[2] = {&Z, &X};
    int best = kNoWeight;
    if (Z.has) best = std::min(best, Z.inner);
    if (X.has) best = std::min(best, X.inner);
    const int maxw = std::max(Z.has ? Z.K : 0, X.has ? X.K : 0);
    for (int d = 1; d <= maxw; ++d) {
        for (Side* S : sides) {
            if (S->resolved || d > S->K) continue;
            if (S->outer >= best) { S->resolved = true; continue; }
            auto tl = clk::now();
            S->step(d);
            best = std::min(best, S->inner);
            if (S->inner <= S->outer) {
                S->proven = S->resolved = true;
                if (opt.verbose) say_final(S->inner, std::chrono::duration<double>(clk::now() - t0).count(), S->tag);
            } else {
                if (opt.verbose) say_bound(std::max(0, S->outer - 1), std::chrono::duration<double>(clk::now() - tl).count(), S->tag);
                if (S->outer >= best) S->resolved = true;
            }
        }
        if (Z.resolved && X.resolved) break;
    }
    auto v = [](const Side& S) { return S.has ? S.inner : kNoWeight; };
    const int dist = std::min(v(Z), v(X));
    BzResult r;
    r.distance = dist >= kNoWeight ? -1 : dist;
    const int lb = std::min(Z.has ? Z.outer : kNoWeight, X.has ? X.outer : kNoWeight);
    r.lower_bound = lb >= kNoWeight ? 0 : lb;
    r.levels = std::max(Z.level, X.level);
    r.proven = Z.resolved && X.resolved && dist < kNoWeight;
    r.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return r;
}

}  // namespace qsa
