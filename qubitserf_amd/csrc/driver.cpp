#include "driver.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <exception>
#include <memory>
#include <thread>
#include <unordered_map>

#include "../../include/qsf_hip.h"
#include "pool.hpp"

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

// Information sets of a generator as the level loop uses them: systematic generators
// computed on the device, ordered by rank, descending, stable (bz.cpp:96-97).
struct Family {
    int ng = 0;
    std::vector<u64> gamma;   // ng x K x stride
    std::vector<int> ranks;
    int full(int K) const { return (int)std::count(ranks.begin(), ranks.end(), K); }
};

// One information-set request: greedy peel of gen (over `order` when given), or gen made
// systematic over each part of `parts`.
struct SetRequest {
    const BitMat* gen;
    const std::vector<int>* order = nullptr;
    const std::vector<std::vector<int>>* parts = nullptr;
    bool ranks_only = false;   // out.gamma stays empty
    Family out;
    // the random information-set bound in the same device call (device_sets with trials > 0)
    const u64* check = nullptr;
    int kcheck = 0, every_nonzero = 0, init = kNoWeight;
    int best = kNoWeight;      // out
};

// All requests in one device call (a request whose peel outgrows the first capacity guess
// is redone with room for every set). trials > 0: the first call also computes every
// request's random information-set bound (check, kcheck, every_nonzero, init -> best).
void device_sets(std::vector<SetRequest*>& reqs, int trials = 0, u64 seed = 0) {
    std::vector<SetRequest*> todo = reqs;
    std::vector<int> caps(reqs.size());
    std::unordered_map<SetRequest*, int> cap;
    for (SetRequest* r : reqs)
        cap[r] = r->parts ? std::max(1, (int)r->parts->size()) : std::min(r->gen->cols + 1, 16);
    while (!todo.empty()) {
        const size_t m = todo.size();
        std::vector<qsf_elim_job> jobs(m);
        std::vector<std::vector<u64>> sets(m);
        std::vector<std::vector<int>> ranks(m), cols(m), off(m);
        for (size_t k = 0; k < m; ++k) {
            SetRequest& r = *todo[k];
            const BitMat& g = *r.gen;
            qsf_elim_job& q = jobs[k];
            q.gen = g.data.data();
            q.K = g.rows;
            q.n = g.cols;
            q.init = trials > 0 ? r.init : kNoWeight;
            if (trials > 0) {
                q.check = r.check;
                q.kcheck = r.kcheck;
                q.every_nonzero = r.every_nonzero;
            }
            q.order = r.order ? r.order->data() : nullptr;
            if (r.parts) {
                off[k].push_back(0);
                for (auto& pc : *r.parts) {
                    cols[k].insert(cols[k].end(), pc.begin(), pc.end());
                    off[k].push_back((int)cols[k].size());
                }
                q.part_cols = cols[k].data();
                q.part_off = off[k].data();
                q.nparts = (int)r.parts->size();
            }
            const int c = cap[todo[k]];
            if (!r.ranks_only) sets[k].assign((size_t)c * g.rows * g.ld + 1, 0);
            ranks[k].assign(c, 0);
            q.sets = r.ranks_only ? nullptr : sets[k].data();
            q.set_cap = c;
            q.ranks = ranks[k].data();
        }
        const int rc = qsf_hip_eliminate(jobs.data(), (int)m, trials, seed,
                                         trials > 0 ? QSF_ELIM_SETS | QSF_ELIM_BOUND : QSF_ELIM_SETS);
        if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
        std::vector<SetRequest*> again;
        for (size_t k = 0; k < m; ++k) {
            SetRequest& r = *todo[k];
            const qsf_elim_job& q = jobs[k];
            if (trials > 0) r.best = q.best;
            if (q.nsets > q.set_cap) {
                cap[todo[k]] = q.nsets;
                again.push_back(todo[k]);
                continue;
            }
            const size_t w = (size_t)q.K * r.gen->ld;
            std::vector<int> idx(q.nsets);
            for (int t = 0; t < q.nsets; ++t) idx[t] = t;
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ranks[k][a] > ranks[k][b]; });
            r.out = Family{};
            r.out.ng = q.nsets;
            if (!r.ranks_only) r.out.gamma.resize((size_t)q.nsets * w);
            for (int t = 0; t < q.nsets; ++t) {
                if (!r.ranks_only)
                    std::memcpy(r.out.gamma.data() + (size_t)t * w, sets[k].data() + (size_t)idx[t] * w, w * 8);
                r.out.ranks.push_back(ranks[k][idx[t]]);
            }
        }
        todo.swap(again);
        trials = 0;   // a redone peel needs no second bound
    }
}

// Sound schedule: the first level D at which the bound over every generator reaches
// `upper` (K if never), and the generators that contribute by then (K - r <= D; ranks
// sorted descending, so a prefix). Those are enumerated from level 1.
int sound_prefix(int K, const std::vector<int>& ranks, bool even, int upper) {
    int D = K;
    for (int d = 1; d <= K; ++d) {
        int o = 0;
        for (int r : ranks) o += std::max(0, (d + 1) - (K - r));
        if (even) o += o & 1;
        if (o >= upper) { D = d; break; }
    }
    int a = 0;
    while (a < (int)ranks.size() && K - ranks[a] <= D) ++a;
    return a;
}

// One subproblem's level-stepping state. Owns the flat generator / detector arrays the
// device caches under `key` for the whole solve.
struct Side {
    const char* tag = "";
    const Problem* prob = nullptr;
    bool has = false, even = false;
    int n = 0, K = 0, stride = 0, ng = 0, kcheck = 0;
    std::vector<u64> gamma, check;
    std::vector<int> ranks;
    u64 key = 0;
    int inner = kNoWeight, outer = 0, level = 0;
    bool proven = false, resolved = false;
    const BzOptions* opt = nullptr;
    u64 work = 0;          // codewords enumerated so far

    // Host part of the set-up; the information sets and the upper bound come from
    // setup_sides, on the device, for many sides at once.
    void prepare(const Problem& p, const BzOptions& o, const char* t) {
        prob = &p;
        opt = &o;
        tag = t;
        n = p.n;
        K = p.gen.rows;
        stride = words(n);
        even = p.even;
        has = !(K == 0 || (!p.every_nonzero && p.detector.rows == 0));
        if (!has) { resolved = true; return; }
        kcheck = p.every_nonzero ? 0 : p.detector.rows;
        check.assign((size_t)std::max(1, kcheck) * stride, 0);
        if (kcheck) std::memcpy(check.data(), p.detector.data.data(), (size_t)kcheck * stride * 8);
        key = fresh_key();
        inner = p.seed_upper > 0 ? p.seed_upper : kNoWeight;
    }

    mutable int nsound = -1;   // sound schedule: generators enumerated from level 1 (a prefix)
    int active(int d) const {
        if (opt->sound) {
            if (nsound < 0) nsound = sound_prefix(K, ranks, even, inner);
            return nsound;
        }
        int a = 0;
        while (a < ng && (d + 1) - (K - ranks[a]) > 0) ++a;
        return a;
    }

    // Level d's plan; false when no generator is active at d (nothing to enumerate).
    bool plan(int d, qsf_enum_plan& pl) const {
        const int act = active(d);
        if (act == 0) return false;
        pl = qsf_enum_plan{};
        pl.n = n; pl.stride = stride; pl.K = K; pl.d = d;
        pl.num_gamma = act; pl.num_gamma_total = ng;
        pl.gamma = gamma.data(); pl.kcheck = kcheck; pl.check = check.data();
        pl.current_best = inner; pl.buffers_key = key;
        return true;
    }

    // The bound after level d (0 before level 1), as apply() computes it.
    int outer_at(int d) const {
        if (d <= 0) return 0;
        const int counted = opt->sound ? active(d) : ng;
        int o = 0;
        for (int g = 0; g < counted; ++g) o += std::max(0, (d + 1) - (K - ranks[g]));
        if (even) o += o & 1;
        return o;
    }

    u64 level_work(int d) const {
        const int act = active(d);
        long double c = 1;
        for (int i = 0; i < d; ++i) c = c * (K - i) / (i + 1);
        const long double w = c * act;
        return w > 1.8e19L ? ~0ull : (u64)(w + 0.5L);
    }

    // Bounds after level d, given what its enumeration found.
    void apply(int d, int found, double ms, const qsf_hip_stats& st) {
        const int act = active(d);
        inner = std::min(inner, found);
        ++level;
        work += st.codewords;
        outer = 0;
        const int counted = opt->sound ? active(d) : ng;   // sound: only the generators enumerated from level 1
        for (int g = 0; g < counted; ++g) outer += std::max(0, (d + 1) - (K - ranks[g]));
        if (even) outer += outer & 1;
        if (opt->profile)
            std::fprintf(stderr, "[prof hip] d=%d K=%d ng=%d/%d stride=%d work=%llu kernel=%d fd=%d/%d medium=%llu slow=%llu -> %.3f ms (device %.3f ms)\n",
                         d, K, act, ng, stride, (unsigned long long)st.codewords, st.kernel, st.fd, st.xd,
                         (unsigned long long)st.medium_hits, (unsigned long long)st.slow_hits, ms, st.kernel_ms);
        if (opt->trace) opt->trace->push_back(LevelTrace{*tag ? tag[0] : 'S', d, act, ng, K, inner, outer, ms, st});
    }

    void step(int d) {
        int found = kNoWeight;
        qsf_hip_stats st{};
        auto t0 = clk::now();
        qsf_enum_plan pl;
        if (plan(d, pl)) {
            // every codeword not enumerated before this level weighs >= outer (the BZ bound
            // after level d-1), so a logical that light ends the level early
            int rc = qsf_hip_enumerate_floor(&pl, outer, &found);
            if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
            qsf_hip_last_stats(&st);
        }
        apply(d, found, std::chrono::duration<double, std::milli>(clk::now() - t0).count(), st);
    }
};

// Codewords BZ enumerates before its lower bound reaches `upper` (or level K), for a
// family of information-set ranks: the sum over levels d of C(K, d) x (sets active at d),
// the bound after level d being sum_g max(0, d + 1 - (K - r_g)), rounded up to even for an
// even code (bz.cpp:142-162). A proposal's candidates are compared by it.
long double predicted_work(int K, const std::vector<int>& ranks, bool even, int upper) {
    long double work = 0, c = 1;
    for (int d = 1; d <= K; ++d) {
        c = c * (K - d + 1) / d;
        int act = 0, outer = 0;
        for (int r : ranks) {
            act += (d + 1) - (K - r) > 0;
            outer += std::max(0, (d + 1) - (K - r));
        }
        if (even) outer += outer & 1;
        work += c * act;
        if (upper <= outer) break;
    }
    return work;
}

// Candidate column orders from a proposal's column scores: descending score + Gumbel(0, 1)
// noise, i.e. samples of the Plackett–Luce distribution the scores define. Candidate t of
// a problem draws its noise from splitmix64 seeded with seed ^ (0x9e3779b97f4a7c15 (t+1)).
std::vector<std::vector<int>> proposal_orders(const std::vector<float>& scores, int count, u64 seed) {
    const int n = (int)scores.size();
    std::vector<std::vector<int>> out;
    for (int t = 0; t < count; ++t) {
        u64 s = seed ^ (0x9e3779b97f4a7c15ull * (u64)(t + 1));
        std::vector<double> key(n);
        for (int j = 0; j < n; ++j) {
            s += 0x9e3779b97f4a7c15ull;
            u64 z = s;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            const double u = ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);   // (0, 1)
            key[j] = scores[j] - std::log(-std::log(u));
        }
        std::vector<int> o(n);
        for (int j = 0; j < n; ++j) o[j] = j;
        std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return key[a] > key[b]; });
        out.push_back(std::move(o));
    }
    return out;
}

// Information sets (greedy peel over the given or proposed order, then the matroid
// partition's sets when n <= 200 and they include more full-rank sets, css.cpp:336-366)
// and the random information-set upper bound of every side, in batched device calls.
// A side with a proposal gets its candidates' peels too; the one with the least predicted
// work is kept, the natural order winning ties.
void setup_sides(std::vector<Side*>& sides, const std::vector<const std::vector<int>*>& orders,
                 const std::vector<const Proposal*>& proposals = {}) {
    std::vector<Side*> live;
    std::vector<qsf_elim_job> jobs;
    for (Side* S : sides) {
        if (!S->has) continue;
        qsf_elim_job q{};
        q.gen = S->prob->gen.data.data();
        q.K = S->K;
        q.n = S->n;
        q.check = S->check.data();
        q.kcheck = S->kcheck;
        q.every_nonzero = S->prob->every_nonzero ? 1 : 0;
        q.init = S->inner;
        jobs.push_back(q);
        live.push_back(S);
    }
    if (live.empty()) return;
    const BzOptions& o = *live[0]->opt;
    // without proposal candidates the bound rides along with the peels (one device call);
    // with them it has a call of its own, so that the candidates' peels run no trials
    bool fuse = true;
    for (size_t i = 0; i < sides.size() && i < proposals.size(); ++i)
        if (sides[i]->has && proposals[i] && proposals[i]->candidates > 0 && (int)proposals[i]->scores.size() == sides[i]->n)
            fuse = false;
    auto ta = clk::now();
    if (!fuse) {
        const int rc = qsf_hip_eliminate(jobs.data(), (int)jobs.size(), o.is_trials, o.seed, QSF_ELIM_BOUND);
        if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
        for (size_t k = 0; k < live.size(); ++k) live[k]->inner = jobs[k].best;
    }
    auto tb = clk::now();
    const double t_bound = std::chrono::duration<double, std::milli>(tb - ta).count();

    // peels: the given order (or 0..n-1), plus every proposal candidate
    const size_t ns = sides.size();
    std::vector<SetRequest> greedy(ns);
    std::vector<std::vector<std::vector<int>>> cand(ns);
    std::vector<std::vector<SetRequest>> cand_req(ns);
    std::vector<SetRequest*> reqs;
    {   // candidate orders of every side with a proposal, on host threads (sorts of n keys)
        std::vector<size_t> with;
        for (size_t i = 0; i < ns; ++i) {
            const Proposal* pr = proposals.empty() ? nullptr : proposals[i];
            if (sides[i]->has && pr && pr->candidates > 0 && (int)pr->scores.size() == sides[i]->n) with.push_back(i);
        }
        HostPool::get().run(with.size(), host_workers(with.size(), 16), [&](size_t k) {
            const Proposal* pr = proposals[with[k]];
            cand[with[k]] = proposal_orders(pr->scores, pr->candidates, pr->seed);
        });
    }
    for (size_t i = 0; i < ns; ++i) {
        if (!sides[i]->has) continue;
        Side& S = *sides[i];
        greedy[i].gen = &S.prob->gen;
        greedy[i].order = orders.empty() ? nullptr : orders[i];
        greedy[i].check = S.check.data();
        greedy[i].kcheck = S.kcheck;
        greedy[i].every_nonzero = S.prob->every_nonzero ? 1 : 0;
        greedy[i].init = S.inner;
        reqs.push_back(&greedy[i]);
        const Proposal* pr = proposals.empty() ? nullptr : proposals[i];
        if (!pr || pr->candidates <= 0 || (int)pr->scores.size() != sides[i]->n) continue;
        cand_req[i].resize(cand[i].size());
        for (size_t t = 0; t < cand[i].size(); ++t) {
            cand_req[i][t].gen = &sides[i]->prob->gen;
            cand_req[i][t].order = &cand[i][t];
            cand_req[i][t].ranks_only = true;
            reqs.push_back(&cand_req[i][t]);
        }
    }
    device_sets(reqs, fuse ? o.is_trials : 0, o.seed);
    if (fuse && o.is_trials > 0)
        for (size_t i = 0; i < ns; ++i)
            if (sides[i]->has) sides[i]->inner = greedy[i].best;
    auto tm = clk::now();
    // candidates were peeled for their ranks only; the winner's generators are peeled again
    std::vector<SetRequest*> winners;
    for (size_t i = 0; i < ns; ++i) {
        Side& S = *sides[i];
        if (!S.has || cand_req[i].empty()) continue;
        long double best = predicted_work(S.K, greedy[i].out.ranks, S.even, S.inner);
        int win = -1;
        for (size_t t = 0; t < cand_req[i].size(); ++t) {
            const long double w = predicted_work(S.K, cand_req[i][t].out.ranks, S.even, S.inner);
            if (w < best) { best = w; win = (int)t; }
        }
        if (win >= 0) {
            greedy[i].order = &cand[i][win];
            winners.push_back(&greedy[i]);
        }
    }
    if (!winners.empty()) device_sets(winners);
    const double t_sets = std::chrono::duration<double, std::milli>(clk::now() - tb).count();
    // matroid partitions (host: an augmenting-path search) of the sides whose peel is short
    // of full-rank sets, several sides at a time
    std::vector<std::vector<std::vector<int>>> parts(ns);
    std::vector<SetRequest> part_req(ns);
    std::vector<SetRequest*> preqs;
    std::vector<size_t> want;
    for (size_t i = 0; i < ns; ++i) {
        Side& S = *sides[i];
        if (!S.has) continue;
        const BitMat& gen = S.prob->gen;
        const int gfull = greedy[i].out.full(S.K);
        if (gfull >= nonzero_columns(gen) / S.K || gen.cols > 200) continue;
        want.push_back(i);
    }
    std::vector<char> ok(ns, 0);
    HostPool::get().run(want.size(), host_workers(want.size(), 2), [&](size_t k) {
        const size_t i = want[k];
        ok[i] = matroid_parts(sides[i]->prob->gen, parts[i]) && !parts[i].empty();
    });
    for (size_t i : want) {
        if (!ok[i]) continue;
        part_req[i].gen = &sides[i]->prob->gen;
        part_req[i].parts = &parts[i];
        preqs.push_back(&part_req[i]);
    }
    const double t_matroid = std::chrono::duration<double, std::milli>(clk::now() - tm).count();
    if (!preqs.empty()) device_sets(preqs);
    if (o.profile)
        std::fprintf(stderr, "[prof setup] sides=%zu bound %.3f ms, peels %.3f ms, matroid %.3f ms (%zu partitions)\n",
                     live.size(), t_bound, t_sets, t_matroid, preqs.size());
    for (size_t i = 0; i < ns; ++i) {
        Side& S = *sides[i];
        if (!S.has) continue;
        Family& f = (part_req[i].gen && part_req[i].out.full(S.K) > greedy[i].out.full(S.K)) ? part_req[i].out
                                                                                              : greedy[i].out;
        S.ng = f.ng;
        S.gamma = std::move(f.gamma);
        S.ranks = std::move(f.ranks);
    }
}

}  // namespace

BzResult bz_distance(const Problem& p, const BzOptions& opt) {
    auto t0 = clk::now();
    BzResult r;
    Side S;
    S.prepare(p, opt, "");
    std::vector<Side*> one{&S};
    setup_sides(one, {opt.order});
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
    Z.prepare(pz, opt, "Z");
    X.prepare(px, opt, "X");
    std::vector<Side*> two{&Z, &X};
    setup_sides(two, {});
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

namespace {

// One problem of a batch, stepped exactly as bz_distance (single side) or
// bz_min_interleaved (pair) would step it, one level per round.
struct Stepper {
    Side A, B;
    bool pair = false;
    int d = 1, turn = 0, best = kNoWeight, maxw = 0;
    bool done = false, capped = false;
    u64 cap = 0;           // this problem's own work cap (0: the batch's)
    Side* cur = nullptr;

    void init() {
        if (!pair) {
            done = !A.has;
            maxw = A.K;
            return;
        }
        if (A.has) best = std::min(best, A.inner);
        if (B.has) best = std::min(best, B.inner);
        maxw = std::max(A.has ? A.K : 0, B.has ? B.K : 0);
        done = maxw < 1;
    }
    void advance() {   // pair: the other side at this level, or the next level
        if (++turn == 2) {
            turn = 0;
            if (A.resolved && B.resolved) { done = true; return; }
            ++d;
        }
    }
    // The side to step next (cur), at level d; done when the solve is over.
    void next() {
        cur = nullptr;
        if (done) return;
        if (!pair) {
            if (d > A.K) done = true;
            else cur = &A;
            return;
        }
        for (;;) {
            if (d > maxw) { done = true; return; }
            Side* S = turn == 0 ? &A : &B;
            if (!(S->resolved || d > S->K)) {
                if (S->outer >= best) S->resolved = true;
                else { cur = S; return; }
            }
            advance();
            if (done) return;
        }
    }
    void after(int found, double ms, const qsf_hip_stats& st) {
        Side* S = cur;
        S->apply(d, found, ms, st);
        if (!pair) {
            if (S->inner <= S->outer) { S->proven = true; done = true; }
            else ++d;
            return;
        }
        best = std::min(best, S->inner);
        if (S->inner <= S->outer) S->proven = S->resolved = true;
        else if (S->outer >= best) S->resolved = true;
        advance();
    }
    BzResult result() const {
        BzResult r;
        if (!pair) {
            if (A.has) {
                r.distance = A.inner >= kNoWeight ? -1 : A.inner;
                r.levels = A.level;
                r.lower_bound = A.outer;
                r.proven = A.proven && !capped;
            }
            return r;
        }
        auto v = [](const Side& S) { return S.has ? S.inner : kNoWeight; };
        const int dist = std::min(v(A), v(B));
        r.distance = dist >= kNoWeight ? -1 : dist;
        const int lb = std::min(A.has ? A.outer : kNoWeight, B.has ? B.outer : kNoWeight);
        r.lower_bound = lb >= kNoWeight ? 0 : lb;
        r.levels = std::max(A.level, B.level);
        r.proven = A.resolved && B.resolved && dist < kNoWeight && !capped;
        return r;
    }
};

}  // namespace

// A batch as a session: problems join between rounds, every round enumerates one level of
// every unfinished problem, finished problems leave with their results.
struct BatchSession::Impl {
    struct Entry {
        long long id;
        std::unique_ptr<BatchProblem> prob;
        std::unique_ptr<Stepper> st;
        clk::time_point added;
    };
    BzOptions opt;
    u64 work_cap;
    long long next_id = 0;
    std::vector<Entry> live;
    std::vector<std::pair<long long, BzResult>> finished;

    void retire_done() {
        std::vector<Entry> keep;
        for (Entry& e : live) {
            if (!e.st->done) { keep.push_back(std::move(e)); continue; }
            BzResult r = e.st->result();
            r.seconds = std::chrono::duration<double>(clk::now() - e.added).count();
            finished.emplace_back(e.id, r);
        }
        live.swap(keep);
    }
};

BatchSession::BatchSession(const BzOptions& opt, unsigned long long work_cap) : impl_(std::make_unique<Impl>()) {
    impl_->opt = opt;
    impl_->work_cap = work_cap;
}
BatchSession::~BatchSession() = default;

int BatchSession::active() const { return (int)impl_->live.size(); }

unsigned long long BatchSession::device_bytes() const {
    unsigned long long b = 0;
    for (const auto& e : impl_->live) {
        if (e.st->A.has) b += qsf_hip_solve_bytes(e.st->A.key);
        if (e.st->pair && e.st->B.has) b += qsf_hip_solve_bytes(e.st->B.key);
    }
    return b;
}

std::vector<long long> BatchSession::add(std::vector<BatchProblem>&& probs) {
    Impl& I = *impl_;
    auto t0 = clk::now();
    std::vector<long long> ids;
    std::vector<Side*> sides;
    std::vector<const std::vector<int>*> orders;
    std::vector<const Proposal*> props;
    const size_t first = I.live.size();
    for (BatchProblem& bp0 : probs) {
        Impl::Entry e{I.next_id++, std::make_unique<BatchProblem>(std::move(bp0)), std::make_unique<Stepper>(), t0};
        BatchProblem& bp = *e.prob;
        Stepper& S = *e.st;
        S.pair = bp.pair;
        S.cap = bp.work_cap;
        S.A.prepare(bp.a, I.opt, bp.pair ? "Z" : "");
        sides.push_back(&S.A);
        orders.push_back(!bp.pair && !bp.order.empty() ? &bp.order : nullptr);
        props.push_back(!bp.pair && bp.proposal.candidates > 0 ? &bp.proposal : nullptr);
        if (bp.pair) {
            S.B.prepare(bp.b, I.opt, "X");
            sides.push_back(&S.B);
            orders.push_back(nullptr);
            props.push_back(nullptr);
        }
        ids.push_back(e.id);
        I.live.push_back(std::move(e));
    }
    setup_sides(sides, orders, props);
    for (size_t k = first; k < I.live.size(); ++k) {
        I.live[k].st->init();
        I.live[k].st->next();
    }
    stats.setup_s += std::chrono::duration<double>(clk::now() - t0).count();
    I.retire_done();
    return ids;
}

// One round. With lookahead (opt-in: QSA_LOOKAHEAD=1) every unfinished
// problem enqueues every level it can still need: each side's levels from its next one up
// to the first level whose bound reaches the current upper bound (the level where it would
// stop or resolve if nothing lighter turns up), within the work cap. All of them run in one
// batched device call (qsf_hip_enumerate_levels): each level on its own bound, so its
// result is the one the level would give alone, and the levels of one problem share a
// bound that skips a level once an earlier level has settled the problem. Then each
// problem replays its levels in the order bz_distance / bz_min_interleaved visit them and
// stops where they would stop. A level with skipped tiles is not replayed: it runs again,
// exactly, in the next round (rare: it needs a light codeword found at a higher level).
// Results therefore equal the level-by-level loop's. Without lookahead (the default) a
// round is one level per problem. Lookahead is off by default: its one decomposition of
// every level runs before any launch, and the rounds' small ones cost less (256 codes
// n ~ U{64..160}: 32.4 ms of kernels against 32.7, 50.8 ms wall against 46.6). Taking
// only the cheap levels after each problem's next one (up to 2^20..2^26 codewords) cut
// the rounds from 10 to 5..7 and made the batch 2-10% slower.
int BatchSession::step() {
    Impl& I = *impl_;
    const char* la_env = std::getenv("QSA_LOOKAHEAD");
    const bool lookahead = la_env && std::atoi(la_env) != 0;
    struct Enq {
        Side* side;
        int d;
        int plan;      // index into plans, -1: no generator active at this level
        u64 cw;
    };
    std::vector<qsf_enum_plan> plans;
    std::vector<int> floors, groups, offs;   // offs: level offset within its problem (for ordering)
    std::vector<std::vector<Enq>> enq(I.live.size());
    for (size_t k = 0; k < I.live.size(); ++k) {
        Stepper& S = *I.live[k].st;
        if (S.done || !S.cur) continue;
        const u64 cap = S.cap ? S.cap : I.work_cap;
        u64 spent = S.A.work + (S.pair ? S.B.work : 0);
        // the visiting order of the level loop from the current position, without results:
        // single side d, d+1, ...; a pair alternates sides at each level
        const int best_now = S.pair ? S.best : S.A.inner;
        auto horizon = [&](const Side& side) {   // last level this side can still need
            for (int d = 1; d <= side.K; ++d)
                if (side.outer_at(d) >= std::min(best_now, side.inner)) return d;
            return side.K;
        };
        const int hA = S.A.has ? horizon(S.A) : 0, hB = (S.pair && S.B.has) ? horizon(S.B) : 0;
        int d = S.d, turn = S.turn, off = 0;
        const int max_levels = lookahead ? 64 : 1;
        for (int steps = 0; (int)enq[k].size() < max_levels && steps < 4 * 256; ++steps) {
            Side* side = !S.pair ? &S.A : (turn == 0 ? &S.A : &S.B);
            const int h = side == &S.A ? hA : hB;
            if (side->has && !side->resolved && d <= side->K && d <= h) {
                const u64 w = side->level_work(d);
                if (cap && (w > cap || spent > cap - w)) break;   // the level loop stops at the cap here
                spent += w;
                qsf_enum_plan pl;
                if (side->plan(d, pl)) {
                    enq[k].push_back(Enq{side, d, (int)plans.size(), w});
                    plans.push_back(pl);
                    floors.push_back(side->outer_at(d - 1));
                    groups.push_back((int)k);
                    offs.push_back(off++);
                } else {
                    enq[k].push_back(Enq{side, d, -1, w});
                }
            }
            if (!S.pair) {
                if (++d > std::min(S.A.K, hA)) break;
            } else {
                if (++turn == 2) {
                    turn = 0;
                    if (++d > std::max(hA, hB)) break;
                }
            }
        }
    }
    // levels in ascending order across problems (the launch order of their tiles), so that
    // low levels, which can settle a problem, run first
    std::vector<int> order(plans.size());
    for (size_t j = 0; j < order.size(); ++j) order[j] = (int)j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return offs[a] < offs[b]; });
    std::vector<int> pos(plans.size());
    std::vector<qsf_enum_plan> oplans(plans.size());
    std::vector<int> ofloors(plans.size()), ogroups(plans.size());
    for (size_t r = 0; r < order.size(); ++r) {
        pos[order[r]] = (int)r;
        oplans[r] = plans[order[r]];
        ofloors[r] = floors[order[r]];
        ogroups[r] = groups[order[r]];
    }
    std::vector<int> found(plans.size(), kNoWeight);
    std::vector<uint64_t> skipped(plans.size(), 0);
    double ms = 0;
    qsf_hip_stats round{};
    if (!oplans.empty()) {
        auto tr = clk::now();
        const int rc = qsf_hip_enumerate_levels(oplans.data(), ofloors.data(), ogroups.data(), (int)oplans.size(),
                                                found.data(), skipped.data());
        if (rc != 0) throw EngineError{rc, qsf_hip_last_error()};
        qsf_hip_last_stats(&round);
        ms = std::chrono::duration<double, std::milli>(clk::now() - tr).count();
        if (I.opt.profile)
            std::fprintf(stderr, "[prof round] %d: %zu levels, %llu codewords, call %.3f ms, device %.3f ms\n", stats.rounds,
                         oplans.size(), (unsigned long long)round.codewords, ms, round.kernel_ms);
        ++stats.rounds;
        stats.codewords += round.codewords;
        stats.device_ms += round.kernel_ms;
    }
    // replay each problem's levels in the level loop's order
    for (size_t k = 0; k < I.live.size(); ++k) {
        Stepper& S = *I.live[k].st;
        size_t fed = 0;
        for (const Enq& q : enq[k]) {
            if (S.done || !S.cur) break;
            const u64 cap = S.cap ? S.cap : I.work_cap;
            const u64 spent = S.A.work + (S.pair ? S.B.work : 0);
            if (cap && (q.cw > cap || spent > cap - q.cw)) { S.capped = S.done = true; break; }
            if (S.cur != q.side || S.d != q.d) {
                // the loop resolved a side or moved on: skip enqueued levels it no longer visits
                if (q.side == S.cur && q.d < S.d) continue;
                if (q.side != S.cur) continue;
                break;
            }
            if (q.plan < 0) {   // no generator active at this level
                S.after(kNoWeight, 0.0, qsf_hip_stats{});
                S.next();
                continue;
            }
            const int r = pos[q.plan];
            if (skipped[r] != 0) break;   // run this level again, exactly, next round
            qsf_hip_stats one{};
            one.codewords = q.cw;
            one.kernel = round.kernel;
            one.kernel_ms = round.kernel_ms;
            S.after(found[r], ms, one);
            S.next();
            ++fed;
        }
        if (I.opt.profile && !enq[k].empty()) {
            u64 waste = 0;
            size_t nq = 0;
            for (const Enq& q : enq[k]) if (q.plan >= 0) ++nq;
            unsigned long long skip = 0;
            for (size_t t = fed; t < enq[k].size(); ++t)
                if (enq[k][t].plan >= 0) {
                    waste += enq[k][t].cw;
                    skip += skipped[pos[enq[k][t].plan]];
                }
            std::fprintf(stderr, "[prof lookahead] problem %zu: %zu levels enqueued, %zu replayed, %llu codewords not replayed, %llu of their tiles skipped%s\n",
                         k, nq, fed, (unsigned long long)waste, skip, S.done ? (S.capped ? " (capped)" : " (done)") : "");
        }
        // the cap check for the next level the loop wants (as the level loop does before it)
        if (!S.done && S.cur) {
            const u64 cap = S.cap ? S.cap : I.work_cap;
            const u64 w = S.cur->level_work(S.d);
            const u64 spent = S.A.work + (S.pair ? S.B.work : 0);
            if (cap && (w > cap || spent > cap - w)) S.capped = S.done = true;
        }
    }
    for (auto& e : I.live) stats.capped += (e.st->done && e.st->capped) ? 1 : 0;
    I.retire_done();
    return (int)I.live.size();
}

std::vector<std::pair<long long, BzResult>> BatchSession::take_finished() {
    std::vector<std::pair<long long, BzResult>> out;
    out.swap(impl_->finished);
    return out;
}

std::vector<BzResult> bz_batch(std::vector<BatchProblem>& probs, const BzOptions& opt, unsigned long long work_cap,
                               BatchStats* stats) {
    auto t0 = clk::now();
    BatchSession S(opt, work_cap);
    const size_t np = probs.size();
    std::vector<long long> ids = S.add(std::move(probs));
    while (S.active() > 0) S.step();
    std::vector<BzResult> out(np);
    std::unordered_map<long long, size_t> at;
    for (size_t i = 0; i < np; ++i) at[ids[i]] = i;
    const double secs = std::chrono::duration<double>(clk::now() - t0).count();
    for (auto& f : S.take_finished()) {
        out[at[f.first]] = f.second;
        out[at[f.first]].seconds = secs;
    }
    S.stats.seconds = secs;
    if (stats) *stats = S.stats;
    return out;
}

// ---------------------------------------------------------------- meet in the middle

namespace {

// One MITM subproblem stepped one weight at a time (state as cpp/distfind/mitm.cpp:403-450).
struct MitmSide {
    bool has = false, resolved = false, confirmed = false, count_all = false;
    int n = 0, nl = 0, target = 0, seed_upper = 0, found = -1, level = 0;
    int qubits = 0;   // > 0: symplectic problem; weights and halves count qubits
    const char* tag = "";
    std::vector<u64> syn, log;

    void setup(const Problem& p, const char* t) {
        tag = t;
        n = p.n;
        qubits = p.qubits;
        const int K = p.gen.rows;
        has = !(K == 0 || (!p.every_nonzero && p.detector.rows == 0));
        if (!has) { resolved = confirmed = true; return; }
        const BitMat H = kernel_basis(p.gen);   // the code's parity checks
        const int coords = qubits > 0 ? qubits : n;
        if (H.rows > 63 || p.detector.rows > 64 || coords > 128 || coords - coords / 2 > 64)
            throw EngineError{4, qubits > 0 ? "symplectic mitm on the device needs at most 63 parity checks, 64 detector "
                                              "rows and 128 qubits"
                                            : "mitm on the device needs at most 63 parity checks, 64 detector rows and "
                                              "n <= 128"};
        count_all = p.every_nonzero;
        syn.assign(n, 0);
        log.assign(n, 0);
        for (int r = 0; r < H.rows; ++r)
            for (int j = 0; j < n; ++j)
                if (H.bit(r, j)) syn[j] |= 1ull << r;
        if (!count_all)
            for (int c = 0; c < p.detector.rows; ++c)
                for (int j = 0; j < n; ++j)
                    if (p.detector.bit(c, j)) log[j] |= 1ull << c;
        nl = coords / 2;
        seed_upper = p.seed_upper;
        target = seed_upper > 0 ? std::min(seed_upper, coords) : coords;
    }

    bool step(int d) {
        ++level;
        int hit = 0;
        const int rc = qubits > 0 ? qsf_hip_mitm_weight_sym(syn.data(), log.data(), qubits, nl, d, &hit)
                                  : qsf_hip_mitm_weight(syn.data(), log.data(), n, nl, d, count_all ? 1 : 0, &hit);
        if (rc != 0) throw EngineError{rc, qsf_hip_mitm_last_error()};
        if (hit) found = d;
        return hit != 0;
    }
};

}  // namespace

BzResult mitm_distance(const Problem& p, const BzOptions& opt) {
    auto t0 = clk::now();
    BzResult r;
    MitmSide S;
    S.setup(p, "");
    if (S.has) {
        for (int d = 1; d <= S.target; ++d) {
            if (S.step(d)) {
                if (opt.verbose) say_final(d, std::chrono::duration<double>(clk::now() - t0).count(), "");
                break;
            }
        }
        if (S.found < 0 && S.seed_upper > 0 && S.target == S.seed_upper) S.found = S.seed_upper;
        r.distance = S.found;
        r.levels = S.level;
        r.lower_bound = S.found >= 0 ? S.found : 0;
        r.proven = S.found >= 0;
    }
    r.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return r;
}

BzResult mitm_min_interleaved(const Problem& pz, const Problem& px, const BzOptions& opt) {
    auto t0 = clk::now();
    MitmSide Z, X;
    Z.setup(pz, "Z");
    X.setup(px, "X");
    MitmSide* sides[2] = {&Z, &X};
    int best = kNoWeight;
    if (Z.has && Z.seed_upper > 0) best = std::min(best, Z.seed_upper);
    if (X.has && X.seed_upper > 0) best = std::min(best, X.seed_upper);
    const int maxd = std::max(Z.has ? Z.target : 0, X.has ? X.target : 0);
    for (int d = 1; d <= maxd; ++d) {
        for (MitmSide* S : sides) {
            if (S->resolved) continue;
            if (d > S->target) {   // its own ceiling without a hit: the seed certifies it
                if (S->seed_upper > 0 && S->target == S->seed_upper) {
                    S->found = S->seed_upper;
                    S->confirmed = true;
                    best = std::min(best, S->found);
                }
                S->resolved = true;
                continue;
            }
            if (d >= best) {   // cannot lower the running minimum
                S->confirmed = true;
                if (S->seed_upper > 0 && best == S->seed_upper) S->found = best;
                S->resolved = true;
                continue;
            }
            if (S->step(d)) {
                best = std::min(best, d);
                S->confirmed = S->resolved = true;
                if (opt.verbose) say_final(d, std::chrono::duration<double>(clk::now() - t0).count(), S->tag);
            }
        }
        if (Z.resolved && X.resolved) break;
    }
    auto v = [](const MitmSide& S) { return S.found >= 0 ? S.found : kNoWeight; };
    const int dist = std::min(v(Z), v(X));
    BzResult r;
    r.distance = dist >= kNoWeight ? -1 : dist;
    r.lower_bound = dist >= kNoWeight ? 0 : dist;
    r.levels = std::max(Z.level, X.level);
    r.proven = Z.confirmed && X.confirmed && dist < kNoWeight;
    r.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return r;
}

}  // namespace qsa
