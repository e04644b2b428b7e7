// distfind_* and codeaut_bz_* C ABI over the gfx950 engine (see include/distfind_capi.h,
// include/codeaut_capi.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/codeaut_capi.h"
#include "../../include/distfind_capi.h"
#include "../../include/qsf_hip.h"
#include "driver.hpp"
#include "pool.hpp"

using namespace qsa;

namespace qsa {
void set_last_error(const std::string& msg);
}

namespace {

thread_local std::string g_err;
thread_local std::vector<LevelTrace> g_trace;
thread_local BatchStats g_batch;   // last distfind_batch_distance call on this thread

constexpr int RC_OK = 0, RC_NULL = 1, RC_ODD = 2, RC_CC_NONCSS = 3, RC_METHOD = 4, RC_BACKEND = 5, RC_DEVICE = 6;

void fill(DistfindResult* out, const BzResult& r) {
    out->distance = r.distance;
    out->lower_bound = r.lower_bound;
    out->proven = r.proven ? 1 : 0;
    out->levels = r.levels;
    out->seconds = r.seconds;
    std::memset(out->backend, 0, sizeof(out->backend));
    std::strncpy(out->backend, "hip", sizeof(out->backend) - 1);
}

// Shared argument policy: methods "bz" and "mitm" (classical, CSS and non-CSS codes);
// backend "gpu"/"auto" only, device present.
int check_args(const char* method, const char* backend, bool noncss = false) {
    std::string m = method ? method : "bz";
    std::string b = backend ? backend : "auto";
    if (m == "cc" && noncss) { g_err = "connected cluster has no non-CSS form"; return RC_CC_NONCSS; }
    if (m != "bz" && m != "mitm") { g_err = "method '" + m + "' is not provided by the gfx950 engine (it runs 'bz' and 'mitm')"; return RC_METHOD; }
    if (b != "gpu" && b != "auto") { g_err = "backend '" + b + "' is not provided by the gfx950 engine (use 'gpu')"; return RC_BACKEND; }
    if (!qsf_hip_available()) { g_err = std::string("no usable gfx950 device: ") + qsf_hip_last_error(); return RC_BACKEND; }
    return RC_OK;
}

// The sound level schedule (BzOptions::sound), process-wide: QSA_SOUND_BOUNDS=1 in the
// environment at first use, or distfind_set_sound_bounds.
std::atomic<int> g_sound{-1};
bool sound_on() {
    int v = g_sound.load();
    if (v < 0) {
        const char* e = std::getenv("QSA_SOUND_BOUNDS");
        v = (e && std::atoi(e) != 0) ? 1 : 0;
        g_sound.store(v);
    }
    return v != 0;
}

BzOptions options(int verbose) {
    BzOptions o;
    o.verbose = verbose != 0;
    o.sound = sound_on();
    o.profile = std::getenv("DISTFIND_PROFILE") != nullptr;
    g_trace.clear();
    o.trace = &g_trace;
    return o;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const EngineError& e) {
        g_err = e.what;
        return e.code == 4 ? RC_METHOD : RC_DEVICE;   // 4: the problem exceeds what a kernel takes
    } catch (const std::exception& e) {
        g_err = e.what();
        return RC_DEVICE;
    }
}

bool is_mitm(const char* method) { return method && std::string(method) == "mitm"; }

// A matrix argument must have rows >= 0 and cols >= 1: negative sizes would otherwise reach
// the bit-matrix allocation as huge unsigned sizes.
bool bad_shape(int rows, int cols) { return rows < 0 || cols <= 0; }

// css side / min dispatch, by method (cpp/distfind/capi.cpp:54-60)
int css_solve(const BitMat& Hx, const BitMat& Hz, char which, int verbose, DistfindResult* out, bool mitm = false) {
    BzOptions o = options(verbose);
    BzResult r;
    if (which == 'Z' || which == 'X') {
        const Problem p = css_side(Hx, Hz, which);
        r = mitm ? mitm_distance(p, o) : bz_distance(p, o);
    } else {
        const Problem pz = css_side(Hx, Hz, 'Z'), px = css_side(Hx, Hz, 'X');
        r = mitm ? mitm_min_interleaved(pz, px, o) : bz_min_interleaved(pz, px, o);
    }
    fill(out, r);
    return RC_OK;
}

}  // namespace

extern "C" {

int distfind_css_distance(const uint8_t* Hx, int hx_rows, int hx_cols, const uint8_t* Hz, int hz_rows,
                          int hz_cols, const char* method, char which, const char* backend, int threads,
                          int verbose, DistfindResult* out) {
    (void)threads;
    if (!out || !Hx || !Hz) { g_err = "null pointer"; return RC_NULL; }
    if (bad_shape(hx_rows, hx_cols) || bad_shape(hz_rows, hz_cols)) { g_err = "negative or empty matrix shape"; return RC_NULL; }
    if (int rc = check_args(method, backend)) return rc;
    return guarded([&] {
        return css_solve(BitMat::from_bytes(Hx, hx_rows, hx_cols), BitMat::from_bytes(Hz, hz_rows, hz_cols), which,
                         verbose, out, is_mitm(method));
    });
}

int distfind_classical_distance(const uint8_t* H, int rows, int cols, const char* method, const char* backend,
                                int threads, int verbose, DistfindResult* out) {
    (void)threads;
    if (!out || !H) { g_err = "null pointer"; return RC_NULL; }
    if (bad_shape(rows, cols)) { g_err = "negative or empty matrix shape"; return RC_NULL; }
    if (int rc = check_args(method, backend)) return rc;
    return guarded([&] {
        BzOptions o = options(verbose);
        const Problem p = classical(BitMat::from_bytes(H, rows, cols));
        fill(out, is_mitm(method) ? mitm_distance(p, o) : bz_distance(p, o));
        return RC_OK;
    });
}

int distfind_stabilizer_distance(const uint8_t* S, int rows, int cols, const char* method, char which,
                                 const char* backend, int threads, int verbose, DistfindResult* out) {
    (void)threads;
    if (!out || !S) { g_err = "null pointer"; return RC_NULL; }
    if (bad_shape(rows, cols)) { g_err = "negative or empty matrix shape"; return RC_NULL; }
    if (cols % 2 != 0) { g_err = "symplectic matrix needs an even column count"; return RC_ODD; }
    BitMat m;
    bool css = false;
    if (int rc = guarded([&] {
            m = BitMat::from_bytes(S, rows, cols);
            css = is_css_symplectic(m);
            return RC_OK;
        }))
        return rc;
    if (int rc = check_args(method, backend, !css)) return rc;
    return guarded([&] {
        if (css) {
            BitMat Hx, Hz;
            split_css(m, Hx, Hz);
            return css_solve(Hx, Hz, which, verbose, out, is_mitm(method));
        }
        BzOptions o = options(verbose);
        if (is_mitm(method)) {   // symplectic meet in the middle (capi.cpp:212-218 of the reference)
            fill(out, mitm_distance(symplectic_plain(m), o));
            return RC_OK;
        }
        if (o.verbose)
            std::fprintf(stderr, "distfind: non-CSS BZ via the (a|b)->(a|b|a^b) isometry; bounds are Hamming "
                                 "weights of the length-%d code (twice the symplectic weight).\n", 3 * (cols / 2));
        BzResult r = bz_distance(symplectic_isometric(m), o);
        if (r.distance > 0) r.distance /= 2;       // Hamming weight = 2 x symplectic weight
        if (r.lower_bound > 0) r.lower_bound /= 2;
        fill(out, r);
        return RC_OK;
    });
}

int distfind_operator_weight(const uint8_t*, int, int, const uint8_t*, int, int, const uint8_t*, const uint8_t*, int,
                             const char*, const char*, int, int, DistfindOpResult* out) {
    if (out) std::memset(out, 0, sizeof(*out));
    g_err = "operator weight is not provided by the gfx950 engine";
    return RC_METHOD;
}

int distfind_subsystem_distance(const uint8_t*, int, int, const uint8_t*, int, int, const char*, char, const char*,
                                int, int, DistfindResult* out) {
    if (out) std::memset(out, 0, sizeof(*out));
    g_err = "subsystem distance is not provided by the gfx950 engine";
    return RC_METHOD;
}

int distfind_subsystem_stabilizer_distance(const uint8_t*, int, int, const char*, char, const char*, int, int,
                                           DistfindResult* out) {
    if (out) std::memset(out, 0, sizeof(*out));
    g_err = "subsystem distance is not provided by the gfx950 engine";
    return RC_METHOD;
}

int distfind_stabilizer_operator_weight(const uint8_t*, int, int, const uint8_t*, int, const char*, const char*, int,
                                        int, DistfindResult* out) {
    if (out) std::memset(out, 0, sizeof(*out));
    g_err = "operator weight is not provided by the gfx950 engine";
    return RC_METHOD;
}

int distfind_backend_available(const char* name) {
    std::string b = name ? name : "auto";
    if (b == "gpu" || b == "auto") return qsf_hip_available();
    return 0;
}

const char* distfind_version(void) { return "0.1.0+gfx950"; }

int distfind_set_sound_bounds(int on) {
    const int old = sound_on() ? 1 : 0;
    g_sound.store(on ? 1 : 0);
    return old;
}

const char* distfind_last_error(void) { return g_err.c_str(); }

int distfind_last_trace(DistfindLevel* out, int cap) {
    const int n = (int)g_trace.size();
    for (int i = 0; out && i < n && i < cap; ++i) {
        const LevelTrace& t = g_trace[i];
        DistfindLevel& o = out[i];
        std::memset(&o, 0, sizeof(o));
        o.side = t.side; o.d = t.d; o.active = t.active; o.ngamma = t.ngamma; o.K = t.K;
        o.inner = t.inner; o.outer = t.outer; o.kernel = t.st.kernel;
        o.fd = t.st.fd; o.xd = t.st.xd; o.rpl = t.st.rpl; o.mfma = t.st.mfma;
        o.codewords = t.st.codewords; o.medium_hits = t.st.medium_hits; o.slow_hits = t.st.slow_hits;
        o.ms = t.ms; o.kernel_ms = t.st.kernel_ms; o.ops_per_codeword = t.st.ops_per_codeword;
        o.side_ms = t.st.side_ms; o.side_codewords = t.st.side_codewords;
        o.side_ops_per_codeword = t.st.side_ops_per_codeword;
    }
    return n;
}

// ---------------------------------------------------------------- host logic, for tests

int distfind_matroid_parts(const uint8_t* G, int rows, int cols, int* part_of_col) {
    if (!G || !part_of_col || rows < 0 || cols <= 0) { g_err = "null or empty generator"; return RC_NULL; }
    std::vector<std::vector<int>> parts;
    if (!matroid_parts(BitMat::from_bytes(G, rows, cols), parts)) { g_err = "partition check failed"; return RC_DEVICE; }
    std::fill(part_of_col, part_of_col + cols, -1);
    for (size_t p = 0; p < parts.size(); ++p)
        for (int c : parts[p]) part_of_col[c] = (int)p;
    return RC_OK;
}

}  // extern "C"

namespace {

// Batch items -> problems (problem build on the host, as the single entry points do).
// One batch item's problem (kernel bases, detectors, proposal, order); rc != 0 with err set
// for a malformed item.
int build_problem(const DistfindBatchItem* items, int i, BatchProblem& bp, std::string& err) {
    const DistfindBatchItem& it = items[i];
    if (!it.A || (it.kind == 1 && !it.B)) { err = "null matrix in batch item " + std::to_string(i); return RC_NULL; }
    if (bad_shape(it.ar, it.ac) || (it.kind == 1 && bad_shape(it.br, it.bc))) {
        err = "negative or empty matrix shape in batch item " + std::to_string(i);
        return RC_NULL;
    }
    if (it.kind == 0) {
        bp.a = classical(BitMat::from_bytes(it.A, it.ar, it.ac));
    } else if (it.kind == 1) {
        if (it.ac != it.bc) { err = "Hx and Hz column counts differ in batch item " + std::to_string(i); return RC_NULL; }
        BitMat Hx = BitMat::from_bytes(it.A, it.ar, it.ac), Hz = BitMat::from_bytes(it.B, it.br, it.bc);
        const char w = it.which == 'z' ? 'Z' : it.which == 'x' ? 'X' : it.which;
        if (w == 'Z' || w == 'X') {
            bp.a = css_side(Hx, Hz, w);
        } else {
            bp.a = css_side(Hx, Hz, 'Z');
            bp.b = css_side(Hx, Hz, 'X');
            bp.pair = true;
        }
    } else {
        err = "batch item kind must be 0 (classical) or 1 (CSS)";
        return RC_METHOD;
    }
    bp.work_cap = it.work_cap > 0 ? (u64)it.work_cap : 0;
    if (it.scores && it.candidates > 0 && !bp.pair) {
        bp.proposal.scores.assign(it.scores, it.scores + it.ac);
        bp.proposal.candidates = std::min(it.candidates, 1024);
        bp.proposal.seed = it.proposal_seed;
    }
    if (it.order && !bp.pair) {
        const int n = it.ac;
        bp.order.assign(it.order, it.order + n);
        std::vector<char> seen(n, 0);
        for (int c : bp.order) {
            if (c < 0 || c >= n || seen[c]) { err = "batch item order is not a permutation"; return RC_NULL; }
            seen[c] = 1;
        }
    }
    return RC_OK;
}

// All items' problems, on up to 16 host threads for large batches; the first malformed
// item (by index) sets the error.
int build_problems(const DistfindBatchItem* items, int count, std::vector<BatchProblem>& probs) {
    probs.assign(count, BatchProblem{});
    std::vector<int> rcs(count, RC_OK);
    std::vector<std::string> errs(count);
    std::vector<std::exception_ptr> exc(count);
    HostPool::get().run((size_t)count, host_workers((size_t)count, 32), [&](size_t i) {
        try {
            rcs[i] = build_problem(items, (int)i, probs[i], errs[i]);
        } catch (...) {
            exc[i] = std::current_exception();
        }
    });
    for (int i = 0; i < count; ++i) {
        if (exc[i]) std::rethrow_exception(exc[i]);
        if (rcs[i] != RC_OK) {
            g_err = errs[i];
            return rcs[i];
        }
    }
    return RC_OK;
}

struct SessionHandle {
    BzOptions opt;
    std::unique_ptr<BatchSession> s;
    std::vector<std::pair<long long, BzResult>> pending;   // finished, not yet polled
};

}  // namespace

extern "C" {

int distfind_batch_distance(const DistfindBatchItem* items, int count, int64_t work_cap, int is_trials, int verbose,
                            DistfindResult* out) {
    if (count < 0 || (count > 0 && (!items || !out))) { g_err = "null batch"; return RC_NULL; }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<BatchProblem> probs;   // malformed items are reported before the device is probed
    if (int rc = guarded([&] { return build_problems(items, count, probs); })) return rc;
    if (int rc = check_args("bz", "gpu")) return rc;
    BzOptions o = options(verbose);
    o.trace = nullptr;
    if (is_trials > 0) o.is_trials = is_trials;
    if (o.profile)
        std::fprintf(stderr, "[prof batch] %d problems built in %.3f ms\n", count,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    return guarded([&] {
        std::vector<BzResult> r = bz_batch(probs, o, work_cap > 0 ? (u64)work_cap : 0, &g_batch);
        for (int i = 0; i < count; ++i) fill(out + i, r[i]);
        return RC_OK;
    });
}

void* distfind_session_new(int64_t work_cap, int is_trials) {
    if (check_args("bz", "gpu")) return nullptr;
    auto* h = new SessionHandle();
    h->opt = options(0);
    h->opt.trace = nullptr;
    if (is_trials > 0) h->opt.is_trials = is_trials;
    h->s = std::make_unique<BatchSession>(h->opt, work_cap > 0 ? (u64)work_cap : 0);
    return h;
}

int distfind_session_add(void* handle, const DistfindBatchItem* items, int count, int64_t* ids) {
    auto* h = static_cast<SessionHandle*>(handle);
    if (!h || count < 0 || (count > 0 && (!items || !ids))) { g_err = "null session or items"; return RC_NULL; }
    std::vector<BatchProblem> probs;
    if (int rc = guarded([&] { return build_problems(items, count, probs); })) return rc;
    return guarded([&] {
        std::vector<long long> got = h->s->add(std::move(probs));
        for (int i = 0; i < count; ++i) ids[i] = got[i];
        return RC_OK;
    });
}

int distfind_session_step(void* handle, int* active) {
    auto* h = static_cast<SessionHandle*>(handle);
    if (!h) { g_err = "null session"; return RC_NULL; }
    return guarded([&] {
        const int a = h->s->step();
        if (active) *active = a;
        return RC_OK;
    });
}

int distfind_session_poll(void* handle, int64_t* ids, DistfindResult* out, int cap) {
    auto* h = static_cast<SessionHandle*>(handle);
    if (!h || cap < 0 || (cap > 0 && (!ids || !out))) { g_err = "null session or outputs"; return -1; }
    auto& pending = h->pending;
    for (auto& f : h->s->take_finished()) pending.push_back(f);
    const int n = std::min<int>(cap, (int)pending.size());
    for (int i = 0; i < n; ++i) {
        ids[i] = pending[i].first;
        fill(out + i, pending[i].second);
    }
    pending.erase(pending.begin(), pending.begin() + n);
    return n;
}

int distfind_session_active(void* handle) {
    auto* h = static_cast<SessionHandle*>(handle);
    return h ? h->s->active() : -1;
}

int64_t distfind_session_device_bytes(void* handle) {
    auto* h = static_cast<SessionHandle*>(handle);
    return h ? (int64_t)h->s->device_bytes() : -1;
}

void distfind_session_stats(void* handle, DistfindBatchStats* out) {
    auto* h = static_cast<SessionHandle*>(handle);
    if (!h || !out) return;
    const BatchStats& b = h->s->stats;
    out->rounds = b.rounds;
    out->capped = b.capped;
    out->codewords = b.codewords;
    out->device_ms = b.device_ms;
    out->setup_s = b.setup_s;
    out->seconds = b.seconds;
}

void distfind_session_free(void* handle) { delete static_cast<SessionHandle*>(handle); }

void distfind_last_batch_stats(DistfindBatchStats* out) {
    if (!out) return;
    out->rounds = g_batch.rounds;
    out->capped = g_batch.capped;
    out->codewords = g_batch.codewords;
    out->device_ms = g_batch.device_ms;
    out->setup_s = g_batch.setup_s;
    out->seconds = g_batch.seconds;
}

// ---------------------------------------------------------------- codeaut collect

namespace {
struct CollectHandle {
    int n = 0, stride = 0;
    bool ok = false;
    int64_t combos = 0, count = 0;
    std::vector<u64> rows;
    std::string backend = "hip";
};
constexpr int64_t kComboCap = (int64_t)1 << 62;
}  // namespace

void* codeaut_bz_collect(const uint8_t* G, int32_t m, int32_t n, int32_t p, int32_t keep_weight, int32_t backend,
                         int32_t threads) {
    (void)threads;
    auto* H = new CollectHandle();
    H->n = n;
    H->stride = words(n);
    if (m <= 0 || n <= 0 || p <= 0) { H->ok = true; return H; }
    if (backend == 0 || !qsf_hip_available()) {
        g_err = backend == 0 ? "cpu collect is not provided by the gfx950 engine" : "no usable gfx950 device";
        H->ok = false;
        return H;
    }
    const int st = H->stride;
    BitMat Gm = BitMat::from_bytes(G, m, n);
    const int maxsw = std::min<int>(p, m);
    static std::atomic<u64> keys{0xc011ec7000000000ull};
    const u64 key = ++keys;
    // every subset size 1..maxsw in one device pass; rows sorted (lexicographic over the
    // packed words, cpp/codeaut/bz.cpp:98-113) and deduplicated on the device
    qsf_enum_plan pl{};
    pl.n = n; pl.stride = st; pl.K = m; pl.d = maxsw;
    pl.num_gamma = 1; pl.num_gamma_total = 1;
    pl.gamma = Gm.data.data(); pl.kcheck = 0; pl.check = nullptr;
    pl.current_best = kNoWeight; pl.buffers_key = key;
    uint64_t* buf = nullptr;
    int64_t rows = 0;
    const int rc = qsf_hip_collect_upto(&pl, keep_weight, &buf, &rows);
    if (rc != 0) {
        g_err = qsf_hip_last_error();
        H->ok = false;
        return H;
    }
    H->rows.assign(buf, buf + rows * st);
    qsf_hip_free(buf);
    for (int sw = 1; sw <= maxsw; ++sw) {
        // combination count, saturating like the reference's int64 counter (codeaut/bz.cpp:88-95)
        long double c = 1;
        for (int i = 0; i < sw; ++i) c = c * (m - i) / (i + 1);
        int64_t total = c > (long double)kComboCap ? kComboCap : (int64_t)(c + 0.5L);
        H->combos = H->combos > kComboCap - total ? kComboCap : H->combos + total;
    }
    H->count = (int64_t)H->rows.size() / st;
    H->ok = true;
    return H;
}

int32_t codeaut_bz_ok(void* h) { return h && ((CollectHandle*)h)->ok ? 1 : 0; }
int64_t codeaut_bz_count(void* h) { return h ? ((CollectHandle*)h)->count : -1; }
int64_t codeaut_bz_combos(void* h) { return h ? ((CollectHandle*)h)->combos : -1; }

void codeaut_bz_copy_rows(void* h, uint8_t* buf) {
    auto* H = (CollectHandle*)h;
    if (!H || !buf) return;
    for (int64_t r = 0; r < H->count; ++r) {
        const u64* src = H->rows.data() + r * H->stride;
        uint8_t* dst = buf + (size_t)r * H->n;
        for (int j = 0; j < H->n; ++j) dst[j] = (uint8_t)((src[j >> 6] >> (j & 63)) & 1);
    }
}

const char* codeaut_bz_backend(void* h) { return h ? ((CollectHandle*)h)->backend.c_str() : ""; }
void codeaut_bz_free(void* h) { delete (CollectHandle*)h; }

int32_t codeaut_backend_available(const char* name) {
    if (!name) return 0;
    std::string b = name;
    if (b == "gpu" || b == "auto") return qsf_hip_available();
    return 0;
}

const char* codeaut_version(void) { return "0.1.0+gfx950"; }

// ---------------------------------------------------------------- Leon (not provided)
// Leon's automorphism search (the reference's cpp/codeaut/leon.cpp) is not part of this
// engine. The symbols exist because the reference's codeaut binding attaches them without
// a guard (python/qubitserf/codeaut/_native.py:31-61): a handle reports ok = 0, so its
// caller (python/qubitserf/codeaut/leon.py:145-151) raises instead of reading a group.

namespace {
struct LeonHandle {
    int32_t n = 0;
};
}  // namespace

void* qaut_leon_run(const uint8_t*, int32_t, int32_t n, int32_t, int32_t) {
    g_err = "Leon's automorphism search is not provided by the gfx950 engine";
    auto* h = new LeonHandle();
    h->n = n;
    return h;
}
int32_t qaut_leon_ok(void*) { return 0; }
int32_t qaut_leon_dim(void*) { return -1; }
int32_t qaut_leon_n(void* h) { return h ? static_cast<LeonHandle*>(h)->n : 0; }
int32_t qaut_leon_num_codewords(void*) { return 0; }
int32_t qaut_leon_num_classes(void*) { return 0; }
int32_t qaut_leon_num_gens(void*) { return 0; }
int32_t qaut_leon_num_factors(void*) { return 0; }
void qaut_leon_copy_gens(void*, int32_t*) {}
void qaut_leon_copy_factors(void*, int64_t*) {}
void qaut_leon_copy_weights(void*, int32_t*) {}
void qaut_leon_free(void* h) { delete static_cast<LeonHandle*>(h); }

}  // extern "C"

namespace qsa {
void set_last_error(const std::string& msg) { g_err = msg; }
}
