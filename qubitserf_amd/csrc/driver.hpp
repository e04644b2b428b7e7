// Brouwer–Zimmermann weight-level driver over the gfx950 engine (host side).
#pragma once
#include <memory>
#include <string>
#include <utility>
#include <vector>
#include "problem.hpp"
#include "../../include/qsf_hip.h"

namespace qsa {

struct BzResult {
    int distance = -1;
    int lower_bound = 0;
    bool proven = false;
    int levels = 0;
    double seconds = 0.0;
};

// One enumerated level, for profiles and the learned proposal's training traces.
struct LevelTrace {
    char side;             // 'S' single, 'Z'/'X' interleaved
    int d, active, ngamma, K;
    int inner, outer;      // bounds after the level
    double ms;             // wall time of the level (host)
    qsf_hip_stats st;      // engine statistics of the level (zero when nothing was enumerated)
};

struct BzOptions {
    bool verbose = false;
    bool profile = false;           // DISTFIND_PROFILE-style per-level lines on stderr
    int is_trials = 256;            // random information-set trials for the upper bound
    // Sound level schedule (off = the reference's). The reference enumerates a generator of
    // rank r < K only at levels d with (d+1) - (K-r) > 0 (cpp/distfind/bz.cpp:142-144) but
    // adds its max(0, (d+1) - (K-r)) to the lower bound (bz.cpp:159-161), which assumes its
    // levels 1..K-r-1 were enumerated too; a lighter codeword there can be missed. On: the
    // generators that contribute by the level where the bound first reaches the upper bound
    // are enumerated from level 1, and only they enter the bound.
    bool sound = false;
    unsigned long long seed = 0x9e3779b9ull;
    std::vector<LevelTrace>* trace = nullptr;
    // Optional column order for the greedy information-set peel of a single-side solve (a
    // proposal's order); null = 0..n-1, the reference's order.
    const std::vector<int>* order = nullptr;
};

// Raised when the device engine reports an error; `code` is its qsf_hip_* return code.
struct EngineError {
    int code;
    std::string what;
};

// Exact minimum over one subproblem, level by level (cpp/distfind/bz.cpp:72-182).
BzResult bz_distance(const Problem& p, const BzOptions& opt);

// Exact min over two subproblems stepped level by level in turn, a side retiring once its
// lower bound reaches the best weight found so far (cpp/distfind/bz.cpp:296-357).
BzResult bz_min_interleaved(const Problem& pz, const Problem& px, const BzOptions& opt);

// Meet-in-the-middle distance of one subproblem on the device (cpp/distfind/mitm.cpp:318-392):
// weights d = 1, 2, ... up to the seeded upper bound, first hit wins. Needs at most 63
// parity checks, 64 detector rows and n <= 128 (EngineError 4 otherwise).
BzResult mitm_distance(const Problem& p, const BzOptions& opt);
// The interleaved CSS minimum of the two sides, capping a side once it cannot lower the
// best (cpp/distfind/mitm.cpp:463-537).
BzResult mitm_min_interleaved(const Problem& pz, const Problem& px, const BzOptions& opt);

// One problem of a batch: a single side (classical, or one CSS side), or the two CSS
// sides of a min-distance solve (pair), with an optional proposal order for a single side.
// A learned proposal for one single-side problem: a score per column; `candidates` column
// orders are sampled from the Plackett–Luce distribution of the scores (seeded), and the
// peel of the order with the least predicted BZ work replaces the natural one when it is
// strictly better.
struct Proposal {
    std::vector<float> scores;
    int candidates = 0;
    unsigned long long seed = 0;
};

struct BatchProblem {
    Problem a, b;
    bool pair = false;
    std::vector<int> order;
    Proposal proposal;
    unsigned long long work_cap = 0;   // > 0: overrides the batch's cap for this problem
};

struct BatchStats {
    int rounds = 0;                 // level-synchronous rounds (device launches of the enumerator)
    int capped = 0;                 // problems stopped by the work cap
    unsigned long long codewords = 0;
    double device_ms = 0.0;         // enumeration kernels
    double setup_s = 0.0;           // information sets and upper bounds (device) incl. host set-up
    double seconds = 0.0;
};

// A batch that problems can join between rounds (continuous batching): add() builds the
// new problems' information sets and bounds on the device; step() enumerates one level of
// every unfinished problem in one batched device call; finished problems (with
// seconds = time since their add) are collected by take_finished(). Results equal the
// single-call results.
class BatchSession {
  public:
    BatchSession(const BzOptions& opt, unsigned long long work_cap);
    ~BatchSession();
    std::vector<long long> add(std::vector<BatchProblem>&& probs);
    int step();                                  // returns the problems still unfinished
    int active() const;
    // Device bytes the session's unfinished problems hold now: the arenas of their sides'
    // cached solves (qsf_hip_solve_bytes), measured, not estimated.
    unsigned long long device_bytes() const;
    std::vector<std::pair<long long, BzResult>> take_finished();
    BatchStats stats;

  private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// Many problems at once: each is stepped level by level exactly as bz_distance /
// bz_min_interleaved step it, and every round enumerates one level of every unfinished
// problem in one batched device call. work_cap > 0 stops a problem before a level would
// take its enumerated codewords past the cap (it then reports proven = false).
std::vector<BzResult> bz_batch(std::vector<BatchProblem>& probs, const BzOptions& opt, unsigned long long work_cap,
                               BatchStats* stats);

}  // namespace qsa
