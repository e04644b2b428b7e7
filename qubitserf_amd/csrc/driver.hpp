// Brouwer–Zimmermann weight-level driver over the gfx950 engine (host side).
#pragma once
#include <string>
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
    unsigned long long seed = 0x9e3779b9ull;
    std::vector<LevelTrace>* trace = nullptr;
    // Optional column partition to build the information sets from (the learned proposal);
    // empty = the greedy / matroid-partition sets.
    const std::vector<std::vector<int>>* parts = nullptr;
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

}  // namespace qsa
