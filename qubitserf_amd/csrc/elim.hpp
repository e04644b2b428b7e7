// Device GF(2) elimination (elim.hip): information sets in systematic form, and the
// random information-set upper bound. Both are batched: one launch serves many problems,
// one wave per (problem, trial) or per problem.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace qsa {

using u64 = std::uint64_t;

constexpr int kElimNone = 1 << 30;
constexpr long kElimLdsWords = 5120;   // 40 KiB of u64 rows per workgroup; column lists add 4-8 n bytes
constexpr int kElimMaxN = 4096;

// One problem: a K x n generator, stride = (n+63)/64 u64 words per row (device pointers).
struct ElimJob {
    const u64* gen;
    const u64* check;    // kcheck x stride detector rows
    int K, n, stride, kcheck;
    int every_nonzero;   // 1: every nonzero row is a logical
    int* best;           // random-IS bound, folded with atomicMin
    u64* scratch;        // matrices that do not fit LDS: trials x K x stride (bound) / K x stride (sets)
    // information sets
    u64* sets;           // out: set_cap x K x stride systematic generators, in peel order
    int* ranks;          // out: set_cap ranks
    int* nsets;          // out: number of sets (> set_cap: not all stored)
    int set_cap;
    // explicit column partition (nparts > 0): set p = gen made systematic over the columns
    // part_cols[part_off[p] .. part_off[p+1]) in that order, each from the original gen
    const int* part_cols;
    const int* part_off;
    int nparts;
    const int* order;    // greedy peel: initial column order (a permutation of [0, n)); null = 0..n-1
    int* set_of_col;     // out (n ints, may be null): the set whose pivot a column became, else -1
};

// Random information-set trials: trials x jobs workgroups of one wave. Trial t of a job
// takes its column order from Fisher–Yates over the splitmix64 stream seeded with
// seed ^ (0xd1b54a32d192ed03 * (t + 1)), position i swapped with floor(hi32(x) (i+1) / 2^32).
// `jobs` is a device array; lds_bytes = max over jobs of is_trials_lds_bytes.
hipError_t launch_is_trials(const ElimJob* jobs, int njobs, size_t lds_bytes, u64 seed, int trial_base, int trials,
                            hipStream_t stream);
size_t is_trials_lds_bytes(const ElimJob& j);
size_t elim_scratch_words(int K, int stride, int trials);   // per job, 0 when it fits LDS

// Greedy information sets (peel): one wave per job; lds_bytes = max of info_sets_lds_bytes.
hipError_t launch_info_sets(const ElimJob* jobs, int njobs, size_t lds_bytes, hipStream_t stream);
size_t info_sets_lds_bytes(const ElimJob& j);

}  // namespace qsa
