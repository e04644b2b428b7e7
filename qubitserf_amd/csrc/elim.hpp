// Device GF(2) elimination: random information-set trials (see elim.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace qsa {

using u64 = std::uint64_t;

constexpr int kElimNone = 1 << 30;
constexpr long kElimLdsWords = 5120;   // 40 KiB of u64 rows per workgroup; the permutation adds 4 n bytes
constexpr int kElimMaxN = 4096;

struct ElimArgs {
    const u64* gen;      // K x stride u64 rows (a basis of the code searched)
    const u64* check;    // kcheck x stride detector rows
    int K, n, stride, kcheck;
    int every_nonzero;   // 1: every nonzero row is a logical
    u64 seed;
    int trial_base;      // index of the launch's first trial (seeds the column order)
    u64* scratch;        // trials x K x stride u64 when the matrix does not fit LDS
    int* best;           // folded with atomicMin
};

size_t elim_lds_bytes(int K, int n, int stride);
size_t elim_scratch_words(int K, int stride, int trials);
hipError_t launch_is_trials(const ElimArgs& a, int trials, hipStream_t stream);

}  // namespace qsa
