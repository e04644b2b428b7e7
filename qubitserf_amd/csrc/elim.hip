// Batched GF(2) Gauss–Jordan elimination on the device: the random information-set upper
// bound that seeds a BZ solve (the heuristic of cpp/distfind/bz.cpp:24-70, as one launch).
//
// One wave per trial. The trial draws a random column order (Fisher–Yates over splitmix64,
// the same stream as the host restatement), puts the K x n generator into systematic form
// over that order — pivot = first row at/below the rank with a 1 in the column, found with
// one ballot per 64 rows; swap; clear the column in every other row — and every row of the
// result is a codeword; the lightest nonzero logical row is an upper bound on the
// distance. Lanes own rows (r = lane + 64 j). The matrix lives in LDS when it fits
// (K x stride <= kElimLdsWords words), else in the trial's slice of a global scratch
// buffer. One atomicMin per trial folds the bound.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "elim.hpp"

namespace qsa {

namespace {

__device__ __forceinline__ u64 splitmix(u64& s) {
    u64 z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(64) void is_trials(ElimArgs a) {
    extern __shared__ u64 lds[];
    const int lane = threadIdx.x;
    const int K = a.K, n = a.n, st = a.stride;
    const long trial = (long)a.trial_base + blockIdx.x;
    const bool in_lds = (long)K * st <= kElimLdsWords;
    u64* M = in_lds ? lds : a.scratch + (size_t)blockIdx.x * K * st;
    int* perm = reinterpret_cast<int*>(lds + (in_lds ? (size_t)K * st : 0));

    if (lane == 0) {
        u64 s = a.seed ^ (0xd1b54a32d192ed03ull * (u64)(trial + 1));
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int i = n - 1; i > 0; --i) {
            const int j = (int)(splitmix(s) % (u64)(i + 1));
            const int t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
    }
    for (int i = lane; i < K * st; i += 64) M[i] = a.gen[i];
    __syncthreads();

    int rank = 0;
    for (int ci = 0; ci < n && rank < K; ++ci) {
        const int c = perm[ci];
        const int w = c >> 6;
        const u64 bit = 1ull << (c & 63);
        int sel = K;
        for (int base = rank; base < K; base += 64) {
            const int r = base + lane;
            const u64 b = __ballot(r < K && (M[(size_t)r * st + w] & bit) != 0);
            if (b) { sel = base + __ffsll((unsigned long long)b) - 1; break; }
        }
        if (sel == K) continue;   // wave-uniform: no pivot in this column
        if (sel != rank) {
            for (int i = lane; i < st; i += 64) {
                const u64 x = M[(size_t)rank * st + i];
                M[(size_t)rank * st + i] = M[(size_t)sel * st + i];
                M[(size_t)sel * st + i] = x;
            }
            __syncthreads();
        }
        const u64* P = M + (size_t)rank * st;
        for (int r = lane; r < K; r += 64)
            if (r != rank && (M[(size_t)r * st + w] & bit))
                for (int i = 0; i < st; ++i) M[(size_t)r * st + i] ^= P[i];
        __syncthreads();
        ++rank;
    }

    int mine = kElimNone;
    for (int r = lane; r < K; r += 64) {
        const u64* R = M + (size_t)r * st;
        int wt = 0;
        for (int i = 0; i < st; ++i) wt += __popcll(R[i]);
        if (wt == 0 || wt >= mine) continue;
        bool logical = a.every_nonzero != 0;
        for (int c = 0; c < a.kcheck && !logical; ++c) {
            u64 acc = 0;
            for (int i = 0; i < st; ++i) acc ^= a.check[(size_t)c * st + i] & R[i];
            logical = __popcll(acc) & 1;
        }
        if (logical) mine = wt;
    }
    for (int off = 32; off > 0; off >>= 1) mine = min(mine, __shfl_xor(mine, off));
    if (lane == 0 && mine < kElimNone) atomicMin(a.best, mine);
}

}  // namespace

size_t elim_lds_bytes(int K, int n, int stride) {
    const long words = (long)K * stride;
    return (size_t)((words <= kElimLdsWords ? words : 0) * 8 + (long)n * 4);
}

size_t elim_scratch_words(int K, int stride, int trials) {
    return (long)K * stride <= kElimLdsWords ? 0 : (size_t)K * stride * (size_t)trials;
}

hipError_t launch_is_trials(const ElimArgs& a, int trials, hipStream_t stream) {
    hipLaunchKernelGGL(is_trials, dim3((unsigned)trials), dim3(64), elim_lds_bytes(a.K, a.n, a.stride), stream, a);
    return hipGetLastError();
}

}  // namespace qsa
