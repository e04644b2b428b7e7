// Batched GF(2) elimination on the device.
//
//  * info_sets: the greedy information sets of a BZ solve (the peel of
//    cpp/distfind/css.cpp:177-203 over restricted_rref, gf2.cpp:42-58): eliminate the
//    generator over the live columns in order; the pivots form a set and the matrix, now
//    systematic on them, is that set's generator; the next round continues from that
//    matrix over the live columns that were neither pivots nor zero columns; stop when a
//    round removes nothing (the live columns start as 0..n-1, or as a given column
//    order). One wave per problem. Given an explicit column partition
//    instead (the matroid partition, css.cpp:274-328), each part is eliminated from the
//    original generator.
//  * is_trials: the random information-set upper bound (the heuristic of
//    cpp/distfind/bz.cpp:24-70): per trial a random column order, systematic form over
//    it, lightest nonzero logical row. One wave per (trial, problem).
//
// Elimination step shared by both: pivot = first row at/below the rank with a 1 in the
// column (found with one ballot per 64 rows), swap it up, clear the column in every other
// row. Lanes own rows r = lane + 64 j. Matrices live in LDS when K x stride <= kElimLdsWords
// words, else in the problem's global scratch.
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

// Gauss–Jordan of the K x st matrix M over columns cols[0..count) in order. Calls
// on_pivot(c) (lane 0) for each pivot column. Returns the rank. All lanes of the wave call it.
template <class F>
__device__ int eliminate(u64* M, int K, int st, const int* cols, int count, F on_pivot) {
    const int lane = threadIdx.x;
    int rank = 0;
    for (int ci = 0; ci < count && rank < K; ++ci) {
        const int c = cols[ci];
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
        if (lane == 0) on_pivot(c);
        __syncthreads();
        ++rank;
    }
    return rank;
}

__global__ __launch_bounds__(64) void is_trials(const ElimJob* __restrict__ jobs, u64 seed, int trial_base) {
    extern __shared__ u64 lds[];
    const ElimJob& J = jobs[blockIdx.y];
    const int lane = threadIdx.x;
    const int K = J.K, n = J.n, st = J.stride;
    if (K == 0 || (!J.every_nonzero && J.kcheck == 0)) return;
    const long trial = (long)trial_base + blockIdx.x;
    const bool in_lds = (long)K * st <= kElimLdsWords;
    u64* M = in_lds ? lds : J.scratch + (size_t)blockIdx.x * K * st;
    int* perm = reinterpret_cast<int*>(lds + (in_lds ? (size_t)K * st : 0));

    if (lane == 0) {   // Fisher–Yates, j = floor(hi32(x) * (i + 1) / 2^32)
        u64 s = seed ^ (0xd1b54a32d192ed03ull * (u64)(trial + 1));
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int i = n - 1; i > 0; --i) {
            const int j = (int)(((splitmix(s) >> 32) * (u64)(i + 1)) >> 32);
            const int t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
    }
    for (int i = lane; i < K * st; i += 64) M[i] = J.gen[i];
    __syncthreads();
    eliminate(M, K, st, perm, n, [](int) {});

    int mine = kElimNone;
    for (int r = lane; r < K; r += 64) {
        const u64* R = M + (size_t)r * st;
        int wt = 0;
        for (int i = 0; i < st; ++i) wt += __popcll(R[i]);
        if (wt == 0 || wt >= mine) continue;
        bool logical = J.every_nonzero != 0;
        for (int c = 0; c < J.kcheck && !logical; ++c) {
            u64 acc = 0;
            for (int i = 0; i < st; ++i) acc ^= J.check[(size_t)c * st + i] & R[i];
            logical = __popcll(acc) & 1;
        }
        if (logical) mine = wt;
    }
    for (int off = 32; off > 0; off >>= 1) mine = min(mine, __shfl_xor(mine, off));
    if (lane == 0 && mine < kElimNone) atomicMin(J.best, mine);
}

__global__ __launch_bounds__(64) void info_sets(const ElimJob* __restrict__ jobs) {
    extern __shared__ u64 lds[];
    __shared__ int s_nlive;
    const ElimJob& J = jobs[blockIdx.x];
    const int lane = threadIdx.x;
    const int K = J.K, n = J.n, st = J.stride;
    const bool in_lds = (long)K * st <= kElimLdsWords;
    u64* W = in_lds ? lds : J.scratch;
    u64* used = lds + (in_lds ? (size_t)K * st : 0);   // st words: pivot columns of this round
    u64* nzc = used + st;                              // st words: nonzero columns of gen
    int* live = reinterpret_cast<int*>(nzc + st);      // n ints

    if (J.nparts > 0) {   // sets of a given partition, each from the original generator
        if (J.set_of_col)
            for (int i = lane; i < n; i += 64) J.set_of_col[i] = -1;
        for (int p = 0; p < J.nparts; ++p) {
            for (int i = lane; i < K * st; i += 64) W[i] = J.gen[i];
            __syncthreads();
            const int o0 = J.part_off[p];
            const int rank = eliminate(W, K, st, J.part_cols + o0, J.part_off[p + 1] - o0, [&](int c) {
                if (J.set_of_col) J.set_of_col[c] = p;
            });
            if (p < J.set_cap) {
                if (J.sets) {
                    u64* out = J.sets + (size_t)p * K * st;
                    for (int i = lane; i < K * st; i += 64) out[i] = W[i];
                }
                if (lane == 0) J.ranks[p] = rank;
            }
            __syncthreads();
        }
        if (lane == 0) *J.nsets = J.nparts;
        return;
    }
    for (int i = lane; i < K * st; i += 64) W[i] = J.gen[i];
    for (int w = lane; w < st; w += 64) {
        u64 o = 0;
        for (int r = 0; r < K; ++r) o |= J.gen[(size_t)r * st + w];
        nzc[w] = o;
        used[w] = 0;
    }
    for (int i = lane; i < n; i += 64) live[i] = J.order ? J.order[i] : i;
    if (J.set_of_col)
        for (int i = lane; i < n; i += 64) J.set_of_col[i] = -1;
    __syncthreads();

    int nlive = n, nsets = 0;
    for (;;) {
        const int rank = eliminate(W, K, st, live, nlive, [&](int c) {
            used[c >> 6] |= 1ull << (c & 63);
            if (J.set_of_col) J.set_of_col[c] = nsets;
        });
        if (nsets < J.set_cap) {
            if (J.sets) {
                u64* out = J.sets + (size_t)nsets * K * st;
                for (int i = lane; i < K * st; i += 64) out[i] = W[i];
            }
            if (lane == 0) J.ranks[nsets] = rank;
        }
        ++nsets;
        if (lane == 0) {   // keep the order: live minus this round's pivots minus zero columns
            int k = 0;
            for (int i = 0; i < nlive; ++i) {
                const int c = live[i];
                const u64 b = 1ull << (c & 63);
                if (!(used[c >> 6] & b) && (nzc[c >> 6] & b)) live[k++] = c;
            }
            s_nlive = k;
        }
        __syncthreads();
        const int next = s_nlive;
        if (next == nlive || next == 0) break;
        nlive = next;
        for (int w = lane; w < st; w += 64) used[w] = 0;
        __syncthreads();
    }
    if (lane == 0) *J.nsets = nsets;
}

}  // namespace

size_t elim_scratch_words(int K, int stride, int trials) {
    return (long)K * stride <= kElimLdsWords ? 0 : (size_t)K * stride * (size_t)trials;
}

static size_t lds_matrix_bytes(const ElimJob& j) {
    const long w = (long)j.K * j.stride;
    return w <= kElimLdsWords ? (size_t)w * 8 : 0;
}

hipError_t launch_is_trials(const ElimJob* jobs, int njobs, size_t lds_bytes, u64 seed, int trial_base, int trials,
                            hipStream_t stream) {
    hipLaunchKernelGGL(is_trials, dim3((unsigned)trials, (unsigned)njobs), dim3(64), lds_bytes, stream, jobs, seed,
                       trial_base);
    return hipGetLastError();
}

hipError_t launch_info_sets(const ElimJob* jobs, int njobs, size_t lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL(info_sets, dim3((unsigned)njobs), dim3(64), lds_bytes, stream, jobs);
    return hipGetLastError();
}

size_t is_trials_lds_bytes(const ElimJob& j) { return lds_matrix_bytes(j) + (size_t)j.n * 4; }
size_t info_sets_lds_bytes(const ElimJob& j) { return lds_matrix_bytes(j) + (size_t)2 * j.stride * 8 + (size_t)j.n * 4; }

}  // namespace qsa
