// Collect mode's device-side sort and deduplication: the rows a level's enumeration
// appended (packed, `stride` u64 words each, duplicates and any order) are sorted
// lexicographically by word (word 0 first, unsigned: the order of the reference's
// RowLess, cpp/codeaut/bz.cpp:41-51) and deduplicated, on the device, before one copy back.
//
// Sort: least-significant-word-first radix passes over a permutation (hipCUB's stable
// DeviceRadixSort on (word, index) pairs, one pass per word from the last to the first).
// Dedup: a row is kept when it differs from its predecessor; an exclusive scan of the keep
// flags gives each kept row its output slot.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <string>

#include "collect.hpp"

namespace qsa {
namespace {

__global__ void iota_u32(uint32_t* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

__global__ void gather_word(const uint64_t* rows, const uint32_t* perm, int stride, int w, uint64_t* key, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = rows[(uint64_t)perm[i] * stride + w];
}

// keep[i] = 1 when sorted row i differs from sorted row i - 1 (row 0 is always kept)
__global__ void mark_distinct(const uint64_t* rows, const uint32_t* perm, int stride, uint32_t* keep, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = 1;
    if (i > 0) {
        const uint64_t* a = rows + (uint64_t)perm[i] * stride;
        const uint64_t* b = rows + (uint64_t)perm[i - 1] * stride;
        k = 0;
        for (int w = 0; w < stride; ++w) k |= a[w] != b[w];
    }
    keep[i] = k;
}

__global__ void scatter_kept(const uint64_t* rows, const uint32_t* perm, const uint32_t* keep, const uint32_t* slot,
                             int stride, uint64_t* out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const uint64_t* a = rows + (uint64_t)perm[i] * stride;
    uint64_t* o = out + (uint64_t)slot[i] * stride;
    for (int w = 0; w < stride; ++w) o[w] = a[w];
}

struct Scratch {
    void* p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e == hipSuccess) n = bytes;
        return e;
    }
};

}  // namespace

// Sorts and deduplicates `n` device rows (n < 2^32) into `out` (device, room for n rows);
// *kept = the distinct row count. Work buffers grow on demand and are kept for later calls
// (the caller serialises calls). 0, 3 (HIP error) or 4 (too many rows).
int sort_dedup_rows(const uint64_t* rows, uint64_t n, int stride, uint64_t* out, uint64_t* kept, hipStream_t stream,
                    std::string& err) {
    static Scratch s_perm[2], s_key[2], s_keep, s_slot, s_tmp;
    *kept = 0;
    if (n == 0) return 0;
    if (n >= (1ull << 32)) { err = "collect: more than 2^32 rows"; return 4; }
#define QSA_CK(x)                                                       \
    do {                                                                \
        hipError_t e_ = (x);                                            \
        if (e_ != hipSuccess) { err = std::string(#x) + ": " + hipGetErrorString(e_); return 3; } \
    } while (0)
    const unsigned blocks = (unsigned)((n + 255) / 256);
    QSA_CK(s_perm[0].reserve(n * 4)); QSA_CK(s_perm[1].reserve(n * 4));
    QSA_CK(s_key[0].reserve(n * 8)); QSA_CK(s_key[1].reserve(n * 8));
    QSA_CK(s_keep.reserve(n * 4)); QSA_CK(s_slot.reserve(n * 4));
    auto* pa = static_cast<uint32_t*>(s_perm[0].p);
    auto* pb = static_cast<uint32_t*>(s_perm[1].p);
    auto* ka = static_cast<uint64_t*>(s_key[0].p);
    auto* kb = static_cast<uint64_t*>(s_key[1].p);
    hipLaunchKernelGGL(iota_u32, dim3(blocks), dim3(256), 0, stream, pa, n);
    QSA_CK(hipGetLastError());
    size_t tmp = 0;
    QSA_CK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, ka, kb, pa, pb, (int)n, 0, 64, stream));
    size_t tmp2 = 0;
    QSA_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, static_cast<uint32_t*>(s_keep.p),
                                            static_cast<uint32_t*>(s_slot.p), (int)n, stream));
    QSA_CK(s_tmp.reserve(tmp > tmp2 ? tmp : tmp2));
    for (int w = stride - 1; w >= 0; --w) {   // least significant word first; radix sort is stable
        hipLaunchKernelGGL(gather_word, dim3(blocks), dim3(256), 0, stream, rows, pa, stride, w, ka, n);
        QSA_CK(hipGetLastError());
        size_t t = s_tmp.n;
        QSA_CK(hipcub::DeviceRadixSort::SortPairs(s_tmp.p, t, ka, kb, pa, pb, (int)n, 0, 64, stream));
        uint32_t* sw = pa; pa = pb; pb = sw;   // sorted permutation now in pa
    }
    auto* keep = static_cast<uint32_t*>(s_keep.p);
    auto* slot = static_cast<uint32_t*>(s_slot.p);
    hipLaunchKernelGGL(mark_distinct, dim3(blocks), dim3(256), 0, stream, rows, pa, stride, keep, n);
    QSA_CK(hipGetLastError());
    size_t t2 = s_tmp.n;
    QSA_CK(hipcub::DeviceScan::ExclusiveSum(s_tmp.p, t2, keep, slot, (int)n, stream));
    hipLaunchKernelGGL(scatter_kept, dim3(blocks), dim3(256), 0, stream, rows, pa, keep, slot, stride, out, n);
    QSA_CK(hipGetLastError());
    uint32_t last_slot = 0, last_keep = 0;
    QSA_CK(hipMemcpyAsync(&last_slot, slot + (n - 1), 4, hipMemcpyDeviceToHost, stream));
    QSA_CK(hipMemcpyAsync(&last_keep, keep + (n - 1), 4, hipMemcpyDeviceToHost, stream));
    QSA_CK(hipStreamSynchronize(stream));
    *kept = (uint64_t)last_slot + last_keep;
    return 0;
#undef QSA_CK
}

}  // namespace qsa
