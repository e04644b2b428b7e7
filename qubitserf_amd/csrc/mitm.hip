// Meet-in-the-middle test of one weight on the device: the independent certifier of the
// distance (the method of cpp/distfind/mitm.cpp:241-314).
//
// A word x of weight d splits into a left part (columns [0, nl)) of weight wL and a right
// part (columns [nl, n)) of weight d - wL. x is a codeword iff both parts have the same
// syndrome under the code's parity checks, and a logical iff, in addition, their detector
// fingerprints differ (or every nonzero codeword counts). For each split the smaller side's
// parts go into an open-addressing table keyed by syndrome, each slot keeping the min and
// max fingerprint seen (two distinct fingerprints exist iff min != max); the other side's
// parts probe it. Parts are generated in colex order, a run per thread, syndromes and
// fingerprints updated by XOR of the columns that change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qsf_hip.h"

namespace {

using u64 = std::uint64_t;
constexpr u64 kEmpty = ~0ull;        // syndromes have at most 63 bits
constexpr int kRun = 64;             // parts per thread

struct Side {
    const u64* syn;     // per column (of this side): syndrome bits
    const u64* log;     // per column: fingerprint bits
    int m, k;           // k-subsets of m columns
    u64 count;          // C(m, k)
};

__device__ __forceinline__ u64 mix(u64 x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    return x ^ (x >> 33);
}

// k-subset of index idx in colex order (mask), from the binomial table B[n * 65 + k]
__device__ u64 unrank(u64 idx, int k, int m, const u64* B) {
    u64 mask = 0;
    int cap = m;
    for (int i = k; i >= 1; --i) {
        int lo = i - 1, hi = cap - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (B[mid * 65 + i] <= idx) lo = mid; else hi = mid - 1;
        }
        mask |= 1ull << lo;
        idx -= B[lo * 65 + i];
        cap = lo;
    }
    return mask;
}

__device__ __forceinline__ u64 next_subset(u64 x) {   // colex successor, same popcount
    const u64 c = x & (~x + 1ull);
    const u64 r = x + c;
    return (((r ^ x) >> 2) >> __builtin_ctzll(c)) | r;
}

template <class F>
__device__ void for_parts(const Side& s, const u64* B, F f) {
    const u64 first = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * kRun;
    if (first >= s.count) return;
    const u64 last = min(s.count, first + kRun);
    u64 mask = s.k ? unrank(first, s.k, s.m, B) : 0ull;
    u64 syn = 0, log = 0;
    for (u64 mm = mask; mm; mm &= mm - 1) {
        syn ^= s.syn[__builtin_ctzll(mm)];
        log ^= s.log[__builtin_ctzll(mm)];
    }
    for (u64 i = first; i < last; ++i) {
        f(syn, log);
        if (i + 1 < last && s.k) {
            const u64 nm = next_subset(mask);
            for (u64 dd = nm ^ mask; dd; dd &= dd - 1) {
                syn ^= s.syn[__builtin_ctzll(dd)];
                log ^= s.log[__builtin_ctzll(dd)];
            }
            mask = nm;
        }
    }
}

// Symplectic parts (cpp/distfind/mitm.cpp:131-237): a coordinate is a qubit and the cost
// is the number of qubits touched. A part of weight k is k qubits of the side, each carrying
// Z, X or Y; its syndrome (fingerprint) is the XOR over those qubits of the Pauli's column:
// the z column, the x column, or both (precomputed per qubit). Part i = r 3^k + c: r ranks
// the qubit subset in colex order, the base-3 digits of c choose the Paulis.
struct SymSide {
    const u64* syn3;    // per qubit of this side: syndromes of Z, X, Y
    const u64* log3;    // per qubit: fingerprints of Z, X, Y
    int m, k;           // k of m qubits
    u64 p3;             // 3^k
    u64 count;          // C(m, k) 3^k
};

template <class F>
__device__ void for_parts(const SymSide& s, const u64* B, F f) {
    const u64 first = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * kRun;
    if (first >= s.count) return;
    const u64 last = min(s.count, first + kRun);
    const u64 r = first / s.p3;
    u64 c = first - r * s.p3;
    u64 mask = s.k ? unrank(r, s.k, s.m, B) : 0ull;
    for (u64 i = first; i < last; ++i) {
        u64 syn = 0, log = 0, cc = c;
        for (u64 mm = mask; mm; mm &= mm - 1) {
            const int q = __builtin_ctzll(mm);
            const int pauli = (int)(cc % 3);
            cc /= 3;
            syn ^= s.syn3[3 * q + pauli];
            log ^= s.log3[3 * q + pauli];
        }
        f(syn, log);
        if (++c == s.p3) {   // every Pauli choice of this subset done: the next subset
            c = 0;
            if (s.k && i + 1 < last) mask = next_subset(mask);
        }
    }
}

__global__ void clear_table(u64* keys, u64* lmin, u64* lmax, u64 cap) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (u64)gridDim.x * blockDim.x) {
        keys[i] = kEmpty;
        lmin[i] = ~0ull;
        lmax[i] = 0;
    }
}

template <class S>
__global__ void insert_parts(S s, const u64* B, u64* keys, unsigned long long* lmin, unsigned long long* lmax,
                             u64 cap) {
    for_parts(s, B, [&](u64 syn, u64 log) {
        u64 h = mix(syn) & (cap - 1);
        for (u64 probe = 0; probe < cap; ++probe) {   // the table is at most half full
            const u64 prev = atomicCAS((unsigned long long*)&keys[h], (unsigned long long)kEmpty,
                                       (unsigned long long)syn);
            if (prev == kEmpty || prev == syn) {
                atomicMin(&lmin[h], (unsigned long long)log);
                atomicMax(&lmax[h], (unsigned long long)log);
                return;
            }
            h = (h + 1) & (cap - 1);
        }
    });
}

template <class S>
__global__ void probe_parts(S s, const u64* B, const u64* keys, const u64* lmin, const u64* lmax, u64 cap,
                            int count_all, int* hit) {
    for_parts(s, B, [&](u64 syn, u64 log) {
        u64 h = mix(syn) & (cap - 1);
        for (u64 probe = 0; probe < cap; ++probe) {
            const u64 k = keys[h];
            if (k == kEmpty) return;
            if (k == syn) {
                if (count_all || lmin[h] != lmax[h] || lmin[h] != log) atomicOr(hit, 1);
                return;
            }
            h = (h + 1) & (cap - 1);
        }
    });
}

struct Mitm {
    std::mutex mu;
    u64* keys = nullptr;
    u64* lmin = nullptr;
    u64* lmax = nullptr;
    u64 cap = 0;
    u64* cols = nullptr;   // syn | log columns, n each
    u64* binom = nullptr;
    int* hit = nullptr;
    int ncols = 0;
    std::string err;
    double ms = 0;

    static u64 C(int n, int k) {
        if (k < 0 || k > n) return 0;
        long double c = 1;
        for (int i = 0; i < k; ++i) c = c * (n - i) / (i + 1);
        return c > 1.8e19L ? ~0ull : (u64)(c + 0.5L);
    }

    hipError_t ensure(u64 want_cap, int n) {
        hipError_t e = hipSuccess;
        if (!binom) {
            std::vector<u64> B(65 * 65, 0);
            for (int i = 0; i <= 64; ++i) {
                B[i * 65] = 1;
                for (int j = 1; j <= i; ++j) B[i * 65 + j] = B[(i - 1) * 65 + j - 1] + (j <= i - 1 ? B[(i - 1) * 65 + j] : 0);
            }
            if ((e = hipMalloc(&binom, B.size() * 8)) != hipSuccess) return e;
            if ((e = hipMemcpy(binom, B.data(), B.size() * 8, hipMemcpyHostToDevice)) != hipSuccess) return e;
            if ((e = hipMalloc(&hit, sizeof(int))) != hipSuccess) return e;
        }
        if (n > ncols) {
            if (cols) (void)hipFree(cols);
            cols = nullptr;
            if ((e = hipMalloc(&cols, (size_t)2 * n * 8)) != hipSuccess) return e;
            ncols = n;
        }
        if (want_cap > cap) {
            for (u64** p : {&keys, &lmin, &lmax})
                if (*p) { (void)hipFree(*p); *p = nullptr; }
            cap = 0;
            for (u64** p : {&keys, &lmin, &lmax})
                if ((e = hipMalloc(p, want_cap * 8)) != hipSuccess) return e;
            cap = want_cap;
        }
        return e;
    }
};

Mitm& mitm() {
    static Mitm* m = new Mitm();
    return *m;
}

unsigned blocks_for(u64 count) {
    const u64 threads = (count + kRun - 1) / kRun;
    return (unsigned)std::max<u64>(1, (threads + 255) / 256);
}

}  // namespace

extern "C" {

int qsf_hip_mitm_weight(const uint64_t* syn, const uint64_t* log, int n, int nl, int d, int count_all, int* hit_out) {
    Mitm& M = mitm();
    std::lock_guard<std::mutex> lk(M.mu);
    if (!syn || !log || !hit_out || n < 1 || nl < 0 || nl > n || n - nl > 64 || nl > 64 || d < 0 || d > n) {
        M.err = "qsf_hip_mitm_weight: bad arguments (each half must have at most 64 columns)";
        return 1;
    }
    *hit_out = 0;
    hipError_t e = hipSuccess;
    M.ms = 0;
    auto fail = [&](const char* what) {
        M.err = std::string("qsf_hip_mitm_weight: ") + what + ": " + hipGetErrorString(e);
        return 3;
    };
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) return fail("event");
    if ((e = M.ensure(0, n)) != hipSuccess) return fail("allocation");
    std::vector<u64> host(2 * (size_t)n);
    std::copy(syn, syn + n, host.begin());
    std::copy(log, log + n, host.begin() + n);
    if ((e = hipMemcpy(M.cols, host.data(), host.size() * 8, hipMemcpyHostToDevice)) != hipSuccess) return fail("copy");
    if ((e = hipMemset(M.hit, 0, sizeof(int))) != hipSuccess) return fail("memset");
    const int nr = n - nl;
    float total = 0;
    for (int wl = 0; wl <= d && !*hit_out; ++wl) {
        const int wr = d - wl;
        const u64 cl = Mitm::C(nl, wl), cr = Mitm::C(nr, wr);
        if (cl == 0 || cr == 0) continue;
        Side L{M.cols, M.cols + n, nl, wl, cl}, R{M.cols + nl, M.cols + n + nl, nr, wr, cr};
        const bool left_small = cl <= cr;
        const Side& ins = left_small ? L : R;
        const Side& prb = left_small ? R : L;
        if (ins.count > (1ull << 27) || prb.count > (1ull << 40)) {
            M.err = "qsf_hip_mitm_weight: weight split too large for the device table";
            return 4;
        }
        u64 cap = 1;
        while (cap < 2 * ins.count) cap <<= 1;
        if ((e = M.ensure(cap, n)) != hipSuccess) return fail("table allocation");
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(clear_table, dim3((unsigned)std::min<u64>(4096, (cap + 255) / 256)), dim3(256), 0, nullptr,
                           M.keys, M.lmin, M.lmax, cap);
        hipLaunchKernelGGL(insert_parts<Side>, dim3(blocks_for(ins.count)), dim3(256), 0, nullptr, ins, M.binom,
                           M.keys, (unsigned long long*)M.lmin, (unsigned long long*)M.lmax, cap);
        hipLaunchKernelGGL(probe_parts<Side>, dim3(blocks_for(prb.count)), dim3(256), 0, nullptr, prb, M.binom,
                           M.keys, M.lmin, M.lmax, cap, count_all, M.hit);
        (void)hipEventRecord(e1, nullptr);
        if ((e = hipGetLastError()) != hipSuccess) return fail("launch");
        if ((e = hipMemcpy(hit_out, M.hit, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess) return fail("read-back");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        total += ms;
    }
    M.ms = total;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int qsf_hip_mitm_weight_sym(const uint64_t* syn, const uint64_t* log, int q, int ql, int d, int* hit_out) {
    Mitm& M = mitm();
    std::lock_guard<std::mutex> lk(M.mu);
    if (!syn || !log || !hit_out || q < 1 || ql < 0 || ql > q || q - ql > 64 || ql > 64 || d < 0 || d > q) {
        M.err = "qsf_hip_mitm_weight_sym: bad arguments (each half must have at most 64 qubits)";
        return 1;
    }
    *hit_out = 0;
    hipError_t e = hipSuccess;
    M.ms = 0;
    auto fail = [&](const char* what) {
        M.err = std::string("qsf_hip_mitm_weight_sym: ") + what + ": " + hipGetErrorString(e);
        return 3;
    };
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) return fail("event");
    if ((e = M.ensure(0, 3 * q)) != hipSuccess) return fail("allocation");
    // per qubit j: Z = column j, X = column q + j, Y = both
    std::vector<u64> host(6 * (size_t)q);
    for (int j = 0; j < q; ++j) {
        host[3 * j] = syn[j];
        host[3 * j + 1] = syn[q + j];
        host[3 * j + 2] = syn[j] ^ syn[q + j];
        host[3 * q + 3 * j] = log[j];
        host[3 * q + 3 * j + 1] = log[q + j];
        host[3 * q + 3 * j + 2] = log[j] ^ log[q + j];
    }
    if ((e = hipMemcpy(M.cols, host.data(), host.size() * 8, hipMemcpyHostToDevice)) != hipSuccess) return fail("copy");
    if ((e = hipMemset(M.hit, 0, sizeof(int))) != hipSuccess) return fail("memset");
    auto parts = [](int m, int k, u64& p3) {   // C(m, k) 3^k, saturating; p3 = 3^k
        p3 = 1;
        for (int i = 0; i < k; ++i) p3 = p3 > (~0ull) / 3 ? ~0ull : p3 * 3;
        const u64 c = Mitm::C(m, k);
        return (c != 0 && p3 > (~0ull) / c) ? ~0ull : c * p3;
    };
    const int qr = q - ql;
    float total = 0;
    for (int wl = 0; wl <= d && !*hit_out; ++wl) {
        const int wr = d - wl;
        if (wl > ql || wr > qr) continue;
        u64 pl = 0, pr = 0;
        const u64 cl = parts(ql, wl, pl), cr = parts(qr, wr, pr);
        SymSide L{M.cols, M.cols + 3 * q, ql, wl, pl, cl};
        SymSide R{M.cols + 3 * ql, M.cols + 3 * q + 3 * ql, qr, wr, pr, cr};
        const bool left_small = cl <= cr;
        const SymSide& ins = left_small ? L : R;
        const SymSide& prb = left_small ? R : L;
        if (ins.count > (1ull << 27) || prb.count > (1ull << 40)) {
            M.err = "qsf_hip_mitm_weight_sym: weight split too large for the device table";
            return 4;
        }
        u64 cap = 1;
        while (cap < 2 * ins.count) cap <<= 1;
        if ((e = M.ensure(cap, 3 * q)) != hipSuccess) return fail("table allocation");
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(clear_table, dim3((unsigned)std::min<u64>(4096, (cap + 255) / 256)), dim3(256), 0, nullptr,
                           M.keys, M.lmin, M.lmax, cap);
        hipLaunchKernelGGL(insert_parts<SymSide>, dim3(blocks_for(ins.count)), dim3(256), 0, nullptr, ins, M.binom,
                           M.keys, (unsigned long long*)M.lmin, (unsigned long long*)M.lmax, cap);
        hipLaunchKernelGGL(probe_parts<SymSide>, dim3(blocks_for(prb.count)), dim3(256), 0, nullptr, prb, M.binom,
                           M.keys, M.lmin, M.lmax, cap, 0, M.hit);
        (void)hipEventRecord(e1, nullptr);
        if ((e = hipGetLastError()) != hipSuccess) return fail("launch");
        if ((e = hipMemcpy(hit_out, M.hit, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess) return fail("read-back");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        total += ms;
    }
    M.ms = total;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

const char* qsf_hip_mitm_last_error(void) { return mitm().err.c_str(); }

}  // extern "C"
