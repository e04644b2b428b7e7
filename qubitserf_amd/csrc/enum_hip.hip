// gfx950 enumeration engine: kernels, device-side solve cache, and the qsf_hip_* C ABI.
// The kernels and the codeword representation are described in kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qsf_hip.h"
#include "kernels.hpp"

namespace qsa {

// =====================================================================================
// device helpers
// =====================================================================================

__device__ __forceinline__ u64 bin(const u64* B, int bk, int n, int k) {
    return (k < 0 || n < 0 || k > n) ? 0ull : B[n * bk + k];
}

// Colex unrank: the idx-th k-subset of [0, m) in colex order (= increasing order of the
// subset's bitmask), as a bitmask. Element c_i (i = k..1, decreasing) is the largest c
// with C(c, i) <= remaining index (combinatorial number system).
__device__ __forceinline__ u64 colex_mask(u64 idx, int k, int m, const u64* B, int bk) {
    u64 mask = 0;
    int cap = m;
    for (int i = k; i >= 1; --i) {
        int lo = i - 1, hi = cap - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (B[mid * bk + i] <= idx) lo = mid; else hi = mid - 1;
        }
        mask |= 1ull << lo;
        idx -= (lo >= i) ? B[lo * bk + i] : 0ull;
        cap = lo;
    }
    return mask;
}

// Next larger integer with the same popcount (colex successor); x != 0.
__device__ __forceinline__ u64 colex_next(u64 x) {
    u64 c = x & (~x + 1ull);
    u64 r = x + c;
    return (((r ^ x) >> 2) >> __builtin_ctzll(c)) | r;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

template <int FD>
struct SElem {
    u32 w[FD];
    int off;
};

// =====================================================================================
// bz_tiled: one workgroup = one wavefront; persistent, tiles pulled from a global counter
// =====================================================================================

template <int FD, int RPL>
__global__ __launch_bounds__(64) void bz_tiled(TiledArgs a) {
    __shared__ u32 rows_fast[kMaxTiledK * FD];
    __shared__ SElem<FD> stream[kStreamMax];
    __shared__ u32 sfull[kMaxRW];

    const int lane = threadIdx.x;
    const int K = a.K, h = a.h, RW = a.RW, RD = a.RD, SD = a.SD;
    int bound = __builtin_amdgcn_readfirstlane(__hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    int cur_g = -1;

    for (;;) {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(a.next_tile, 1ull);
        t = __shfl(t, 0, 64);
        if (t >= a.total_tiles) break;
        {
            int gb = __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bound = min(bound, __builtin_amdgcn_readfirstlane(gb));
        }
        if (bound <= a.floor) break;

        // locate the rectangle holding tile t (rects sorted by tile_base)
        int lo = 0, hi = a.nrects - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (a.rects[mid].tile_base <= t) lo = mid; else hi = mid - 1;
        }
        const Rect R = a.rects[lo];
        const u64 local = t - R.tile_base;
        const u64 bx = local / R.stream_chunks;
        const u64 by = local - bx * R.stream_chunks;
        const int g = R.g;

        const int lk = R.lanes_on_a ? R.ka : R.kb;
        const int lbase = R.lanes_on_a ? 0 : h;
        const int lm = R.lanes_on_a ? h : K - h;
        const u64 lflags = a.flags[2 * g + (R.lanes_on_a ? 0 : 1)];
        const int sk = R.lanes_on_a ? R.kb : R.ka;
        const int sbase = R.lanes_on_a ? h : 0;
        const int sm = R.lanes_on_a ? K - h : h;
        const u64 sflags = a.flags[2 * g + (R.lanes_on_a ? 1 : 0)];
        const u32* gtbl = a.tbl + (size_t)g * K * RW;

        __syncthreads();   // previous tile's readers of LDS are done
        if (g != cur_g) {
            for (int i = lane; i < K * FD; i += 64) rows_fast[i] = gtbl[(i / FD) * RW + (i % FD)];
            cur_g = g;
        }

        // ---- stream chunk -> LDS (each lane a contiguous run, colex successor steps)
        const u64 c0 = by * (u64)a.S;
        const int len = (int)min((u64)a.S, R.n_stream - c0);
        __syncthreads();   // rows_fast visible
        {
            const int per = (len + 63) >> 6;
            const int i0 = lane * per;
            const int i1 = min(len, i0 + per);
            if (i0 < i1) {
                u64 m = colex_mask(c0 + (u64)i0, sk, sm, a.binom, a.bk);
                for (int i = i0; i < i1; ++i) {
                    SElem<FD> e;
#pragma unroll
                    for (int f = 0; f < FD; ++f) e.w[f] = 0;
                    u64 mm = m;
                    while (mm) {
                        int b = __builtin_ctzll(mm);
                        mm &= mm - 1;
#pragma unroll
                        for (int f = 0; f < FD; ++f) e.w[f] ^= rows_fast[(sbase + b) * FD + f];
                    }
                    e.off = __popcll(m & sflags);
                    stream[i] = e;
                    if (i + 1 < i1) m = colex_next(m);
                }
            }
        }

        // ---- register tile: RPL consecutive list elements per lane
        u32 xw[RPL][FD];
        int xoff[RPL];
        const u64 b0 = bx * (u64)(64 * RPL) + (u64)lane * RPL;
        {
            u64 m = (b0 < R.n_lane) ? colex_mask(b0, lk, lm, a.binom, a.bk) : 0ull;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const bool valid = b0 + k < R.n_lane;
#pragma unroll
                for (int f = 0; f < FD; ++f) xw[k][f] = 0;
                if (valid) {
                    u64 mm = m;
                    while (mm) {
                        int b = __builtin_ctzll(mm);
                        mm &= mm - 1;
#pragma unroll
                        for (int f = 0; f < FD; ++f) xw[k][f] ^= rows_fast[(lbase + b) * FD + f];
                    }
                    xoff[k] = __popcll(m & lflags);
                    if (b0 + k + 1 < R.n_lane && lk > 0) m = colex_next(m);
                } else {
                    xoff[k] = kBigOffset;
                }
            }
        }
        __syncthreads();   // stream chunk visible

        // ---- hot loop: every (lane element, stream element) pair
        for (int s = 0; s < len; ++s) {
            const SElem<FD> e = stream[s];
            int acc = kBigOffset;
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                int tk = xoff[k];
#pragma unroll
                for (int f = 0; f < FD; ++f) tk += __popc(xw[k][f] ^ e.w[f]);
                acc = min(acc, tk);
            }
            const int thr = bound - e.off;
            if (__builtin_expect(__ballot(acc < thr) != 0ull, 0)) {
                // ---- slow path: exact weight and logical test of the light candidates
                const u64 smask = colex_mask(c0 + (u64)s, sk, sm, a.binom, a.bk);
                if (lane < RW) {
                    u32 v = 0;
                    u64 mm = smask;
                    while (mm) {
                        int b = __builtin_ctzll(mm);
                        mm &= mm - 1;
                        v ^= gtbl[(sbase + b) * RW + lane];
                    }
                    sfull[lane] = v;
                }
                __syncthreads();
                u32 cand = 0;
#pragma unroll
                for (int k = 0; k < RPL; ++k) {
                    int tk = xoff[k];
#pragma unroll
                    for (int f = 0; f < FD; ++f) tk += __popc(xw[k][f] ^ e.w[f]);
                    if (tk < thr) cand |= 1u << k;
                }
                int mine = QSF_WEIGHT_NONE;
                while (cand) {
                    const int k = __builtin_ctz(cand);
                    cand &= cand - 1;
                    const u64 lmask = colex_mask(b0 + k, lk, lm, a.binom, a.bk);
                    int w = __popcll(lmask & lflags) + e.off;
                    for (int f = 0; f < RD && w < bound; ++f) {
                        u32 v = sfull[f];
                        u64 mm = lmask;
                        while (mm) {
                            int b = __builtin_ctzll(mm);
                            mm &= mm - 1;
                            v ^= gtbl[(lbase + b) * RW + f];
                        }
                        w += __popc(v);
                    }
                    if (w < bound && w < mine) {
                        bool logical = (SD == 0);
                        for (int f = 0; f < SD && !logical; ++f) {
                            u32 v = sfull[RD + f];
                            u64 mm = lmask;
                            while (mm) {
                                int b = __builtin_ctzll(mm);
                                mm &= mm - 1;
                                v ^= gtbl[(lbase + b) * RW + RD + f];
                            }
                            logical = v != 0;
                        }
                        if (logical) mine = w;
                    }
                }
                const int wm = __builtin_amdgcn_readfirstlane(wave_min(mine));
                if (wm < bound) {
                    bound = wm;
                    if (lane == 0) atomicMin(a.best, wm);
                }
                __syncthreads();
            }
        }
        if (bound <= a.floor) break;
    }
}

// =====================================================================================
// bz_rowsweep: full rows, lexicographic runs (small levels, K > 128, collect mode)
// =====================================================================================

__global__ __launch_bounds__(256) void bz_rowsweep(SweepArgs a) {
    const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 per_g = (a.total + a.chunk - 1) / a.chunk;
    const u64 g = tid / per_g;
    if (g >= (u64)a.ngamma) return;
    const u64 start = (tid - g * per_g) * a.chunk;
    if (start >= a.total) return;
    const u64 end = min(a.total, start + a.chunk);
    const int K = a.K, D = a.d, M = D - 1, st = a.stride;
    const u64* G = a.gamma + (size_t)g * K * st;

    int pos[kMaxD];
    u64 base[16], cw[16];
    {   // lexicographic unrank of `start`
        u64 r = start;
        int x = 0;
        for (int i = 0; i < D; ++i)
            for (;;) {
                u64 cnt = bin(a.binom, a.bk, K - 1 - x, D - 1 - i);
                if (r < cnt) { pos[i] = x++; break; }
                r -= cnt;
                ++x;
            }
    }
    for (int w = 0; w < st; ++w) base[w] = 0;
    for (int i = 0; i < M; ++i)
        for (int w = 0; w < st; ++w) base[w] ^= G[(size_t)pos[i] * st + w];

    int bound = a.collect ? a.keep + 1 : __hip_atomic_load(a.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int mine = QSF_WEIGHT_NONE;
    int last = pos[D - 1];
    u64 done = 0;
    const u64 need = end - start;
    while (done < need) {
        int hi = K;
        if ((u64)(K - last) > need - done) hi = last + (int)(need - done);
        const int first = last;
        for (; last < hi; ++last) {
            int wt = 0;
            for (int w = 0; w < st; ++w) {
                cw[w] = base[w] ^ G[(size_t)last * st + w];
                wt += __popcll(cw[w]);
            }
            if (wt >= bound) continue;
            if (a.collect) {
                if (wt == 0) continue;
                unsigned long long slot = atomicAdd(a.out_rows, 1ull);
                if (slot < a.out_cap)
                    for (int w = 0; w < st; ++w) a.out[slot * st + w] = cw[w];
                continue;
            }
            bool logical = a.kcheck == 0;
            for (int c = 0; c < a.kcheck && !logical; ++c) {
                u64 acc = 0;
                for (int w = 0; w < st; ++w) acc ^= a.check[(size_t)c * st + w] & cw[w];
                logical = __popcll(acc) & 1;
            }
            if (logical) { bound = wt; mine = wt; }
        }
        done += (u64)(hi - first);
        if (done >= need) break;
        int j = M - 1;
        while (j >= 0 && pos[j] == K - D + j) --j;
        if (j < 0) break;
        for (int t = j; t < M; ++t)
            for (int w = 0; w < st; ++w) base[w] ^= G[(size_t)pos[t] * st + w];
        ++pos[j];
        for (int t = j + 1; t < M; ++t) pos[t] = pos[t - 1] + 1;
        for (int t = j; t < M; ++t)
            for (int w = 0; w < st; ++w) base[w] ^= G[(size_t)pos[t] * st + w];
        last = pos[M - 1] + 1;
    }
    if (!a.collect && mine < QSF_WEIGHT_NONE) atomicMin(a.best, mine);
}

// =====================================================================================
// host side
// =====================================================================================

namespace {

constexpr u64 kBinInf = ~0ull;

struct HipError {
    int code;
    std::string what;
};

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw HipError{3, std::string(#expr) + ": " + hipGetErrorString(e_)};      \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void reserve(size_t count) {
        if (count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        HIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

u64 sat_add(u64 a, u64 b) { return (a == kBinInf || b == kBinInf || a > kBinInf - b) ? kBinInf : a + b; }
u64 sat_mul(u64 a, u64 b) {
    if (a == 0 || b == 0) return 0;
    if (a == kBinInf || b == kBinInf) return kBinInf;
    unsigned __int128 p = (unsigned __int128)a * b;
    return p > kBinInf ? kBinInf : (u64)p;
}

// Everything the device needs for one solve, built once per buffers_key.
struct Solve {
    u64 key = 0;
    int n = 0, stride = 0, K = 0, ngt = 0, kcheck = 0;
    int h = 0, RD = 0, SD = 0, RW = 0, bk = 0;
    std::vector<u64> binom;       // (K+1) x bk
    DevBuf<u64> gamma, check, dbinom, flags;
    DevBuf<u32> tbl;
    u64 C(int nn, int kk) const {
        if (kk < 0 || nn < 0 || kk > nn) return 0;
        if (kk >= bk) return kBinInf;
        return binom[(size_t)nn * bk + kk];
    }
};

struct Engine {
    std::mutex mu;
    bool probed = false, ok = false;
    std::string probe_msg;
    int device = 0, cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::list<std::unique_ptr<Solve>> cache;   // most recent first
    DevBuf<Rect> rects;
    DevBuf<int> best;
    DevBuf<unsigned long long> counter;
    DevBuf<u64> out;
    qsf_hip_stats stats{};
    std::string err;
    int policy = 0;   // 0 auto, 1 row sweep, 2 tiled where it applies

    bool probe() {
        if (probed) return ok;
        probed = true;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) { probe_msg = "no HIP device"; return ok = false; }
        if (hipGetDevice(&device) != hipSuccess) { probe_msg = "hipGetDevice failed"; return ok = false; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { probe_msg = "hipGetDeviceProperties failed"; return ok = false; }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            probe_msg = std::string("device is ") + prop.gcnArchName + ", kernels are built for gfx950";
            return ok = false;
        }
        cus = prop.multiProcessorCount;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess) {
            probe_msg = "stream/event creation failed";
            return ok = false;
        }
        return ok = true;
    }

    Solve& solve_for(const qsf_enum_plan& p) {
        for (auto it = cache.begin(); it != cache.end(); ++it) {
            Solve& s = **it;
            if (p.buffers_key != 0 && s.key == p.buffers_key && s.n == p.n && s.K == p.K &&
                s.ngt == p.num_gamma_total && s.kcheck == p.kcheck && s.stride == p.stride) {
                cache.splice(cache.begin(), cache, it);
                return s;
            }
        }
        auto s = std::make_unique<Solve>();
        build(*s, p);
        cache.push_front(std::move(s));
        while (cache.size() > 6) cache.pop_back();
        return *cache.front();
    }

    void build(Solve& s, const qsf_enum_plan& p) {
        s.key = p.buffers_key;
        s.n = p.n; s.stride = p.stride; s.K = p.K; s.ngt = p.num_gamma_total; s.kcheck = p.kcheck;
        const int K = p.K, st = p.stride, n = p.n;
        s.h = K / 2;
        s.bk = std::min(K, kMaxD) + 1;
        s.binom.assign((size_t)(K + 1) * s.bk, 0);
        for (int i = 0; i <= K; ++i) {
            s.binom[(size_t)i * s.bk] = 1;
            for (int j = 1; j < s.bk && j <= i; ++j)
                s.binom[(size_t)i * s.bk + j] = sat_add(s.binom[(size_t)(i - 1) * s.bk + j - 1],
                                                        j <= i - 1 ? s.binom[(size_t)(i - 1) * s.bk + j] : 0);
        }
        // residual form of every generator
        std::vector<std::vector<int>> resid(s.ngt);
        std::vector<std::vector<char>> flagged(s.ngt, std::vector<char>(K, 0));
        int maxres = 0;
        std::vector<int> cnt(n), who(n);
        for (int g = 0; g < s.ngt; ++g) {
            const u64* G = p.gamma + (size_t)g * K * st;
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int r = 0; r < K; ++r)
                for (int j = 0; j < n; ++j)
                    if ((G[(size_t)r * st + (j >> 6)] >> (j & 63)) & 1) { ++cnt[j]; who[j] = r; }
            for (int j = 0; j < n; ++j) {
                if (cnt[j] == 1 && !flagged[g][who[j]]) flagged[g][who[j]] = 1;
                else resid[g].push_back(j);
            }
            maxres = std::max(maxres, (int)resid[g].size());
        }
        s.RD = std::max(1, (maxres + 31) / 32);
        s.SD = (p.kcheck + 31) / 32;
        s.RW = s.RD + s.SD;
        std::vector<u32> tbl((size_t)s.ngt * K * s.RW, 0);
        std::vector<u64> fl((size_t)s.ngt * 2, 0);
        for (int g = 0; g < s.ngt; ++g) {
            const u64* G = p.gamma + (size_t)g * K * st;
            for (int r = 0; r < K; ++r) {
                const u64* row = G + (size_t)r * st;
                u32* out = tbl.data() + ((size_t)g * K + r) * s.RW;
                for (size_t t = 0; t < resid[g].size(); ++t) {
                    int j = resid[g][t];
                    if ((row[j >> 6] >> (j & 63)) & 1) out[t >> 5] |= 1u << (t & 31);
                }
                for (int c = 0; c < p.kcheck; ++c) {
                    const u64* ch = p.check + (size_t)c * st;
                    u64 acc = 0;
                    for (int w = 0; w < st; ++w) acc ^= ch[w] & row[w];
                    if (__builtin_popcountll(acc) & 1) out[s.RD + (c >> 5)] |= 1u << (c & 31);
                }
                if (flagged[g][r]) {
                    if (r < s.h) fl[2 * g] |= 1ull << r;
                    else fl[2 * g + 1] |= 1ull << (r - s.h);
                }
            }
        }
        s.gamma.reserve((size_t)s.ngt * K * st);
        HIPCHK(hipMemcpyAsync(s.gamma.p, p.gamma, (size_t)s.ngt * K * st * 8, hipMemcpyHostToDevice, stream));
        s.check.reserve((size_t)std::max(1, p.kcheck) * st);
        if (p.kcheck)
            HIPCHK(hipMemcpyAsync(s.check.p, p.check, (size_t)p.kcheck * st * 8, hipMemcpyHostToDevice, stream));
        s.dbinom.reserve(s.binom.size());
        HIPCHK(hipMemcpyAsync(s.dbinom.p, s.binom.data(), s.binom.size() * 8, hipMemcpyHostToDevice, stream));
        s.tbl.reserve(tbl.size());
        HIPCHK(hipMemcpyAsync(s.tbl.p, tbl.data(), tbl.size() * 4, hipMemcpyHostToDevice, stream));
        s.flags.reserve(fl.size());
        HIPCHK(hipMemcpyAsync(s.flags.p, fl.data(), fl.size() * 8, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));   // host vectors above go out of scope
    }

    // ---- min mode -------------------------------------------------------------------
    int enumerate(const qsf_enum_plan& p, int floor, int* best_out) {
        stats = qsf_hip_stats{};
        Solve& s = solve_for(p);
        const u64 per_g = s.C(p.K, p.d);
        stats.codewords = sat_mul(per_g, (u64)p.num_gamma);
        if (per_g == 0 || p.num_gamma == 0) { *best_out = p.current_best; return 0; }
        if (stats.codewords > (1ull << 62)) throw HipError{4, "level has more than 2^62 codewords"};
        best.reserve(1);
        counter.reserve(1);
        HIPCHK(hipMemcpyAsync(best.p, &p.current_best, sizeof(int), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemsetAsync(counter.p, 0, sizeof(unsigned long long), stream));

        const bool can_tile = p.K <= kMaxTiledK && s.RW <= kMaxRW && p.d <= 32;
        const bool tiled = can_tile && (policy == 2 || (policy == 0 && stats.codewords >= (1ull << 16)));
        if (tiled) launch_tiled(s, p, floor);
        else launch_sweep(s, p, false, 0, 0);

        int result = 0;
        HIPCHK(hipMemcpyAsync(&result, best.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.kernel_ms = ms;
        *best_out = std::min(result, p.current_best);
        return 0;
    }

    void launch_tiled(Solve& s, const qsf_enum_plan& p, int floor) {
        const int K = p.K, d = p.d, h = s.h;
        const int FD = s.RD >= 2 ? 2 : 1;
        const int RPL = 16;
        const int S = kStreamMax;
        const u64 lane_tile = 64ull * RPL;
        std::vector<Rect> rv;
        u64 tiles = 0;
        for (int g = 0; g < p.num_gamma; ++g)
            for (int ka = std::max(0, d - (K - h)); ka <= std::min(d, h); ++ka) {
                const int kb = d - ka;
                const u64 na = s.C(h, ka), nb = s.C(K - h, kb);
                if (na == 0 || nb == 0) continue;
                Rect r{};
                r.g = g; r.ka = ka; r.kb = kb;
                r.lanes_on_a = na >= nb ? 1 : 0;
                r.n_lane = r.lanes_on_a ? na : nb;
                r.n_stream = r.lanes_on_a ? nb : na;
                r.lane_blocks = (r.n_lane + lane_tile - 1) / lane_tile;
                r.stream_chunks = (r.n_stream + S - 1) / S;
                r.tile_base = tiles;
                tiles += r.lane_blocks * r.stream_chunks;
                rv.push_back(r);
            }
        rects.reserve(rv.size());
        HIPCHK(hipMemcpyAsync(rects.p, rv.data(), rv.size() * sizeof(Rect), hipMemcpyHostToDevice, stream));
        TiledArgs a{};
        a.tbl = s.tbl.p; a.flags = s.flags.p; a.rects = rects.p; a.nrects = (int)rv.size();
        a.total_tiles = tiles; a.binom = s.dbinom.p; a.bk = s.bk;
        a.K = K; a.h = h; a.RD = s.RD; a.SD = s.SD; a.RW = s.RW; a.S = S;
        a.best = best.p; a.next_tile = counter.p; a.floor = floor;
        const u64 grid = std::min<u64>(tiles, (u64)std::max(1, cus) * 16);
        stats.kernel = 2;
        stats.tiles = (int)std::min<u64>(tiles, 0x7fffffff);
        HIPCHK(hipEventRecord(ev0, stream));
        if (grid > 0) {
            if (FD == 1) hipLaunchKernelGGL((bz_tiled<1, 16>), dim3((unsigned)grid), dim3(64), 0, stream, a);
            else hipLaunchKernelGGL((bz_tiled<2, 16>), dim3((unsigned)grid), dim3(64), 0, stream, a);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(ev1, stream));
        (void)RPL;
    }

    void launch_sweep(Solve& s, const qsf_enum_plan& p, bool collect, int keep, u64 cap) {
        if (p.stride > 16) throw HipError{4, "row-sweep kernel supports n <= 1024"};
        if (p.d > kMaxD) throw HipError{4, "weight level above 64"};
        SweepArgs a{};
        a.gamma = s.gamma.p; a.check = s.check.p; a.kcheck = p.kcheck; a.stride = p.stride;
        a.K = p.K; a.d = p.d; a.ngamma = p.num_gamma; a.binom = s.dbinom.p; a.bk = s.bk;
        a.total = s.C(p.K, p.d);
        const u64 target_threads = (u64)std::max(1, cus) * 2048;
        u64 chunk = std::max<u64>(64, (a.total * (u64)p.num_gamma + target_threads - 1) / target_threads);
        a.chunk = chunk;
        a.best = best.p; a.collect = collect ? 1 : 0; a.keep = keep;
        a.out_rows = counter.p; a.out = out.p; a.out_cap = cap;
        const u64 threads = ((a.total + chunk - 1) / chunk) * (u64)p.num_gamma;
        const u64 blocks = (threads + 255) / 256;
        stats.kernel = 1;
        HIPCHK(hipEventRecord(ev0, stream));
        if (blocks > 0) {
            hipLaunchKernelGGL(bz_rowsweep, dim3((unsigned)blocks), dim3(256), 0, stream, a);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(ev1, stream));
    }

    // ---- collect mode ---------------------------------------------------------------
    int collect(const qsf_enum_plan& p, int keep, uint64_t** hits, int64_t* rows) {
        stats = qsf_hip_stats{};
        Solve& s = solve_for(p);
        const u64 per_g = s.C(p.K, p.d);
        stats.codewords = sat_mul(per_g, (u64)p.num_gamma);
        *hits = nullptr;
        *rows = 0;
        if (per_g == 0 || p.num_gamma == 0) return 0;
        if (stats.codewords > (1ull << 62)) throw HipError{4, "level has more than 2^62 codewords"};
        best.reserve(1);
        counter.reserve(1);
        u64 cap = 1 << 16;
        for (;;) {
            out.reserve(cap * (u64)p.stride);
            HIPCHK(hipMemsetAsync(counter.p, 0, sizeof(unsigned long long), stream));
            launch_sweep(s, p, true, keep, cap);
            unsigned long long got = 0;
            HIPCHK(hipMemcpyAsync(&got, counter.p, sizeof(got), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            if (got > cap) { cap = got; continue; }   // overflow: rerun with room for all
            u64* h = (u64*)std::malloc(std::max<u64>(1, got) * p.stride * 8);
            if (!h) throw HipError{3, "host allocation failed"};
            if (got) HIPCHK(hipMemcpy(h, out.p, got * p.stride * 8, hipMemcpyDeviceToHost));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            stats.kernel_ms = ms;
            *hits = h;
            *rows = (int64_t)got;
            return 0;
        }
    }
};

Engine& engine() {
    static Engine* e = new Engine();   // never destroyed: device teardown order at exit is unsafe
    return *e;
}

bool plan_ok(const qsf_enum_plan* p, std::string& why) {
    if (!p) { why = "null plan"; return false; }
    if (p->K < 0 || p->n <= 0 || p->stride != (p->n + 63) / 64) { why = "inconsistent n/stride/K"; return false; }
    if (p->num_gamma < 0 || p->num_gamma_total < p->num_gamma) { why = "num_gamma > num_gamma_total"; return false; }
    if (p->num_gamma_total > 0 && p->K > 0 && !p->gamma) { why = "null gamma"; return false; }
    if (p->kcheck < 0 || (p->kcheck > 0 && !p->check)) { why = "bad check rows"; return false; }
    if (p->d < 1 || p->d > p->K) { why = "weight level outside [1, K]"; return false; }
    return true;
}

}  // namespace
}  // namespace qsa

using qsa::engine;

extern "C" {

int qsf_hip_available(void) {
    auto& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    if (e.probe()) return 1;
    e.err = "qsf_hip_available: " + e.probe_msg;
    return 0;
}

static int run_guarded(const char* what, const std::function<int(qsa::Engine&)>& f) {
    auto& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    if (!e.probe()) { e.err = std::string(what) + ": " + e.probe_msg; return 2; }
    try {
        e.err.clear();
        return f(e);
    } catch (const qsa::HipError& x) {
        e.err = std::string(what) + ": " + x.what;
        return x.code;
    } catch (const std::exception& x) {
        e.err = std::string(what) + ": " + x.what();
        return 3;
    }
}

int qsf_hip_enumerate_floor(const qsf_enum_plan* plan, int floor, int* best_out) {
    std::string why;
    if (!best_out || !qsa::plan_ok(plan, why)) {
        engine().err = "qsf_hip_enumerate: " + (best_out ? why : std::string("null best_out"));
        return 1;
    }
    return run_guarded("qsf_hip_enumerate", [&](qsa::Engine& e) { return e.enumerate(*plan, floor, best_out); });
}

int qsf_hip_enumerate(const qsf_enum_plan* plan, int* best_out) {
    return qsf_hip_enumerate_floor(plan, -1, best_out);
}

int qsf_hip_collect(const qsf_enum_plan* plan, int keep_weight, uint64_t** hits, int64_t* rows) {
    std::string why;
    if (!hits || !rows || !qsa::plan_ok(plan, why)) {
        engine().err = "qsf_hip_collect: " + (hits && rows ? why : std::string("null output"));
        return 1;
    }
    return run_guarded("qsf_hip_collect", [&](qsa::Engine& e) { return e.collect(*plan, keep_weight, hits, rows); });
}

void qsf_hip_free(void* p) { std::free(p); }

int qsf_hip_set_kernel_policy(int policy) {
    auto& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    int old = e.policy;
    e.policy = (policy >= 0 && policy <= 2) ? policy : 0;
    return old;
}

void qsf_hip_last_stats(qsf_hip_stats* out) {
    if (out) *out = engine().stats;
}

const char* qsf_hip_last_error(void) { return engine().err.c_str(); }

}  // extern "C"
