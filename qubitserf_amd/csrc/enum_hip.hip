// gfx950 enumeration engine: kernels, device-side solve cache, and the qsf_hip_* C ABI.
// The kernels and the codeword representation are described in kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <atomic>
#include <string>
#include <vector>

#include "../../include/qsf_hip.h"
#include "collect.hpp"
#include "elim.hpp"
#include "kernels.hpp"
#include "masks.hpp"
#include "pool.hpp"
#include "tuning.hpp"

// Device memory the library owns for the life of the process (one copy per GPU it runs on).
// Host code never uses these names as addresses: at set-up the engine reads each array's
// device address with hipGetSymbolAddress and passes that pointer to kernels and copies.
//  * tiled_binom: saturating C(n, k), n <= kMaxTiledK, k <= kMaxD, read by the tiled kernels.
//  * scratch: all device memory of qsf_hip_enumerate_floor_noalloc (bound and counters, the
//    solve's tables, one run's block), so that entry point allocates nothing.
constexpr size_t kScratchBytes = 16u << 20;
static __device__ __align__(256) unsigned long long tiled_binom[(qsa::kMaxTiledK + 1) * (qsa::kMaxD + 1)];
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace qsa {

// =====================================================================================
// device helpers
// =====================================================================================

__device__ __forceinline__ u64 bin(const u64* B, int bk, int n, int k) {
    return (k < 0 || n < 0 || k > n) ? 0ull : B[n * bk + k];
}

// Values every lane already agrees on, moved to scalar registers so that loops and branches
// on them are scalar (the compiler cannot see uniformity through vector loads).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
// (readfirstlane returns int: each half goes through u32 so the low half is not sign-extended)
__device__ __forceinline__ u64 uni64(u64 v) {
    return ((u64)(u32)__builtin_amdgcn_readfirstlane((u32)(v >> 32)) << 32) |
           (u64)(u32)__builtin_amdgcn_readfirstlane((u32)v);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

// Collect mode: one atomic per wave for the rows its active lanes append. Every active
// lane calls it; the lanes with `hit` get consecutive slots (lane order), the others ~0.
__device__ __forceinline__ unsigned long long wave_append(unsigned long long* rows, bool hit) {
    const u64 m = __ballot(hit);
    if (m == 0ull) return ~0ull;
    const int leader = __builtin_ctzll(m);
    const int lane = (int)(threadIdx.x & 63);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(rows, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    return hit ? base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull)) : ~0ull;
}

// popcount(x) + acc in one instruction. The compiler forms popcount(a) + popcount(b) + c
// as two popcounts into 0 and a three-input add; chained through the popcount's own
// addend, the hot test's distance over FD dwords is FD XORs and FD popcounts.
__device__ __forceinline__ int popc_acc(u32 x, int acc) {
    int r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

template <int FD>
struct SElem {
    u32 w[FD];
    int off;
};

// =====================================================================================
// bz_tiled: one workgroup = one wavefront; persistent, tiles pulled from a global counter
// =====================================================================================

// FD residual dwords are tested in the hot loop; XD >= FD dwords are held per codeword and
// re-tested (the "medium" test) before a stream step drops into the exact slow path.
// W: words per list subset mask (1: lists of up to 64 rows, K <= 128; 2: up to 128 rows,
// K <= 256). The masks enter only where list elements are made or re-made, not the hot loop.
template <int FD, int XD, int RPL, int W>
__global__ __launch_bounds__(64) void bz_tiled(TiledArgs a) {
    using Mask = typename std::conditional<W == 2, M2, u64>::type;
    __shared__ u32 rows_fast[(W == 2 ? kMaxTiledK : kMaxMatrixK) * XD];
    __shared__ SElem<XD> stream[kStreamMax];
    __shared__ u32 sfull[kMaxRW];
    __shared__ Mask run_start[64];           // colex mask of the first stream element of each lane's run
    __shared__ int run_fix[64];              // and its fixed set (a rectangle with a list of them)
    extern __shared__ u32 rows_full[];       // whole residual+syndrome table when it fits (dynamic size)
    __shared__ u32 sbin[kSmallBinRows * kSmallBinCols];   // the unrank's binomials C(n < 64, k <= 8)

    const int lane = threadIdx.x;
    for (int i = lane; i < kSmallBinRows * kSmallBinCols; i += 64)
        sbin[i] = (u32)a.binom[(i / kSmallBinCols) * a.bk + i % kSmallBinCols];
    __syncthreads();
    int cur_job = -1, cur_g = -1;
    int K = 0, RW = 0, RD = 0, SD = 0, floor = -1, bound = 0;
    const TJob* J = nullptr;
    unsigned slow_entries = 0, medium_entries = 0;
    unsigned job_hits = 0;   // this wave's medium entries on the current job (flushed to J->hits)

    for (;;) {
        unsigned long long t0 = 0;
        if (lane == 0) t0 = atomicAdd(a.next_tile, 1ull);
        // readfirstlane (lane 0 is active) makes the tile index, and everything derived from
        // it, provably wave-uniform: scalar registers and scalar loops, no exec masking
        const u64 t = uni64(t0);
        if (t >= a.total_tiles) break;

        const Rect& Rr = a.rects[uni(find_rect(a, t))];   // the rectangle holding tile t
        const int job = uni(Rr.job);
        if (job != cur_job) {   // the problem's parameters, and its running bound
            if (lane == 0 && job_hits) atomicAdd(J->hits, (unsigned long long)job_hits);
            job_hits = 0;
            J = a.jobs + job;
            K = uni(J->K); RW = uni(J->RW); RD = uni(J->RD); SD = uni(J->SD); floor = uni(J->floor);
            bound = a.collect ? a.keep + 1
                              : uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            cur_job = job;
            cur_g = -1;
        }
        if (!a.collect) {
            bound = min(bound, uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
            if (bound <= floor) {
                if (a.njobs == 1) break;
                continue;   // this problem is settled; its remaining tiles are skipped
            }
            // another level of this problem, in the same launch, settled it
            if (J->choice) {   // one of the level's candidate variants: run only the one picked
                int sb = bound;
                if (J->shared != J->best)
                    sb = min(sb, uni(__hip_atomic_load(J->shared, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                int c = 0;
                if (lane == 0) c = pick_variant(J, sb);
                if (uni(c) != J->variant) continue;
            }
            if (J->shared != J->best) {
                const int sb = uni(__hip_atomic_load(J->shared, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (sb <= floor) {
                    if (lane == 0) atomicAdd(J->skipped, 1ull);
                    continue;
                }
                // The level loop's bound here is min(bound, shared): the levels before this
                // one (earlier launch phases) are complete, and in this phase only this level
                // moves the shared bound. Pruning on it keeps every level's result the one
                // the level loop would get.
                bound = min(bound, sb);
            }
        }
        struct {
            u64 tile_base, stream_groups, stream_chunks, n_lane, n_stream;
            int fixed, nfix;
        } R = {uni64(Rr.tile_base), uni64(Rr.stream_groups), uni64(Rr.stream_chunks), uni64(Rr.n_lane),
               uni64(Rr.n_stream), uni(Rr.fixed), uni(Rr.nfix)};
        // a list of fixed sets: the stream is (fixed set, subset) pairs, ns subsets each
        const bool flist = R.nfix > 1;
        const u64 ns = flist ? R.n_stream / (u64)R.nfix : R.n_stream;
        const u64 local = t - R.tile_base;
        const u64 bx = local / R.stream_groups;
        const u64 bg = local - bx * R.stream_groups;
        const int g = uni(Rr.g);

        const int lk = uni(Rr.lk), lbase = uni(Rr.lbase), lm = uni(Rr.lm);
        const int sk = uni(Rr.sk), sbase = uni(Rr.sbase), sm = uni(Rr.sm);
        const Mask lflags = range_mask<Mask>(J->flags + 4 * g, lbase, lm);
        const Mask sflags = range_mask<Mask>(J->flags + 4 * g, sbase, sm);
        const u32* fx = R.fixed >= 0 ? J->fixed + (size_t)R.fixed * (RW + 1) : nullptr;
        u32 fw[XD];
#pragma unroll
        for (int f = 0; f < XD; ++f) fw[f] = (fx && f < RD) ? fx[f] : 0u;
        const int foff = fx ? (int)fx[RW] : 0;
        const u32* gtbl = J->tbl + (size_t)g * K * RW;

        const bool full_in_lds = K * RW <= a.lds_full;
        const u32* slow_tbl = full_in_lds ? rows_full : gtbl;
        __syncthreads();   // previous tile's readers of LDS are done
        if (g != cur_g) {
            for (int i = lane; i < K * XD; i += 64) rows_fast[i] = (i % XD) < RD ? gtbl[(i / XD) * RW + (i % XD)] : 0u;
            if (full_in_lds)
                for (int i = lane; i < K * RW; i += 64) rows_full[i] = gtbl[i];
            cur_g = g;
        }
        __syncthreads();   // rows_fast visible

        // ---- register tile: RPL consecutive list elements per lane, built once per tile
        u32 xw[RPL][XD];
        int xoff[RPL];
        const u64 b0 = bx * (u64)(64 * RPL) + (u64)lane * RPL;
        const Mask m0 = (b0 < R.n_lane) ? colex_unrank_any<Mask>(b0, lk, lm, sbin, a.binom, a.bk) : Mask{};
        {
            // the first element's rows XORed in full, each next one by the rows that leave
            // or enter between colex neighbours (usually two)
            Mask m = m0;
            u32 acc[XD];
#pragma unroll
            for (int f = 0; f < XD; ++f) acc[f] = 0;
            for (Mask mm = m; nonzero(mm);) {
                const int b = take_low(mm);
#pragma unroll
                for (int f = 0; f < XD; ++f) acc[f] ^= rows_fast[(lbase + b) * XD + f];
            }
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const bool valid = b0 + k < R.n_lane;
#pragma unroll
                for (int f = 0; f < XD; ++f) xw[k][f] = valid ? acc[f] : 0u;
                if (valid) {
                    xoff[k] = pop_and(m, lflags);
                    if (b0 + k + 1 < R.n_lane && lk > 0) {
                        const Mask nm = colex_next(m);
                        for (Mask mm = mask_xor(nm, m); nonzero(mm);) {
                            const int b = take_low(mm);
#pragma unroll
                            for (int f = 0; f < XD; ++f) acc[f] ^= rows_fast[(lbase + b) * XD + f];
                        }
                        m = nm;
                    }
                } else {
                    xoff[k] = kBigOffset;
                }
            }
        }

        const u64 ch0 = bg * (u64)a.G;
        const u64 ch1 = min(ch0 + (u64)a.G, R.stream_chunks);
        for (u64 ch = ch0; ch < ch1; ++ch) {
            // ---- stream chunk -> LDS (each lane a contiguous run, colex successor steps)
            const u64 c0 = ch * (u64)a.S;
            const int len = (int)min((u64)a.S, R.n_stream - c0);
            if (ch != ch0) __syncthreads();   // previous chunk's readers are done
            const int per = (len + 63) >> 6;
            {
                const int i0 = lane * per;
                const int i1 = min(len, i0 + per);
                if (i0 < i1) {
                    const u64 e0 = c0 + (u64)i0;
                    int fi = flist ? (int)(e0 / ns) : 0;         // this run's fixed set
                    u64 si = flist ? e0 - (u64)fi * ns : e0;     // and its subset's rank
                    Mask m = colex_unrank_any<Mask>(si, sk, sm, sbin, a.binom, a.bk);
                    run_start[lane] = m;
                    run_fix[lane] = fi;
                    u32 fwe[XD];
                    int foffe = foff;
#pragma unroll
                    for (int f = 0; f < XD; ++f) fwe[f] = fw[f];
                    auto load_fix = [&](int k) {   // fixed set k of the rectangle's list
                        const u32* fk = J->fixed + (size_t)(R.fixed + k) * (RW + 1);
#pragma unroll
                        for (int f = 0; f < XD; ++f) fwe[f] = f < RD ? fk[f] : 0u;
                        foffe = (int)fk[RW];
                    };
                    if (flist) load_fix(fi);
                    u32 acc[XD];   // the XOR of the current subset's rows, kept by successor steps
#pragma unroll
                    for (int f = 0; f < XD; ++f) acc[f] = 0;
                    auto xor_rows = [&](Mask mm) {
                        while (nonzero(mm)) {
                            const int b = take_low(mm);
#pragma unroll
                            for (int f = 0; f < XD; ++f) acc[f] ^= rows_fast[(sbase + b) * XD + f];
                        }
                    };
                    xor_rows(m);
                    for (int i = i0; i < i1; ++i) {
                        SElem<XD> e;
#pragma unroll
                        for (int f = 0; f < XD; ++f) e.w[f] = fwe[f] ^ acc[f];
                        e.off = foffe + pop_and(m, sflags);
                        stream[i] = e;
                        if (i + 1 < i1) {
                            if (flist && si + 1 == ns) {   // next fixed set, first subset
                                ++fi;
                                si = 0;
                                const Mask nm = colex_first<Mask>(sk);
                                xor_rows(mask_xor(nm, m));
                                m = nm;
                                load_fix(fi);
                            } else {
                                const Mask nm = colex_next(m);
                                xor_rows(mask_xor(nm, m));
                                m = nm;
                                ++si;
                            }
                        }
                    }
                }
            }
            __syncthreads();   // stream chunk visible

            // ---- hot loop: every (lane element, stream element) pair. Two steps per trip
            // with two element registers: the LDS read of the next element is in flight
            // while the current one is tested, without register copies.
            auto step = [&](const u32 (&ew)[FD], const int eoff, const int s) {
                int acc = kBigOffset;
#pragma unroll
                for (int k = 0; k < RPL; ++k) {
                    int tk = xoff[k];
#pragma unroll
                    for (int f = 0; f < FD; ++f) tk = popc_acc(xw[k][f] ^ ew[f], tk);
                    acc = min(acc, tk);
                }
                const int thr = bound - eoff;
                if (__builtin_expect(__ballot(acc < thr) != 0ull, 0)) {
                    // ---- medium test: all XD register dwords
                    u32 ex[XD];
#pragma unroll
                    for (int f = 0; f < XD; ++f) ex[f] = stream[s].w[f];
                    u32 cand = 0;
#pragma unroll
                    for (int k = 0; k < RPL; ++k) {
                        int tk = xoff[k];
#pragma unroll
                        for (int f = 0; f < XD; ++f) tk += __popc(xw[k][f] ^ ex[f]);
                        if (tk < thr) cand |= 1u << k;
                    }
                    ++medium_entries;
                    ++job_hits;
                    if (__ballot(cand != 0) == 0ull) return;
                    // ---- slow path: exact weight and logical test of the light candidates
                    ++slow_entries;
                    const int owner = s / per;
                    Mask smask = run_start[owner];
                    const u32* fxe = fx;   // the candidate stream element's fixed set
                    if (flist) {
                        int fi = run_fix[owner];
                        u64 si = c0 + (u64)(owner * per) - (u64)fi * ns;
                        for (int j = owner * per; j < s; ++j) {
                            if (si + 1 == ns) { ++fi; si = 0; smask = colex_first<Mask>(sk); }
                            else { smask = colex_next(smask); ++si; }
                        }
                        fxe = J->fixed + (size_t)(R.fixed + fi) * (RW + 1);
                    } else {
                        for (int j = owner * per; j < s; ++j) smask = colex_next(smask);
                    }
                    if (lane < RW) {
                        u32 v = fxe ? fxe[lane] : 0u;
                        Mask mm = smask;
                        while (nonzero(mm)) v ^= slow_tbl[(sbase + take_low(mm)) * RW + lane];
                        sfull[lane] = v;
                    }
                    __syncthreads();
                    int mine = QSF_WEIGHT_NONE;
                    while (cand) {
                        const int k = __builtin_ctz(cand);
                        cand &= cand - 1;
                        Mask lmask = m0;
                        for (int j = 0; j < k; ++j) lmask = colex_next(lmask);
                        int w = pop_and(lmask, lflags) + eoff;
                        for (int f = 0; f < RD && w < bound; ++f) {
                            u32 v = sfull[f];
                            Mask mm = lmask;
                            while (nonzero(mm)) v ^= slow_tbl[(lbase + take_low(mm)) * RW + f];
                            w += __popc(v);
                        }
                        if (a.collect) {
                            // bound = keep + 1; the lanes of this candidate round append together
                            const unsigned long long slot = wave_append(a.out_rows, w >= 1 && w < bound);
                            if (slot != ~0ull) {
                                if (slot < a.out_cap) {
                                    const int stride = J->stride;
                                    const u64* G = J->gamma + (size_t)g * K * stride;
                                    for (int wd = 0; wd < stride; ++wd) {
                                        u64 v = 0;
                                        for (int q = 0; q < 4; ++q)
                                            for (u64 mm = Rr.fmask[q]; mm; mm &= mm - 1)
                                                v ^= G[(size_t)(64 * q + __builtin_ctzll(mm)) * stride + wd];
                                        for (Mask mm = smask; nonzero(mm);) v ^= G[(size_t)(sbase + take_low(mm)) * stride + wd];
                                        for (Mask mm = lmask; nonzero(mm);) v ^= G[(size_t)(lbase + take_low(mm)) * stride + wd];
                                        a.out[slot * stride + wd] = v;
                                    }
                                }
                            }
                            continue;   // next candidate of this lane
                        }
                        if (w < bound && w < mine) {
                            bool logical = (SD == 0);
                            for (int f = 0; f < SD && !logical; ++f) {
                                u32 v = sfull[RD + f];
                                Mask mm = lmask;
                                while (nonzero(mm)) v ^= slow_tbl[(lbase + take_low(mm)) * RW + RD + f];
                                logical = v != 0;
                            }
                            if (logical) mine = w;
                        }
                    }
                    const int wm = __builtin_amdgcn_readfirstlane(wave_min(mine));
                    if (!a.collect && wm < bound) {
                        bound = wm;
                        if (lane == 0) atomicMin(J->best, wm);
                        if (lane == 0 && J->shared != J->best) atomicMin(J->shared, wm);
                    }
                    __syncthreads();
                }
            };
            u32 aw[FD], bw[FD];
            int aoff = 0, boff = 0;
#pragma unroll
            for (int f = 0; f < FD; ++f) aw[f] = stream[0].w[f];
            aoff = stream[0].off;
            for (int s = 0; s < len; s += 2) {
                if (s + 1 < len) {
#pragma unroll
                    for (int f = 0; f < FD; ++f) bw[f] = stream[s + 1].w[f];
                    boff = stream[s + 1].off;
                }
                step(aw, aoff, s);
                if (s + 1 >= len) break;
                if (s + 2 < len) {
#pragma unroll
                    for (int f = 0; f < FD; ++f) aw[f] = stream[s + 2].w[f];
                    aoff = stream[s + 2].off;
                }
                step(bw, boff, s + 1);
            }
            if (bound <= floor) break;
        }
    }
    if (lane == 0 && slow_entries) atomicAdd(a.slow_hits, (unsigned long long)slow_entries);
    if (lane == 0 && medium_entries) atomicAdd(a.slow_hits + 1, (unsigned long long)medium_entries);
    if (lane == 0 && job_hits) atomicAdd(J->hits, (unsigned long long)job_hits);
}

// =====================================================================================
// bz_mfma: the hot test as int8 matrix products (min mode)
// =====================================================================================
//
// The weight of a codeword on its flagged rows and on 30 of its residual columns is a lower
// bound on its weight, and it is a dot product: for residual bits a (lane element) and b
// (stream element), hd(a, b) = |a| + sum_i (1 - 2 a_i) b_i. With the lane side as +1 / -1
// bytes plus the two dimensions (1, off_lane + |a|) and the stream side as 0 / 1 bytes plus
// (off_stream, 1), one v_mfma_i32_32x32x32_i8 yields w30 = off_lane + off_stream + hd for a
// 32 x 32 block of codewords. Residual bit j + 8 i (j < 8, i < 4) sits in byte i of dword j
// of a codeword's 32-byte vector, so a stream element's bytes are (w >> j) & 0x01010101;
// bits 30 and 31 (dword 6 and 7, byte 3) carry the two extra dimensions. Any byte order
// shared by both operands gives the same products, so the instruction's lane layout does
// not enter.
//
// A workgroup is four waves; a tile is 1024 lane elements (256 per wave, held as eight A
// operands: lane l holds row l % 32 and the 16 dimensions of half l / 32) times the stream
// range of nch chunks of kMfmaStream elements. Thread t owns the run of 2 nch consecutive
// stream elements starting at 2 nch t (one unrank per tile, then colex successors whose
// first residual dword changes by the rows that enter or leave) and contributes two of
// them to each chunk. A chunk is made once in LDS and tested by all four waves: per group
// of 32 stream elements a wave reads its B operand once, issues eight products and reduces
// them with v_min3. A group whose minimum is below the bound is tested exactly: the
// products are formed again and each lane tests the pairs of its values below the bound
// (XD residual dwords, then the full residual and syndrome).
template <int XD>
__global__ __launch_bounds__(256, 4) void bz_mfma(TiledArgs a) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef int v16i __attribute__((ext_vector_type(16)));
    __shared__ u32 rows_fast[kMaxMatrixK * XD];
    __shared__ v4i sbytes[2 * kMfmaStream];                  // half h of chunk position e at [h * kMfmaStream + e]
    __shared__ u64 smk[kMfmaStream];                          // the chunk's stream-element subsets
    __shared__ u64 lmk[kMfmaWaves * 256];                     // each wave's lane-element subsets
    __shared__ unsigned short lay[64 * 16];                   // product layout: (row, column) of lane l, value v
    __shared__ unsigned long long s_tile;
    __shared__ int s_choice;                                   // the candidate picked for the tile's level
    __shared__ u32 sbin[kSmallBinRows * kSmallBinCols];        // the unrank's binomials C(n < 64, k <= 8)
    extern __shared__ u32 rows_full[];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int half = lane >> 5, row = lane & 31;
    if (wave == 0) {
        // The (row, column) that value v of lane l holds, read off two products whose
        // entries are their own row and column index (row index x ones, ones x column
        // index, both on dimension 0), so the exact tests need no assumed layout.
        const v4i idx = v4i{half == 0 ? row : 0, 0, 0, 0};
        const v4i one = v4i{half == 0 ? 1 : 0, 0, 0, 0};
        const v16i zero = {};
        const v16i pr = __builtin_amdgcn_mfma_i32_32x32x32_i8(idx, one, zero, 0, 0, 0);
        const v16i pc = __builtin_amdgcn_mfma_i32_32x32x32_i8(one, idx, zero, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 16; ++v) lay[lane * 16 + v] = (unsigned short)((pr[v] & 31) | ((pc[v] & 31) << 8));
    }
    for (int i = tid; i < kSmallBinRows * kSmallBinCols; i += 64 * kMfmaWaves)
        sbin[i] = (u32)a.binom[(i / kSmallBinCols) * a.bk + i % kSmallBinCols];
    // (visible to every wave after the loop's first barrier)
    int cur_job = -1, cur_g = -1;
    int K = 0, RW = 0, RD = 0, SD = 0, floor = -1, bound = 0;
    const TJob* J = nullptr;
    unsigned slow_entries = 0, medium_entries = 0;
    unsigned job_hits = 0;   // this wave's medium entries on the current job (flushed to J->hits)

    for (;;) {
        __syncthreads();   // every wave is done with the previous tile's LDS
        if (tid == 0) s_tile = atomicAdd(a.next_tile, 1ull);
        __syncthreads();
        const u64 t = uni64(s_tile);
        if (t >= a.total_tiles) break;
        const Rect& Rr = a.rects[uni(find_rect(a, t))];   // the rectangle holding tile t
        const int job = uni(Rr.job);
        if (job != cur_job) {
            if (lane == 0 && job_hits) atomicAdd(J->hits, (unsigned long long)job_hits);
            job_hits = 0;
            J = a.jobs + job;
            K = uni(J->K); RW = uni(J->RW); RD = uni(J->RD); SD = uni(J->SD); floor = uni(J->floor);
            bound = uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            cur_job = job;
            cur_g = -1;
        }
        bound = min(bound, uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
        if (__syncthreads_or(bound <= floor)) {
            if (a.njobs == 1) break;
            continue;
        }
        {   // another level of this problem, in the same launch, settled it
            const int sb = uni(__hip_atomic_load(J->shared, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (J->choice) {   // one of the level's candidate variants: run only the one picked
                if (tid == 0) s_choice = pick_variant(J, J->shared != J->best ? min(bound, sb) : bound);
                __syncthreads();
                if (s_choice != J->variant) continue;
            }
            if (__syncthreads_or(J->shared != J->best && sb <= floor)) {
                if (tid == 0) atomicAdd(J->skipped, 1ull);
                continue;
            }
            if (J->shared != J->best) bound = min(bound, sb);   // the level loop's bound (as bz_tiled)
        }
        const u64 tile_base = uni64(Rr.tile_base), stream_groups = uni64(Rr.stream_groups);
        const u64 stream_chunks = uni64(Rr.stream_chunks), n_lane = uni64(Rr.n_lane), n_stream = uni64(Rr.n_stream);
        const int fixed = uni(Rr.fixed);
        const u64 local = t - tile_base;
        const u64 bx = local / stream_groups;
        const u64 bg = local - bx * stream_groups;
        const int g = uni(Rr.g);
        const int lk = uni(Rr.lk), lbase = uni(Rr.lbase), lm = uni(Rr.lm);
        const int sk = uni(Rr.sk), sbase = uni(Rr.sbase), sm = uni(Rr.sm);
        const u64 lflags = range_mask<u64>(J->flags + 4 * g, lbase, lm);
        const u64 sflags = range_mask<u64>(J->flags + 4 * g, sbase, sm);
        const u32* fx = fixed >= 0 ? J->fixed + (size_t)fixed * (RW + 1) : nullptr;
        const u32 fw0 = fx ? fx[0] : 0u;
        const int foff = fx ? (int)fx[RW] : 0;
        const u32* gtbl = J->tbl + (size_t)g * K * RW;
        const bool full_in_lds = K * RW <= a.lds_full;
        const u32* slow_tbl = full_in_lds ? rows_full : gtbl;
        if (g != cur_g) {   // same tile in every wave: the branch is workgroup-uniform
            for (int i = tid; i < K * XD; i += 256) rows_fast[i] = (i % XD) < RD ? gtbl[(i / XD) * RW + (i % XD)] : 0u;
            if (full_in_lds)
                for (int i = tid; i < K * RW; i += 256) rows_full[i] = gtbl[i];
            cur_g = g;
        }
        __syncthreads();

        // ---- this wave's 256 lane elements: row r of block b is element
        // bx * kMfmaLane + 256 wave + 8 r + b (lanes l and l + 32 share a row)
        v4i av[kMfmaBlocks];
        u64* wlmk = lmk + 256 * wave;
        {
            const u64 first = bx * (u64)kMfmaLane + (u64)(256 * wave + kMfmaBlocks * row);
            u64 mnext = (first < n_lane && lk) ? colex_unrank_any<u64>(first, lk, lm, sbin, a.binom, a.bk) : 0ull;
#pragma unroll
            for (int b = 0; b < kMfmaBlocks; ++b) {
                const u64 idx = first + (u64)b;
                const bool valid = idx < n_lane;
                const u64 m = valid ? mnext : 0ull;
                if (valid && lk && idx + 1 < n_lane) mnext = colex_next(mnext);
                if (half == 0) wlmk[kMfmaBlocks * row + b] = m;
                u32 x0 = 0;
                for (u64 mm = m; mm; mm &= mm - 1) x0 ^= rows_fast[(lbase + __builtin_ctzll(mm)) * XD];
                const int xoff = __popcll(m & lflags);
                u32 q[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {   // bytes 0 / 1 -> +1 / -1 (0xFE * x = (x << 8) - (x << 1))
                    const u32 x = (x0 >> (4 * half + i)) & 0x01010101u;
                    q[i] = 0x01010101u ^ ((x << 8) - (x << 1));
                }
                if (half == 1) {   // dimension 27: 1, dimension 31: off_lane + |a|
                    const int w = min(xoff + __popc(x0 & 0x3fffffffu), 127);
                    q[2] = (q[2] & 0x00ffffffu) | (1u << 24);
                    q[3] = (q[3] & 0x00ffffffu) | ((u32)(valid ? w : 127) << 24);
                }
                if (!valid) {
                    q[0] = q[1] = 0u;
                    if (half == 0) q[2] = q[3] = 0u; else { q[2] = 1u << 24; q[3] = 127u << 24; }
                }
                av[b] = v4i{(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
            }
        }

        // ---- this thread's stream run: elements s0 + 2 nch tid + [0, 2 nch)
        const u64 ch0 = bg * (u64)a.G;
        const int nch = (int)(min(ch0 + (u64)a.G, stream_chunks) - ch0);
        const u64 s0 = ch0 * (u64)kMfmaStream;
        u64 sidx = s0 + (u64)(2 * nch) * (u64)tid;
        // threads whose run holds a real element; every chunk position from 2 t_real on is
        // padding (products >= 127), so the tile's groups past it are not tested
        const u64 t_real = min((u64)(64 * kMfmaWaves), (n_stream - s0 + (u64)(2 * nch) - 1) / (u64)(2 * nch));
        const int gmax = uni((int)min((u64)(kMfmaStream / 32), (2 * t_real + 31) / 32));
        u64 smask = (sidx < n_stream && sk) ? colex_unrank_any<u64>(sidx, sk, sm, sbin, a.binom, a.bk) : 0ull;
        u32 sx0 = fw0;
        for (u64 mm = smask; mm; mm &= mm - 1) sx0 ^= rows_fast[(sbase + __builtin_ctzll(mm)) * XD];

        for (int k = 0; k < nch; ++k) {
            if (k != 0) {
                bound = min(bound, uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                if (J->shared != J->best)
                    bound = min(bound, uni(__hip_atomic_load(J->shared, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                if (__syncthreads_or(bound <= floor)) break;   // also: the previous chunk is consumed
            }
#pragma unroll
            for (int e = 0; e < 2; ++e) {   // chunk positions 2 tid, 2 tid + 1
                const int pos = 2 * tid + e;
                if (sidx < n_stream) {
                    u32 q[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) q[j] = (sx0 >> j) & 0x01010101u;
                    const int off = foff + __popcll(smask & sflags);
                    q[6] = (q[6] & 0x00ffffffu) | ((u32)min(off, 127) << 24);   // dimension 27: off_stream
                    q[7] = (q[7] & 0x00ffffffu) | (1u << 24);                  // dimension 31: 1
                    sbytes[pos] = v4i{(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
                    sbytes[kMfmaStream + pos] = v4i{(int)q[4], (int)q[5], (int)q[6], (int)q[7]};
                    smk[pos] = smask;
                    if (sk && sidx + 1 < n_stream) {   // colex successor; XOR the rows that change
                        const u64 nm = colex_next(smask);
                        for (u64 mm = nm ^ smask; mm; mm &= mm - 1) sx0 ^= rows_fast[(sbase + __builtin_ctzll(mm)) * XD];
                        smask = nm;
                    }
                    ++sidx;
                } else {   // past the list: products of 127, at or above any threshold
                    sbytes[pos] = v4i{0, 0, 0, 0};
                    sbytes[kMfmaStream + pos] = v4i{0, 0, (int)(127u << 24), 0};
                }
            }
            __syncthreads();

            int thrd = min(bound, 127);
            // B operands one group ahead; the lane blocks in two passes of four independent
            // products, each reduced by a chain of v_min3
            const v4i* sb = sbytes + half * kMfmaStream + row;
            v4i bnext = sb[0];
            for (int gi = 0; gi < gmax; ++gi) {
                const v4i bv = bnext;
                if (gi + 1 < gmax) bnext = sb[(gi + 1) * 32];
                const v16i zero = {};
                auto red = [](const v16i& c) {
                    int m = min(min(c[0], c[1]), c[2]);
#pragma unroll
                    for (int v = 3; v < 15; v += 2) m = min(min(m, c[v]), c[v + 1]);
                    return min(m, c[15]);
                };
                int mh[2];   // minimum over blocks 0..3 and 4..7, kept for the exact path
#pragma unroll
                for (int b0 = 0; b0 < kMfmaBlocks; b0 += 4) {
                    const v16i p0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[b0 + 0], bv, zero, 0, 0, 0);
                    const v16i p1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[b0 + 1], bv, zero, 0, 0, 0);
                    const v16i p2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[b0 + 2], bv, zero, 0, 0, 0);
                    const v16i p3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[b0 + 3], bv, zero, 0, 0, 0);
                    // one v_min3 chain over the half's 64 values (32 instructions)
                    int m = min(min(p0[0], p0[1]), p0[2]);
#pragma unroll
                    for (int v = 3; v < 15; v += 2) m = min(min(m, p0[v]), p0[v + 1]);
                    m = min(min(m, p0[15]), p1[0]);
#pragma unroll
                    for (int v = 1; v < 15; v += 2) m = min(min(m, p1[v]), p1[v + 1]);
                    m = min(min(m, p1[15]), p2[0]);
#pragma unroll
                    for (int v = 1; v < 15; v += 2) m = min(min(m, p2[v]), p2[v + 1]);
                    m = min(min(m, p2[15]), p3[0]);
#pragma unroll
                    for (int v = 1; v < 15; v += 2) m = min(min(m, p3[v]), p3[v + 1]);
                    mh[b0 / 4] = min(m, p3[15]);
                }
                if (__builtin_expect(__ballot(min(mh[0], mh[1]) < thrd) == 0ull, 1)) continue;
                // ---- exact tests of this group: the products of the half (four blocks) whose
                // minimum is below the threshold are formed again, each lane marks its values
                // below it (bit 16 (b % 4) + v of cand[b / 4]), then tests those (lane
                // element, stream element) pairs; both elements are then real (padding gives
                // products >= 127)
                ++medium_entries;
                ++job_hits;
                int mine = QSF_WEIGHT_NONE;
                u64 cand[2] = {0ull, 0ull};
#pragma unroll
                for (int b = 0; b < kMfmaBlocks; ++b) {
                    if (__ballot(mh[b / 4] < thrd) == 0ull) continue;
                    const v16i p = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[b], bv, zero, 0, 0, 0);
                    if (__ballot(red(p) < thrd) == 0ull) continue;
                    u64 bits = 0;
#pragma unroll
                    for (int v = 0; v < 16; ++v) bits |= (p[v] < thrd ? 1ull : 0ull) << (16 * (b & 3) + v);
                    cand[b >> 2] |= bits;
                }
                for (int part = 0; part < 2; ++part) {
                    for (u64 cm = part ? cand[1] : cand[0]; cm; cm &= cm - 1) {
                        const int q = __builtin_ctzll(cm);
                        const int b = 4 * part + (q >> 4);
                        const u32 ij = lay[lane * 16 + (q & 15)];
                        const u64 lmask = wlmk[kMfmaBlocks * (ij & 31) + b];
                        const int pos = gi * 32 + (int)(ij >> 8);
                        const u64 sm_ = smk[pos];
                        const int base = __popcll(lmask & lflags) + foff + __popcll(sm_ & sflags);
                        int tk = base;
                        {   // XD residual dwords of the pair's codeword
                            u32 xw[XD];
#pragma unroll
                            for (int f = 0; f < XD; ++f) xw[f] = (fx && f < RD) ? fx[f] : 0u;
                            for (u64 mm = lmask; mm; mm &= mm - 1) {
                                const int r = lbase + __builtin_ctzll(mm);
#pragma unroll
                                for (int f = 0; f < XD; ++f) xw[f] ^= rows_fast[r * XD + f];
                            }
                            for (u64 mm = sm_; mm; mm &= mm - 1) {
                                const int r = sbase + __builtin_ctzll(mm);
#pragma unroll
                                for (int f = 0; f < XD; ++f) xw[f] ^= rows_fast[r * XD + f];
                            }
#pragma unroll
                            for (int f = 0; f < XD; ++f) tk += __popc(xw[f]);
                        }
                        if (tk >= bound || tk >= mine) continue;
                        // exact weight and logical test of the pair
                        ++slow_entries;
                        int w = base;
                        for (int f = 0; f < RD && w < bound && w < mine; ++f) {
                            u32 x = fx ? fx[f] : 0u;
                            for (u64 mm = lmask; mm; mm &= mm - 1) x ^= slow_tbl[(lbase + __builtin_ctzll(mm)) * RW + f];
                            for (u64 mm = sm_; mm; mm &= mm - 1) x ^= slow_tbl[(sbase + __builtin_ctzll(mm)) * RW + f];
                            w += __popc(x);
                        }
                        if (w >= bound || w >= mine) continue;
                        bool logical = (SD == 0);
                        for (int f = RD; f < RW && !logical; ++f) {
                            u32 x = fx ? fx[f] : 0u;
                            for (u64 mm = lmask; mm; mm &= mm - 1) x ^= slow_tbl[(lbase + __builtin_ctzll(mm)) * RW + f];
                            for (u64 mm = sm_; mm; mm &= mm - 1) x ^= slow_tbl[(sbase + __builtin_ctzll(mm)) * RW + f];
                            logical = x != 0;
                        }
                        if (logical) mine = w;
                    }
                }
                const int wm = uni(wave_min(mine));
                if (wm < bound) {
                    bound = wm;
                    if (lane == 0) atomicMin(J->best, wm);
                    if (lane == 0 && J->shared != J->best) atomicMin(J->shared, wm);
                    thrd = min(bound, 127);
                }
                if (bound <= floor) break;
            }
        }
    }
    if (lane == 0 && slow_entries) atomicAdd(a.slow_hits, (unsigned long long)slow_entries);
    if (lane == 0 && medium_entries) atomicAdd(a.slow_hits + 1, (unsigned long long)medium_entries);
    if (lane == 0 && job_hits) atomicAdd(J->hits, (unsigned long long)job_hits);
}

// =====================================================================================
// bz_f4: the hot test as fp4 (e2m1) matrix products (min mode)
// =====================================================================================
//
// As bz_mfma, on the fp4 matrix instructions with unit block scales:
//   MM = 2  v_mfma_scale_f32_32x32x64_f8f6f4   32 x 32 codewords per product, 64 dimensions
//   MM = 3  v_mfma_scale_f32_16x16x128_f8f6f4  16 x 16 codewords per product, 128 dimensions
//   MM = 4  two 32x32x64 products chained through the accumulator, 128 dimensions
//   MM = 5  32x32x64 products over two lane elements per row (packed, 30 test bits)
// gfx950 runs fp4 products at twice the int8 rate, so a product tests 60 (MM = 2) or 124
// (MM = 3) residual bits for about the time the int8 test spends on 30. MM = 4 tests 124
// bits on 32 x 32 outputs: twice the products of MM = 2 for the same reduction, with four
// lane blocks per wave (the second product's operands take the other four's registers). The lane side holds
// +1 / -1 (codes 0x2 / 0xA) per test bit, the stream side 1 / 0 (0x2 / 0x0). Dimensions
// 28..31 of the last 32-dimension block carry the lane constant L = off_lane + |a| as four
// fp4 digits (d0, d1, d2, d3) against the fixed stream values (6, 6, 6, 1), so the f32
// products are exact small integers:
//     out = L + sum_i (1 - 2 a_i) b_i = off_lane + hd(a, b)   over the test bits.
// Every value a lane holds belongs to the stream element of its column, lane mod NR: that
// is the instructions' output layout, read off index products at start-up and checked (a
// launch that finds another layout reports it and the host raises). The stream constant
// off_stream then enters as a per-lane threshold, min(bound, 127) - off_stream.
// Block kb of the 32-dimension blocks (lanes kb NR .. kb NR + NR - 1) of product p holds
// residual word p NKB + kb of both elements: nibble i of dword j is bit j + 4 i, i.e.
// (w >> j) & 0x11111111 moved into place (two VALU per dword); the last word keeps its bits
// 0..27.
// Padding: a lane element past its list holds only the constant 114 (digits 6, 6, 6, 6), a
// stream element past its list only zeros with off_stream = 127; the host uses these
// kernels only while the bound is at most 100, so neither can pass.
// Packed (MM = 5): lanes 0..31 hold lane element (row, 0) and lanes 32..63 element (row, 1),
// the second with block scale 2^8; both halves of B hold the stream element's 30 test bits.
// Lane side -a_i per bit and -off_lane (clamped to the tile's threshold T, two digits
// against the fixed (6, 1)), stream side 1 - 2 b_i; the accumulator starts at
// 2^23 + 257 (127 + T - off_stream' - |b|), a per-column constant. Each f32 value is then
// 2^23 + F0 + 256 F1 with F = 127 + T - (off_lane' + off_stream' + hd(a, b)) in [37, 157],
// an exact integer whose bits are the value's low mantissa bits: bit 7 (15) is set exactly
// when pair (row, 0) ((row, 1)) passes, so the reduction is an OR of the raw bits
// (v_or3_b32). Padding elements hold no test bits and the constants that keep F <= 127.
// The host runs it only at bounds <= kPackMaxT.

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// One product (NP = 1) or NP chained products (the second accumulates onto the first);
// NB lane blocks per wave.
template <int MM> struct F4Prod;
template <> struct F4Prod<2> {
    typedef f32x16 acc;
    static constexpr int NR = 32, NP = 1, NB = kMfmaBlocks;
    __device__ static acc mm(i32x4 a, i32x4 b, acc c = acc{}) {
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(i32x8{a[0], a[1], a[2], a[3], 0, 0, 0, 0},
                                                               i32x8{b[0], b[1], b[2], b[3], 0, 0, 0, 0}, c,
                                                               4, 4, 0, 127, 0, 127);
    }
};
template <> struct F4Prod<3> {
    typedef f32x4 acc;
    static constexpr int NR = 16, NP = 1, NB = kMfmaBlocks;
    __device__ static acc mm(i32x4 a, i32x4 b, acc c = acc{}) {
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(i32x8{a[0], a[1], a[2], a[3], 0, 0, 0, 0},
                                                                i32x8{b[0], b[1], b[2], b[3], 0, 0, 0, 0}, c,
                                                                4, 4, 0, 127, 0, 127);
    }
};
template <> struct F4Prod<4> : F4Prod<2> {
    static constexpr int NP = 2, NB = f4_blocks(4);
};
// Two lane elements per row: lanes 0..31 carry one, lanes 32..63 the other with block scale
// 2^8 (E8M0 135), each lane's scale applying to the 32 dimensions it holds.
template <> struct F4Prod<5> : F4Prod<2> {
    __device__ static acc mms(i32x4 a, i32x4 b, acc c, int scale_a) {
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(i32x8{a[0], a[1], a[2], a[3], 0, 0, 0, 0},
                                                               i32x8{b[0], b[1], b[2], b[3], 0, 0, 0, 0}, c,
                                                               4, 4, 0, scale_a, 0, 127);
    }
};

// fp4 e2m1 code of v in {-1, 0, 1, 2, 3, 4, 6}
__device__ __forceinline__ u32 f4_code(int v) {
    return v < 0 ? 0xAu : v == 0 ? 0u : v == 1 ? 2u : v == 2 ? 4u : v == 3 ? 5u : v == 4 ? 6u : 7u;
}

// The lane constant L (clamped to [0, 100]; a smaller value only weakens the test) as
// 6 (d0 + d1 + d2) + d3, digits d0..d2 in {0, 1, 2, 3, 4, 6} and d3 in {-1, 0, .., 4}:
// nibble j of the result is the code of d_j.
__device__ __forceinline__ u32 f4_lane_const(int L) {
    L = max(0, min(L, 100));
    int t = L / 6, u = L - 6 * t;
    if (u == 5) { t += 1; u = -1; }
    const int d0 = t >= 6 ? 6 : (t == 5 ? 4 : t);
    const int r = t - d0;
    const int d1 = r >= 6 ? 6 : (r == 5 ? 4 : r);
    return f4_code(d0) | (f4_code(d1) << 4) | (f4_code(r - d1) << 8) | (f4_code(u) << 12);
}

// The packed test's lane constant, negated: -L (L clamped to [0, kPackMaxT], 29 taken as 28;
// a smaller value only weakens the test) as -(6 d0 + d1), d0 in {0, .., 4} and d1 in
// {-1, 0, .., 4, 6}: nibble 0 is the code of -d0, nibble 1 that of -d1.
__device__ __forceinline__ u32 f4_neg_code(int v) { return v == 0 ? 0u : v > 0 ? 8u | f4_code(v) : f4_code(-v); }
__device__ __forceinline__ u32 f4_pack_const(int L) {
    L = max(0, min(L, kPackMaxT));
    if (L == 29) L = 28;
    int t = L / 6, u = L - 6 * t;
    if (u == 5) { t += 1; u = -1; }
    if (t == 5) { t = 4; u += 6; }
    return f4_neg_code(t) | (f4_neg_code(u) << 4);
}

template <int XD, int MM>
__global__ __launch_bounds__(256, 4) void bz_f4(TiledArgs a) {
    typedef F4Prod<MM> P;
    typedef typename P::acc acc;
    constexpr int NR = P::NR;                     // rows (lane elements) = columns (stream elements) per product
    constexpr int NKB = 64 / NR;                  // 32-dimension blocks across the lanes
    constexpr int NP = P::NP;                     // chained products per lane block
    constexpr bool PK = MM == 5;                  // two lane elements per row, one per half
    constexpr int NW = PK ? 1 : NKB * NP;         // residual words tested
    constexpr int NB = P::NB;                     // lane blocks per wave
    constexpr int NV = NR * NR / 64;              // product values per lane
    constexpr int LW = NB * NR * f4_per_row(MM);  // lane elements per wave
    constexpr int NCB = (NB * NV * f4_per_row(MM) + 63) / 64;   // candidate words of a group
    constexpr int SL = f4_stream(MM);             // stream chunk length
    constexpr int EPT = SL / (64 * kMfmaWaves);   // stream elements each thread makes per chunk
    static_assert(XD >= NW, "the held residual dwords must cover the test bits");
    static_assert(NB * NV * f4_per_row(MM) <= 256, "candidate bits of a group");
    __shared__ u32 rows_fast[kMaxMatrixK * XD];
    __shared__ i32x4 sbytes[NW * SL];             // word w of chunk position e at [w * SL + e]
    __shared__ float soff[SL];                    // off_stream of each chunk position (127: padding)
    __shared__ u64 smk[SL];                       // the chunk's stream-element subsets
    __shared__ u64 lmk[kMfmaWaves * LW];          // each wave's lane-element subsets
    __shared__ unsigned short lay[64 * NV];       // product layout: (row, column) of lane l, value v
    __shared__ int s_bad;
    __shared__ unsigned long long s_tile;
    __shared__ int s_choice;                                   // the candidate picked for the tile's level
    __shared__ u32 sbin[kSmallBinRows * kSmallBinCols];        // the unrank's binomials C(n < 64, k <= 8)
    extern __shared__ u32 rows_full[];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int kb = lane / NR, row = lane % NR;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    if (wave == 0) {
        // Index products on dimensions 0..2 (block 0): row-index digits (r & 3, (r >> 2) & 3,
        // 4 (r >> 4)) against weights (1, 4, 4) give r; weights against column digits give c.
        const u32 idx = kb == 0 ? (f4_code(row & 3) | (f4_code((row >> 2) & 3) << 4) | (f4_code(4 * (row >> 4)) << 8)) : 0u;
        const u32 wts = kb == 0 ? (f4_code(1) | (f4_code(4) << 4) | (f4_code(4) << 8)) : 0u;
        const i32x4 I = {(int)idx, 0, 0, 0}, W = {(int)wts, 0, 0, 0};
        const acc pr = P::mm(I, W), pc = P::mm(W, I);
        bool bad = false;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int r = (int)pr[v], c = (int)pc[v];
            bad = bad || c != row || r < 0 || r >= NR;
            lay[lane * NV + v] = (unsigned short)((r & 31) | ((c & 31) << 8));
        }
        if constexpr (PK) {   // each lane's block scale applies to the 32 dimensions it holds
            const i32x4 one = {0x22222222, 0x22222222, 0x22222222, 0x22222222};
            const acc ps = P::mms(one, one, acc{}, kb == 0 ? 127 : 135);
#pragma unroll
            for (int v = 0; v < NV; ++v) bad = bad || ps[v] != 8224.0f;   // 32 + 256 * 32
        }
        if (__ballot(bad) != 0ull && lane == 0) s_bad = 1;
    }
    __syncthreads();
    if (s_bad) {   // another output layout or scale rule: report it, enumerate nothing
        if (tid == 0) atomicAdd(a.out_rows, 1ull);
        return;
    }

    for (int i = tid; i < kSmallBinRows * kSmallBinCols; i += 64 * kMfmaWaves)
        sbin[i] = (u32)a.binom[(i / kSmallBinCols) * a.bk + i % kSmallBinCols];
    // (visible to every wave after the loop's first barrier)
    int cur_job = -1, cur_g = -1;
    int K = 0, RW = 0, RD = 0, SD = 0, floor = -1, bound = 0;
    const TJob* J = nullptr;
    unsigned slow_entries = 0, medium_entries = 0;
    unsigned job_hits = 0;   // this wave's medium entries on the current job (flushed to J->hits)

    for (;;) {
        __syncthreads();   // every wave is done with the previous tile's LDS
        if (tid == 0) s_tile = atomicAdd(a.next_tile, 1ull);
        __syncthreads();
        const u64 t = uni64(s_tile);
        if (t >= a.total_tiles) break;
        const Rect& Rr = a.rects[uni(find_rect(a, t))];   // the rectangle holding tile t
        const int job = uni(Rr.job);
        if (job != cur_job) {
            if (lane == 0 && job_hits) atomicAdd(J->hits, (unsigned long long)job_hits);
            job_hits = 0;
            J = a.jobs + job;
            K = uni(J->K); RW = uni(J->RW); RD = uni(J->RD); SD = uni(J->SD); floor = uni(J->floor);
            bound = uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            cur_job = job;
            cur_g = -1;
        }
        bound = min(bound, uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
        if (__syncthreads_or(bound <= floor)) {
            if (a.njobs == 1) break;
            continue;
        }
        {   // another level of this problem, in the same launch, settled it
            const int sb = uni(__hip_atomic_load(J->shared, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (J->choice) {   // one of the level's candidate variants: run only the one picked
                if (tid == 0) s_choice = pick_variant(J, J->shared != J->best ? min(bound, sb) : bound);
                __syncthreads();
                if (s_choice != J->variant) continue;
            }
            if (__syncthreads_or(J->shared != J->best && sb <= floor)) {
                if (tid == 0) atomicAdd(J->skipped, 1ull);
                continue;
            }
            if (J->shared != J->best) bound = min(bound, sb);   // the level loop's bound (as bz_tiled)
        }
        // the packed test's tile threshold: it clamps the lane and stream constants to it,
        // which keeps every 8-bit field in range; the host runs it only at bounds <= kPackMaxT
        if (PK && bound > kPackMaxT) {
            if (tid == 0) atomicAdd(a.out_rows, 1ull);
            break;
        }
        const int ptile = PK ? bound : 0;
        const u64 tile_base = uni64(Rr.tile_base), stream_groups = uni64(Rr.stream_groups);
        const u64 stream_chunks = uni64(Rr.stream_chunks), n_lane = uni64(Rr.n_lane), n_stream = uni64(Rr.n_stream);
        const int fixed = uni(Rr.fixed);
        const u64 local = t - tile_base;
        const u64 bx = local / stream_groups;
        const u64 bg = local - bx * stream_groups;
        const int g = uni(Rr.g);
        const int lk = uni(Rr.lk), lbase = uni(Rr.lbase), lm = uni(Rr.lm);
        const int sk = uni(Rr.sk), sbase = uni(Rr.sbase), sm = uni(Rr.sm);
        const u64 lflags = range_mask<u64>(J->flags + 4 * g, lbase, lm);
        const u64 sflags = range_mask<u64>(J->flags + 4 * g, sbase, sm);
        const u32* fx = fixed >= 0 ? J->fixed + (size_t)fixed * (RW + 1) : nullptr;
        const int foff = fx ? (int)fx[RW] : 0;
        const u32* gtbl = J->tbl + (size_t)g * K * RW;
        const bool full_in_lds = K * RW <= a.lds_full;
        const u32* slow_tbl = full_in_lds ? rows_full : gtbl;
        if (g != cur_g) {   // same tile in every wave: the branch is workgroup-uniform
            for (int i = tid; i < K * XD; i += 256) rows_fast[i] = (i % XD) < RD ? gtbl[(i / XD) * RW + (i % XD)] : 0u;
            if (full_in_lds)
                for (int i = tid; i < K * RW; i += 256) rows_full[i] = gtbl[i];
            cur_g = g;
        }
        __syncthreads();

        // ---- this wave's LW lane elements: row r of block b is element
        // bx * (4 LW) + LW wave + NB r + b (the NKB lanes of a row share it)
        i32x4 av[NB][NP];
        u64* wlmk = lmk + LW * wave;
        {
            const int lrow = PK ? 2 * row + kb : row;   // this lane's element row in the wave
            const u64 first = bx * (u64)(kMfmaWaves * LW) + (u64)(LW * wave + NB * lrow);
            u64 mnext = (first < n_lane && lk) ? colex_unrank_any<u64>(first, lk, lm, sbin, a.binom, a.bk) : 0ull;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const u64 idx = first + (u64)b;
                const bool valid = idx < n_lane;
                const u64 m = valid ? mnext : 0ull;
                if (valid && lk && idx + 1 < n_lane) mnext = colex_next(mnext);
                if (PK || kb == 0) wlmk[NB * lrow + b] = m;
                u32 x[NW];
#pragma unroll
                for (int w = 0; w < NW; ++w) x[w] = 0u;
                for (u64 mm = m; mm; mm &= mm - 1) {
                    const int r = lbase + __builtin_ctzll(mm);
#pragma unroll
                    for (int w = 0; w < NW; ++w) x[w] ^= rows_fast[r * XD + w];
                }
                x[NW - 1] &= PK ? 0x3fffffffu : 0x0fffffffu;
                int L = __popcll(m & lflags);
#pragma unroll
                for (int w = 0; w < NW; ++w) L += __popc(x[w]);
                if constexpr (PK) {
                    // -a_i per test bit (codes 0xA / 0x0), dimensions 30 and 31 the negated
                    // off_lane (clamped to the tile threshold) against the stream's fixed (6, 1)
                    u32 q[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) q[j] = valid ? ((x[0] >> j) & 0x11111111u) * 0xAu : 0u;
                    const u32 c = f4_pack_const(valid ? min(__popcll(m & lflags), ptile) : kPackMaxT);
                    q[2] = (q[2] & 0x0fffffffu) | ((c & 15u) << 28);
                    q[3] = (q[3] & 0x0fffffffu) | ((c >> 4) << 28);
                    av[b][0] = i32x4{(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
                    continue;
                }
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    u32 xk = x[p * NKB];
#pragma unroll
                    for (int w = 1; w < NKB; ++w)
                        if (kb == w) xk = x[p * NKB + w];
                    u32 q[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) q[j] = 0x22222222u | (((xk >> j) & 0x11111111u) << 3);
                    if (p == NP - 1 && kb == NKB - 1) {
                        const u32 c = valid ? f4_lane_const(L) : 0x7777u;
#pragma unroll
                        for (int j = 0; j < 4; ++j) q[j] = (valid ? (q[j] & 0x0fffffffu) : 0u) | (((c >> (4 * j)) & 15u) << 28);
                    } else if (!valid) {
                        q[0] = q[1] = q[2] = q[3] = 0u;
                    }
                    av[b][p] = i32x4{(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
                }
            }
        }

        // ---- this thread's stream run: elements s0 + EPT nch tid + [0, EPT nch)
        const u64 ch0 = bg * (u64)a.G;
        const int nch = (int)(min(ch0 + (u64)a.G, stream_chunks) - ch0);
        const u64 s0 = ch0 * (u64)SL;
        u64 sidx = s0 + (u64)(EPT * nch) * (u64)tid;
        // threads whose run holds a real element; every chunk position from EPT t_real on is
        // padding, so the tile's groups past it are not tested
        const u64 t_real = min((u64)(64 * kMfmaWaves), (n_stream - s0 + (u64)(EPT * nch) - 1) / (u64)(EPT * nch));
        const int gmax = uni((int)min((u64)(SL / NR), (EPT * t_real + NR - 1) / NR));
        u64 smask = (sidx < n_stream && sk) ? colex_unrank_any<u64>(sidx, sk, sm, sbin, a.binom, a.bk) : 0ull;
        u32 sx[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) sx[w] = (fx && w < RD) ? fx[w] : 0u;
        for (u64 mm = smask; mm; mm &= mm - 1) {
            const int r = sbase + __builtin_ctzll(mm);
#pragma unroll
            for (int w = 0; w < NW; ++w) sx[w] ^= rows_fast[r * XD + w];
        }

        for (int k = 0; k < nch; ++k) {
            if (k != 0) {
                bound = min(bound, uni(__hip_atomic_load(J->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                if (J->shared != J->best)
                    bound = min(bound, uni(__hip_atomic_load(J->shared, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                if (__syncthreads_or(bound <= floor)) break;   // also: the previous chunk is consumed
            }
#pragma unroll
            for (int e = 0; e < EPT; ++e) {   // chunk positions EPT tid + e
                const int pos = EPT * tid + e;
                if (sidx < n_stream) {
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        const u32 sw = w == NW - 1 ? (sx[w] & (PK ? 0x3fffffffu : 0x0fffffffu)) : sx[w];
                        u32 q[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j)   // packed: 1 - 2 b_i (0x2 / 0xA), else b_i (0x2 / 0x0)
                            q[j] = PK ? 0x22222222u ^ (((sw >> j) & 0x11111111u) << 3) : ((sw >> j) & 0x11111111u) << 1;
                        if (w == NW - 1 && PK) {   // the fixed values (6, 1) of the lane-constant dimensions
                            q[2] = (q[2] & 0x0fffffffu) | (7u << 28); q[3] = (q[3] & 0x0fffffffu) | (2u << 28);
                        } else if (w == NW - 1) {   // the fixed values (6, 6, 6, 1) of the lane-constant dimensions
                            q[0] |= 7u << 28; q[1] |= 7u << 28; q[2] |= 7u << 28; q[3] |= 2u << 28;
                        }
                        sbytes[w * SL + pos] = i32x4{(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
                    }
                    // packed: the column constant min(off_stream, tile threshold) + |b|
                    soff[pos] = PK ? (float)(min(foff + __popcll(smask & sflags), ptile) + __popc(sx[0] & 0x3fffffffu))
                                   : (float)min(foff + __popcll(smask & sflags), 127);
                    smk[pos] = smask;
                    if (sk && sidx + 1 < n_stream) {   // colex successor; XOR the rows that change
                        const u64 nm = colex_next(smask);
                        for (u64 mm = nm ^ smask; mm; mm &= mm - 1) {
                            const int r = sbase + __builtin_ctzll(mm);
#pragma unroll
                            for (int w = 0; w < NW; ++w) sx[w] ^= rows_fast[r * XD + w];
                        }
                        smask = nm;
                    }
                    ++sidx;
                } else {   // past the list: zeros, and a threshold no product reaches
#pragma unroll
                    for (int w = 0; w < NW; ++w) sbytes[w * SL + pos] = i32x4{0, 0, 0, 0};
                    soff[pos] = PK ? (float)ptile : 127.0f;
                    smk[pos] = 0ull;
                }
            }
            __syncthreads();

            float thrf = (float)min(bound, 127);
            int tpk = min(bound, ptile);   // the packed test's threshold (as thrf)
            const int sa = kb == 0 ? 127 : 135;
            // B operands and stream constant one group ahead (product p: word p NKB + kb; the
            // chained test loads its second product's operand in the group, for registers)
            const i32x4* sb = sbytes + (PK ? 0 : kb) * SL + row;
            i32x4 bnext = sb[0];
            float onext = soff[row];
            // lane block b's product with the group's B operands
            auto prod = [&](int b, const i32x4 (&bv)[NP]) {
                acc c = P::mm(av[b][0], bv[0]);
#pragma unroll
                for (int p = 1; p < NP; ++p) c = P::mm(av[b][p], bv[p], c);
                return c;
            };
            for (int gi = 0; gi < gmax; ++gi) {
                i32x4 bv[NP];
                bv[0] = bnext;
#pragma unroll
                for (int p = 1; p < NP; ++p) bv[p] = sb[p * NKB * SL + gi * NR];
                const float thr = thrf - onext;   // this lane's column: L + sum < bound - off_stream
                // packed: every value is 2^23 + F1 + 256 F2 with the fields
                // F = 127 + tpk - off_stream' - |b| - off_lane' - sum_i a_i (1 - 2 b_i)
                //   = 127 + tpk - (off_stream' + off_lane' + hd(a, b))   in [37, 157]
                // (off' clamped to the tile threshold), so bit 7 of a field is set exactly
                // when the element pair's weight on the test bits is below tpk
                const int cx = (127 + tpk - (int)onext) * 257;
                const float cf = __int_as_float(0x4B000000 + cx);
                if (gi + 1 < gmax) {
                    bnext = sb[(gi + 1) * NR];
                    onext = soff[(gi + 1) * NR + row];
                }
                // The minimum over the products' bit patterns: non-negative floats order like
                // their bits as integers (v_min3_i32, no NaN canonicalisation); a negative value
                // (only from a clamped lane constant) becomes a negative integer and sends the
                // group to the exact tests, which compare in f32.
                int mn = 0x7fffffff;
                bool pass;
                int pk_pair = 0;
                if constexpr (PK) {
                    acc cv;
#pragma unroll
                    for (int v = 0; v < NV; ++v) cv[v] = cf;
                    auto red = [](const acc& c) {
                        int m = __float_as_int(c[0]) | __float_as_int(c[1]) | __float_as_int(c[2]);
#pragma unroll
                        for (int v = 3; v < 15; v += 2) m = m | __float_as_int(c[v]) | __float_as_int(c[v + 1]);
                        return m | __float_as_int(c[15]);
                    };
                    int orr = 0;
#pragma unroll
                    for (int b0 = 0; b0 < NB; b0 += 2) {
                        const acc p0 = P::mms(av[b0][0], bv[0], cv, sa);
                        const acc p1 = P::mms(av[b0 + 1][0], bv[0], cv, sa);
                        const int o = red(p0) | red(p1);
                        pk_pair |= ((o & 0x8080) != 0 ? 1 : 0) << (b0 / 2);   // pairs of blocks that pass
                        orr |= o;
                    }
                    pass = __ballot((orr & 0x8080) != 0) != 0ull;
                } else if constexpr (NV == 16) {
                    auto red = [](const acc& c) {
                        int m = min(min(__float_as_int(c[0]), __float_as_int(c[1])), __float_as_int(c[2]));
#pragma unroll
                        for (int v = 3; v < 15; v += 2) m = min(min(m, __float_as_int(c[v])), __float_as_int(c[v + 1]));
                        return min(m, __float_as_int(c[15]));
                    };
                    if constexpr (NP == 1) {
#pragma unroll
                        for (int b0 = 0; b0 < NB; b0 += 4) {
                            const acc p0 = prod(b0 + 0, bv);
                            const acc p1 = prod(b0 + 1, bv);
                            const acc p2 = prod(b0 + 2, bv);
                            const acc p3 = prod(b0 + 3, bv);
                            mn = min(mn, min(min(red(p0), red(p1)), min(red(p2), red(p3))));
                        }
                    } else {   // chained products in passes of two (two accumulators live)
#pragma unroll
                        for (int b0 = 0; b0 < NB; b0 += 2) {
                            const acc p0 = prod(b0 + 0, bv);
                            const acc p1 = prod(b0 + 1, bv);
                            mn = min(mn, min(red(p0), red(p1)));
                        }
                    }
                } else {
                    acc p[NB];
#pragma unroll
                    for (int b = 0; b < NB; ++b) p[b] = prod(b, bv);
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        mn = min(min(mn, min(__float_as_int(p[b][0]), __float_as_int(p[b][1]))),
                                 min(__float_as_int(p[b][2]), __float_as_int(p[b][3])));
                }
                // thr < 0 (a padded column) has a negative pattern below every non-negative one
                if constexpr (!PK) pass = __ballot(mn < __float_as_int(thr)) != 0ull;
                if (__builtin_expect(!pass, 1)) continue;
                // ---- exact tests of this group: the products are formed again, each lane
                // marks its values below its threshold (bit NV b + v), then tests those
                // (lane element, stream element) pairs; both elements are then real
                ++medium_entries;
                ++job_hits;
                int mine = QSF_WEIGHT_NONE;
                u64 cand[NCB] = {};
                // an opaque copy of the B operands: the products are formed again here instead
                // of all eight being kept live from the test above (which would spill)
                i32x4 bx[NP];
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    bx[p] = bv[p];
                    asm volatile("" : "+v"(bx[p]));
                }
                if constexpr (PK)
#pragma unroll
                    for (int b = 0; b < NB; ++b) {   // packed: bit 7 of field s marks element (row, s)
                    if (__ballot((pk_pair >> (b / 2)) & 1) == 0ull) continue;   // only the passing pairs
                    // the products without the bias (exact integers in f32), the fields added back
                    const acc p = P::mms(av[b][0], bx[0], acc{}, sa);
                    int o = 0;
#pragma unroll
                    for (int v = 0; v < NV; ++v) o |= (int)p[v] + cx;
                    if (__ballot((o & 0x8080) != 0) == 0ull) continue;
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const int q = 2 * (NV * b + v), bits = (int)p[v] + cx;
                        if (bits & 0x80) cand[q >> 6] |= 1ull << (q & 63);
                        if (bits & 0x8000) cand[(q + 1) >> 6] |= 1ull << ((q + 1) & 63);
                    }
                }
#pragma unroll
                for (int b = 0; b < NB && !PK; ++b) {
                    const acc p = prod(b, bx);
                    float pm = p[0];
#pragma unroll
                    for (int v = 1; v < NV; ++v) pm = fminf(pm, p[v]);
                    if (__ballot(pm < thr) == 0ull) continue;
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const int q = NV * b + v;
                        if (p[v] < thr) cand[q >> 6] |= 1ull << (q & 63);
                    }
                }
                for (int part = 0; part < NCB; ++part) {
                    for (u64 cm = cand[part]; cm; cm &= cm - 1) {
                        const int q = (64 * part + __builtin_ctzll(cm)) >> (PK ? 1 : 0);
                        const int half = PK ? (__builtin_ctzll(cm) & 1) : 0;
                        const int b = q / NV;
                        const u32 ij = lay[lane * NV + (q % NV)];
                        const u64 lmask = wlmk[NB * (PK ? 2 * (int)(ij & 31) + half : (int)(ij & 31)) + b];
                        const int pos = gi * NR + (int)(ij >> 8);
                        const u64 sm_ = smk[pos];
                        const int base = __popcll(lmask & lflags) + foff + __popcll(sm_ & sflags);
                        int tk = base;
                        {   // XD residual dwords of the pair's codeword
                            u32 xw[XD];
#pragma unroll
                            for (int f = 0; f < XD; ++f) xw[f] = (fx && f < RD) ? fx[f] : 0u;
                            for (u64 mm = lmask; mm; mm &= mm - 1) {
                                const int r = lbase + __builtin_ctzll(mm);
#pragma unroll
                                for (int f = 0; f < XD; ++f) xw[f] ^= rows_fast[r * XD + f];
                            }
                            for (u64 mm = sm_; mm; mm &= mm - 1) {
                                const int r = sbase + __builtin_ctzll(mm);
#pragma unroll
                                for (int f = 0; f < XD; ++f) xw[f] ^= rows_fast[r * XD + f];
                            }
#pragma unroll
                            for (int f = 0; f < XD; ++f) tk += __popc(xw[f]);
                        }
                        if (tk >= bound || tk >= mine) continue;
                        // exact weight and logical test of the pair
                        ++slow_entries;
                        int w = base;
                        for (int f = 0; f < RD && w < bound && w < mine; ++f) {
                            u32 x = fx ? fx[f] : 0u;
                            for (u64 mm = lmask; mm; mm &= mm - 1) x ^= slow_tbl[(lbase + __builtin_ctzll(mm)) * RW + f];
                            for (u64 mm = sm_; mm; mm &= mm - 1) x ^= slow_tbl[(sbase + __builtin_ctzll(mm)) * RW + f];
                            w += __popc(x);
                        }
                        if (w >= bound || w >= mine) continue;
                        bool logical = (SD == 0);
                        for (int f = RD; f < RW && !logical; ++f) {
                            u32 x = fx ? fx[f] : 0u;
                            for (u64 mm = lmask; mm; mm &= mm - 1) x ^= slow_tbl[(lbase + __builtin_ctzll(mm)) * RW + f];
                            for (u64 mm = sm_; mm; mm &= mm - 1) x ^= slow_tbl[(sbase + __builtin_ctzll(mm)) * RW + f];
                            logical = x != 0;
                        }
                        if (logical) mine = w;
                    }
                }
                const int wm = uni(wave_min(mine));
                if (wm < bound) {
                    bound = wm;
                    if (lane == 0) atomicMin(J->best, wm);
                    if (lane == 0 && J->shared != J->best) atomicMin(J->shared, wm);
                    thrf = (float)min(bound, 127);
                    tpk = min(bound, ptile);
                }
                if (bound <= floor) break;
            }
        }
    }
    if (lane == 0 && slow_entries) atomicAdd(a.slow_hits, (unsigned long long)slow_entries);
    if (lane == 0 && medium_entries) atomicAdd(a.slow_hits + 1, (unsigned long long)medium_entries);
    if (lane == 0 && job_hits) atomicAdd(J->hits, (unsigned long long)job_hits);
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
            if (a.collect) {   // the lanes of this step append together (one atomic per wave)
                const unsigned long long slot = wave_append(a.out_rows, wt >= 1 && wt < bound);
                if (slot < a.out_cap)
                    for (int w = 0; w < st; ++w) a.out[slot * st + w] = cw[w];
                continue;
            }
            if (wt >= bound) continue;
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
// bz_sweep_small: the small levels (< 2^16 codewords) of many problems in one launch
// =====================================================================================
//
// The row-sweep kernel above gives each thread a run of at least 64 combinations, which
// suits large levels; a level of a few thousand combinations then runs on a few hundred
// threads, each waiting on one global load after another (the d = 3 level of a [64, 32]
// code, two generators: 83 us for 9920 combinations each). Here each workgroup stages its
// generator's rows and the binomials of the unrank in LDS when they fit, each thread takes
// a run of at most 32 combinations, and every small level of a batched round shares one
// launch and one read-back. Same two-level sweep and the same result as bz_rowsweep.
__global__ __launch_bounds__(256) void bz_sweep_small(SweepJobsArgs a) {
    __shared__ u64 srows[kSweepRowsLds];
    __shared__ u64 sbin[kSweepBinLds];
    const unsigned b = blockIdx.x;
    int lo = 0, hi = a.njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.jobs[mid].block_base <= b) lo = mid; else hi = mid - 1;
    }
    const SJob& J = a.jobs[lo];
    const int K = J.K, D = J.d, M = D - 1, st = J.stride, bpg = J.blocks_per_g;
    const int g = (int)((b - J.block_base) / (unsigned)bpg);
    const int lb = (int)((b - J.block_base) - (u64)g * bpg);
    const u64* G = J.gamma + (size_t)g * K * st;
    const bool rows_lds = K * st <= kSweepRowsLds, bin_lds = K * D <= kSweepBinLds;
    if (rows_lds)
        for (int i = threadIdx.x; i < K * st; i += 256) srows[i] = G[i];
    if (bin_lds)
        for (int i = threadIdx.x; i < K * D; i += 256) sbin[i] = J.binom[(size_t)(i / D) * J.bk + (i % D)];
    __syncthreads();
    const u64* R = rows_lds ? srows : G;
    auto binom = [&](int n, int k) -> u64 {   // n in [0, K), k in [0, D)
        if (k > n) return 0ull;
        return bin_lds ? sbin[n * D + k] : J.binom[(size_t)n * J.bk + k];
    };

    const u64 start = ((u64)lb * 256 + threadIdx.x) * (u64)J.chunk;
    int mine = QSF_WEIGHT_NONE;
    if (start < J.total) {
        const u64 end = min(J.total, start + (u64)J.chunk);
        int pos[kMaxD];
        u64 base[16], cw[16];
        {   // lexicographic unrank of `start`
            u64 r = start;
            int x = 0;
            for (int i = 0; i < D; ++i)
                for (;;) {
                    const u64 cnt = binom(K - 1 - x, D - 1 - i);
                    if (r < cnt) { pos[i] = x++; break; }
                    r -= cnt;
                    ++x;
                }
        }
        for (int w = 0; w < st; ++w) base[w] = 0;
        for (int i = 0; i < M; ++i)
            for (int w = 0; w < st; ++w) base[w] ^= R[(size_t)pos[i] * st + w];
        int bound = __hip_atomic_load(J.best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int last = pos[D - 1];
        u64 done = 0;
        const u64 need = end - start;
        while (done < need) {
            int top = K;
            if ((u64)(K - last) > need - done) top = last + (int)(need - done);
            const int first = last;
            for (; last < top; ++last) {
                int wt = 0;
                for (int w = 0; w < st; ++w) {
                    cw[w] = base[w] ^ R[(size_t)last * st + w];
                    wt += __popcll(cw[w]);
                }
                if (wt >= bound) continue;
                bool logical = J.kcheck == 0;
                for (int c = 0; c < J.kcheck && !logical; ++c) {
                    u64 acc = 0;
                    for (int w = 0; w < st; ++w) acc ^= J.check[(size_t)c * st + w] & cw[w];
                    logical = __popcll(acc) & 1;
                }
                if (logical) { bound = wt; mine = wt; }
            }
            done += (u64)(top - first);
            if (done >= need) break;
            int j = M - 1;
            while (j >= 0 && pos[j] == K - D + j) --j;
            if (j < 0) break;
            for (int t = j; t < M; ++t)
                for (int w = 0; w < st; ++w) base[w] ^= R[(size_t)pos[t] * st + w];
            ++pos[j];
            for (int t = j + 1; t < M; ++t) pos[t] = pos[t - 1] + 1;
            for (int t = j; t < M; ++t)
                for (int w = 0; w < st; ++w) base[w] ^= R[(size_t)pos[t] * st + w];
            last = pos[M - 1] + 1;
        }
    }
    const int wm = wave_min(mine);
    if ((threadIdx.x & 63) == 0 && wm < QSF_WEIGHT_NONE) atomicMin(J.best, wm);
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
    int h = 0, RD = 0, SD = 0, RW = 0, bk = 0, maxres = 0;
    std::vector<int> nflag;       // flagged rows per generator
    // Measured structure factor of this solve's codewords: hot-test passes over the
    // fair-coin expectation at the last level that had enough expected passes to tell
    // (sf_logp: log of that level's fair per-codeword pass probability; sf = 0: not
    // measured yet). Deeper in the tail the factor grows: at a fair probability p it is
    // taken as sf (p / p_observed)^-kElasticity. The kernel choice scales pass rates by it.
    double sf = 0.0;
    double sf_logp = 0.0;
    double factor_at(double p) const {
        if (sf <= 0.0 || p <= 0.0) return tuning::kPriorFactor;
        return std::min(4096.0, std::max(1.0, sf * std::exp(-tuning::kElasticity * (std::log(p) - sf_logp))));
    }
    std::vector<u32> htbl;        // host copy of the residual+syndrome table
    std::vector<u64> hflags;      // host copy of the flag masks
    std::vector<u64> binom;       // (K+1) x bk
    // device copies, carved from one allocation: [gamma | check | binom | flags | tbl]; the
    // binomials only for K > kMaxTiledK (below that the kernels read the library's static
    // table, whose rows are dbk = kMaxD + 1 wide)
    struct Ptr { u64* p = nullptr; };
    struct Ptr32 { u32* p = nullptr; };
    DevBuf<u64> arena;
    std::vector<u64> stage;       // host image of the arena (source of its upload)
    size_t n_gamma = 0, n_check = 0, n_binom = 0, n_flags = 0;
    int dbk = 0;                  // row stride of the device binomial table
    Ptr gamma, check, dbinom, flags;
    Ptr32 tbl;
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
    static constexpr int kVariants = 24;  // kernel variants of a tiled run (vtab in run_tiled)
    u64* gbinom = nullptr;               // device address of tiled_binom (filled at probe)
    unsigned char* scratch_dev = nullptr;   // device address of scratch (no-allocation calls)
    DevBuf<unsigned char> blk;           // one tiled run's uploaded block (see run_tiled)
    unsigned char* pin = nullptr;        // its pinned host staging
    size_t pin_cap = 0;
    size_t cache_cap = 6;
    DevBuf<int> best;
    DevBuf<unsigned long long> counter;   // [0] tile counter, [1] slow-path entries
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
        for (int k = 0; k < 2 * kVariants; ++k) {   // start/stop events of a tiled run's launch groups
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) { probe_msg = "event creation failed"; return ok = false; }
            gev_dyn.push_back(e);
        }
        try {
            // the library's static device arrays on this device (the engine runs on the
            // device that is current here, and keeps using it)
            HIPCHK(hipGetSymbolAddress((void**)&gbinom, HIP_SYMBOL(tiled_binom)));
            HIPCHK(hipGetSymbolAddress((void**)&scratch_dev, HIP_SYMBOL(scratch)));
            const int bk = kMaxD + 1;
            std::vector<u64> B((size_t)(kMaxTiledK + 1) * bk, 0);
            for (int i = 0; i <= kMaxTiledK; ++i) {
                B[(size_t)i * bk] = 1;
                for (int j = 1; j < bk && j <= i; ++j)
                    B[(size_t)i * bk + j] = sat_add(B[(size_t)(i - 1) * bk + j - 1], j <= i - 1 ? B[(size_t)(i - 1) * bk + j] : 0);
            }
            static_assert(sizeof(tiled_binom) == (size_t)(kMaxTiledK + 1) * (kMaxD + 1) * sizeof(u64), "binomial table size");
            HIPCHK(hipMemcpy(gbinom, B.data(), B.size() * 8, hipMemcpyHostToDevice));
        } catch (const HipError& e) {
            probe_msg = e.what;
            return ok = false;
        }
        return ok = true;
    }

    // ---- no-allocation calls ------------------------------------------------------------
    // While `noalloc` is set (inside qsf_hip_enumerate_floor_noalloc) every byte of device
    // memory comes from `scratch`, laid out as
    //   [0, kScratchHead): the row sweep's bound and counters
    //   [kScratchHead, run_at): the tables of the one solve resident there (na_solve)
    //   [run_at, kScratchBytes): one kernel run's block (bounds, jobs, rectangles, ...)
    // and host staging is pageable memory (na_host). Nothing is allocated on the device or
    // pinned; a call that needs more than the array holds is refused.
    static constexpr size_t kScratchHead = 256;
    bool noalloc = false;
    std::unique_ptr<Solve> na_solve;      // host tables of the solve whose arena is in scratch
    size_t na_run_at = kScratchHead;      // first byte of scratch after that arena
    std::vector<unsigned char> na_host;   // pageable staging of a no-allocation run block

    Solve& solve_noalloc(const qsf_enum_plan& p) {
        if (na_solve) {
            const Solve& s = *na_solve;
            if (p.buffers_key != 0 && s.key == p.buffers_key && s.n == p.n && s.K == p.K &&
                s.ngt == p.num_gamma_total && s.kcheck == p.kcheck && s.stride == p.stride)
                return *na_solve;
        }
        na_solve.reset();
        na_run_at = kScratchHead;
        auto s = std::make_unique<Solve>();
        build_host(*s, p);
        const size_t bytes = s->stage.size() * 8;
        if (kScratchHead + bytes > kScratchBytes)
            throw HipError{4, "the solve's tables (" + std::to_string(bytes) + " bytes) exceed the " +
                                  std::to_string(kScratchBytes >> 20) + " MiB no-allocation scratch"};
        u64* d = reinterpret_cast<u64*>(scratch_dev + kScratchHead);
        s->gamma.p = d;
        s->check.p = d + s->n_gamma;
        s->dbinom.p = s->n_binom ? d + s->n_gamma + s->n_check : gbinom;
        s->flags.p = d + s->n_gamma + s->n_check + s->n_binom;
        s->tbl.p = reinterpret_cast<u32*>(d + s->n_gamma + s->n_check + s->n_binom + s->n_flags);
        // pageable source: the staging vector stays alive (in na_solve) until the stream drains
        HIPCHK(hipMemcpyAsync(d, s->stage.data(), bytes, hipMemcpyHostToDevice, stream));
        na_run_at = (kScratchHead + bytes + 255) & ~(size_t)255;
        na_solve = std::move(s);
        return *na_solve;
    }

    // Host staging and device block of one kernel run: pinned staging and a device buffer
    // grown on demand, or (no-allocation calls) pageable staging and the tail of scratch.
    void stage_block(size_t total, unsigned char*& host, unsigned char*& dev) {
        if (!noalloc) {
            pin_reserve(total);
            blk.reserve(total);
            host = pin;
            dev = blk.p;
            return;
        }
        if (na_run_at + total > kScratchBytes)
            throw HipError{4, "the level's run block (" + std::to_string(total) + " bytes) does not fit the " +
                                  std::to_string(kScratchBytes >> 20) + " MiB no-allocation scratch beside the solve's tables"};
        if (na_host.size() < total) na_host.resize(total);
        host = na_host.data();
        dev = scratch_dev + na_run_at;
    }
    int* dev_best() {
        if (noalloc) return reinterpret_cast<int*>(scratch_dev);
        best.reserve(1);
        return best.p;
    }
    unsigned long long* dev_counter() {
        if (noalloc) return reinterpret_cast<unsigned long long*>(scratch_dev + 64);
        counter.reserve(4);
        return counter.p;
    }

    Solve& solve_for(const qsf_enum_plan& p) {
        if (noalloc) return solve_noalloc(p);
        for (auto it = cache.begin(); it != cache.end(); ++it) {
            Solve& s = **it;
            if (p.buffers_key != 0 && s.key == p.buffers_key && s.n == p.n && s.K == p.K &&
                s.ngt == p.num_gamma_total && s.kcheck == p.kcheck && s.stride == p.stride) {
                cache.splice(cache.begin(), cache, it);
                return s;
            }
        }
        auto s = std::make_unique<Solve>();
        build_host(*s, p);
        evict(cache_cap > 0 ? cache_cap - 1 : 0);   // room for the new solve, which may take a leaver's arena
        std::vector<Solve*> one{s.get()};
        upload_all(one);
        cache.push_front(std::move(s));
        return *cache.front();
    }

    // Solves for many plans: the missing ones' host tables are built on several threads,
    // then allocated and uploaded one by one. cache_cap must exceed n.
    std::vector<Solve*> solves_for(const qsf_enum_plan* plans, int n) {
        std::vector<Solve*> out(n, nullptr);
        std::vector<std::pair<Solve*, const qsf_enum_plan*>> fresh;
        std::unordered_map<u64, Solve*> made;
        for (int j = 0; j < n; ++j) {
            const qsf_enum_plan& p = plans[j];
            for (auto it = cache.begin(); it != cache.end(); ++it) {
                Solve& c = **it;
                if (p.buffers_key != 0 && c.key == p.buffers_key && c.n == p.n && c.K == p.K &&
                    c.ngt == p.num_gamma_total && c.kcheck == p.kcheck && c.stride == p.stride) {
                    cache.splice(cache.begin(), cache, it);
                    out[j] = &c;
                    break;
                }
            }
            if (out[j]) continue;
            if (p.buffers_key != 0) {
                auto m = made.find(p.buffers_key);
                if (m != made.end()) { out[j] = m->second; continue; }
            }
            cache.push_front(std::make_unique<Solve>());
            out[j] = cache.front().get();
            fresh.emplace_back(out[j], &p);
            if (p.buffers_key != 0) made[p.buffers_key] = out[j];
        }
        std::vector<Solve*> up;
        for (auto& f : fresh) up.push_back(f.first);
        try {   // a solve whose build fails leaves the cache
            HostPool::get().run(fresh.size(), host_workers(fresh.size(), 2),
                                [&](size_t i) { build_host(*fresh[i].first, *fresh[i].second); });
            evict(cache_cap);   // older solves leave first (the new ones are at the front), so
            upload_all(up);     // that the new ones can take their arenas
        } catch (...) {
            for (Solve* f : up) cache.remove_if([&](const std::unique_ptr<Solve>& u) { return u.get() == f; });
            throw;
        }
        return out;
    }

    void build_host(Solve& s, const qsf_enum_plan& p) {
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
            // Residual column order; the first columns are the hot tests' bits. A column with
            // c ones among the K rows is 1 in the XOR of d random rows with probability
            // (1 - (1 - 2c/K)^d) / 2, so near-constant columns (c < K/8 or > 7K/8) go last,
            // densest first. The informative ones come first in a fixed pseudo-random order:
            // ordered densest-first, the columns of a structured code are correlated on its
            // light codewords (the gross code's Z side: 4.5e6 groups into the exact tests at
            // d = 10 instead of ~5e5). The weight is a column-order invariant, so results do
            // not depend on the order.
            std::stable_sort(resid[g].begin(), resid[g].end(), [&](int x, int y) {
                return std::abs(K - 2 * cnt[x]) < std::abs(K - 2 * cnt[y]);
            });
            const char* order_env = std::getenv("QSA_RESID_ORDER");   // tuning: "dense" or a seed
            const bool dense = order_env && std::strcmp(order_env, "dense") == 0;
            size_t informative = 0;
            while (informative < resid[g].size() && 4 * std::abs(K - 2 * cnt[resid[g][informative]]) <= 3 * K)
                ++informative;
            u64 z = (order_env && !dense ? std::strtoull(order_env, nullptr, 10) : 1ull) * 0x9E3779B97F4A7C15ull + (u64)g;
            for (size_t i = dense ? 0 : informative; i > 1; --i) {   // Fisher-Yates, splitmix64 stream
                z += 0x9E3779B97F4A7C15ull;
                u64 x = z;
                x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
                x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
                x ^= x >> 31;
                std::swap(resid[g][i - 1], resid[g][(size_t)(x % i)]);
            }
            maxres = std::max(maxres, (int)resid[g].size());
        }
        s.maxres = maxres;
        s.nflag.assign(s.ngt, 0);
        for (int g = 0; g < s.ngt; ++g)
            for (int r = 0; r < K; ++r) s.nflag[g] += flagged[g][r];
        s.RD = std::max(1, (maxres + 31) / 32);
        s.SD = (p.kcheck + 31) / 32;
        s.RW = s.RD + s.SD;
        std::vector<u32> tbl((size_t)s.ngt * K * s.RW, 0);
        std::vector<u64> fl((size_t)s.ngt * 4, 0);
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
                // four mask words per generator cover rows 0..255, the tiled kernels' K range
                if (flagged[g][r] && r < kMaxTiledK) fl[4 * g + (r >> 6)] |= 1ull << (r & 63);
            }
        }
        // one staging vector, one allocation, one copy
        const size_t n_gamma = (size_t)s.ngt * K * st, n_check = (size_t)std::max(1, p.kcheck) * st;
        // K <= kMaxTiledK: the static table (C(n <= 256, k <= 64)) serves the device
        const size_t n_binom = K <= kMaxTiledK ? 0 : s.binom.size(), n_flags = fl.size(), n_tbl = (tbl.size() + 1) / 2;
        s.dbk = n_binom ? s.bk : kMaxD + 1;
        std::vector<u64>& stage = s.stage;   // host image of the device arena
        stage.assign(n_gamma + n_check + n_binom + n_flags + n_tbl, 0);
        u64* at = stage.data();
        std::memcpy(at, p.gamma, n_gamma * 8);
        if (p.kcheck) std::memcpy(at + n_gamma, p.check, (size_t)p.kcheck * st * 8);
        if (n_binom) std::memcpy(at + n_gamma + n_check, s.binom.data(), n_binom * 8);
        std::memcpy(at + n_gamma + n_check + n_binom, fl.data(), n_flags * 8);
        if (!tbl.empty()) std::memcpy(at + n_gamma + n_check + n_binom + n_flags, tbl.data(), tbl.size() * 4);
        s.n_gamma = n_gamma; s.n_check = n_check; s.n_binom = n_binom; s.n_flags = n_flags;
        s.htbl = std::move(tbl);
        s.hflags = std::move(fl);
    }

    // Solves past the cache capacity leave; their device arenas go to a spare list that new
    // solves take from (a batch of new codes would otherwise pay one hipMalloc and one
    // hipFree per code). Eviction happens between calls, after the stream has drained.
    std::vector<std::pair<u64*, size_t>> spare_arenas;
    void evict(size_t keep) {
        while (cache.size() > keep) {
            Solve& s = *cache.back();
            if (s.arena.p && spare_arenas.size() < 1024) {
                spare_arenas.push_back({s.arena.p, s.arena.n});
                s.arena.p = nullptr;
                s.arena.n = 0;
            }
            cache.pop_back();
        }
    }
    void take_spare(DevBuf<u64>& a, size_t need) {
        if (a.n >= need) return;
        int pick = -1;
        for (int k = 0; k < (int)spare_arenas.size(); ++k)
            if (spare_arenas[k].second >= need && (pick < 0 || spare_arenas[k].second < spare_arenas[pick].second)) pick = k;
        if (pick < 0) return;
        if (a.p) (void)hipFree(a.p);
        a.p = spare_arenas[pick].first;
        a.n = spare_arenas[pick].second;
        spare_arenas[pick] = spare_arenas.back();
        spare_arenas.pop_back();
    }

    // Device arenas of new solves, filled from one pinned staging buffer (asynchronous DMA;
    // the buffer is reused once the previous uploads have completed).
    u64* up_pin = nullptr;
    size_t up_pin_cap = 0;   // u64
    hipEvent_t up_done = nullptr;
    void upload_all(const std::vector<Solve*>& ss) {
        if (ss.empty()) return;
        size_t total = 0;
        for (const Solve* s : ss) total += s->stage.size();
        if (up_done) HIPCHK(hipEventSynchronize(up_done));
        else HIPCHK(hipEventCreateWithFlags(&up_done, hipEventDisableTiming));
        if (total > up_pin_cap) {
            if (up_pin) (void)hipHostFree(up_pin);
            up_pin = nullptr;
            up_pin_cap = 0;
            HIPCHK(hipHostMalloc((void**)&up_pin, std::max<size_t>(total, 1) * 8));
            up_pin_cap = total;
        }
        size_t at = 0;
        for (Solve* sp : ss) {
            Solve& s = *sp;
            const std::vector<u64>& stage = s.stage;
            const size_t n_gamma = s.n_gamma, n_check = s.n_check, n_binom = s.n_binom, n_flags = s.n_flags;
            take_spare(s.arena, stage.size());
            s.arena.reserve(stage.size());
            u64* d = s.arena.p;
            s.gamma.p = d;
            s.check.p = d + n_gamma;
            s.dbinom.p = n_binom ? d + n_gamma + n_check : gbinom;
            s.flags.p = d + n_gamma + n_check + n_binom;
            s.tbl.p = reinterpret_cast<u32*>(d + n_gamma + n_check + n_binom + n_flags);
            std::memcpy(up_pin + at, stage.data(), stage.size() * 8);
            HIPCHK(hipMemcpyAsync(d, up_pin + at, stage.size() * 8, hipMemcpyHostToDevice, stream));
            at += stage.size();
        }
        HIPCHK(hipEventRecord(up_done, stream));
    }

    // ---- elimination: information sets and random-IS bounds, many problems per launch --
    DevBuf<u64> el_u64;        // gens, checks, set outputs
    DevBuf<int> el_int;        // bounds, set counts, ranks, partitions
    DevBuf<u64> el_scratch;
    DevBuf<ElimJob> el_jobs;
    unsigned char* el_pin = nullptr;   // pinned staging of an elimination call: [u64 area | int area | jobs]
    size_t el_pin_cap = 0;
    hipEvent_t el_b0 = nullptr, el_b1 = nullptr;   // random-IS kernels' start/stop
    qsf_hip_elim_stats el_stats{};

    int eliminate(qsf_elim_job* jobs, int njobs, int trials, u64 seed, int what) {
        el_stats = qsf_hip_elim_stats{};
        el_stats.jobs = njobs;
        const bool want_sets = (what & QSF_ELIM_SETS) != 0, want_bound = (what & QSF_ELIM_BOUND) != 0 && trials > 0;
        // layout: per job [gen | check | sets]  and  [best | nsets | ranks | part_off | part_cols]
        std::vector<size_t> u_off(njobs), i_off(njobs);
        size_t nu = 0, ni = 0, scratch_sets = 0, scratch_trial = 0;
        int max_n = 0;
        for (int j = 0; j < njobs; ++j) {
            const qsf_elim_job& q = jobs[j];
            const int st = (q.n + 63) / 64;
            u_off[j] = nu;
            nu += (size_t)q.K * st + (size_t)std::max(0, q.kcheck) * st + (want_sets && q.sets ? (size_t)q.set_cap * q.K * st : 0);
            i_off[j] = ni;
            ni += 2 + (want_sets ? (size_t)q.set_cap : 0) + (q.nparts > 0 ? (size_t)q.nparts + 1 + q.part_off[q.nparts] : 0) +
                  (q.order ? (size_t)q.n : 0) + (want_sets && q.set_of_col ? (size_t)q.n : 0);
            scratch_sets += elim_scratch_words(q.K, st, 1);
            scratch_trial += elim_scratch_words(q.K, st, 1);
            max_n = std::max(max_n, q.n);
        }
        // random-IS trials per launch: bounded so that the scratch stays under 1 GiB
        const int batch = !want_bound ? 0
                          : (int)std::max<size_t>(1, std::min<size_t>((size_t)trials, scratch_trial ? (1ull << 27) / scratch_trial : (size_t)trials));
        const size_t nscratch = std::max(scratch_sets, scratch_trial * (size_t)batch);
        // staging in pinned memory: the copies are asynchronous DMA, not runtime-staged
        const size_t o_hi = align16(nu * 8), o_hj = align16(o_hi + ni * 4);
        const size_t pin_bytes = o_hj + (size_t)njobs * sizeof(ElimJob) + 16;
        if (pin_bytes > el_pin_cap) {
            if (el_pin) (void)hipHostFree(el_pin);
            el_pin = nullptr;
            el_pin_cap = 0;
            HIPCHK(hipHostMalloc((void**)&el_pin, pin_bytes));
            el_pin_cap = pin_bytes;
        }
        std::memset(el_pin, 0, o_hj);
        u64* hu = reinterpret_cast<u64*>(el_pin);
        int* hi = reinterpret_cast<int*>(el_pin + o_hi);
        ElimJob* hj = reinterpret_cast<ElimJob*>(el_pin + o_hj);
        for (int j = 0; j < njobs; ++j) {
            const qsf_elim_job& q = jobs[j];
            const int st = (q.n + 63) / 64;
            std::memcpy(hu + u_off[j], q.gen, (size_t)q.K * st * 8);
            if (q.kcheck > 0) std::memcpy(hu + u_off[j] + (size_t)q.K * st, q.check, (size_t)q.kcheck * st * 8);
            int* h = hi + i_off[j];
            h[0] = q.init;
            int* po = h + 2 + (want_sets ? q.set_cap : 0);
            if (q.nparts > 0) {
                std::memcpy(po, q.part_off, (size_t)(q.nparts + 1) * 4);
                std::memcpy(po + q.nparts + 1, q.part_cols, (size_t)q.part_off[q.nparts] * 4);
                po += q.nparts + 1 + q.part_off[q.nparts];
            }
            if (q.order) std::memcpy(po, q.order, (size_t)q.n * 4);
        }
        el_u64.reserve(std::max<size_t>(nu, 1));
        el_int.reserve(std::max<size_t>(ni, 1));
        if (nscratch) el_scratch.reserve(nscratch);
        el_jobs.reserve(njobs);
        for (int j = 0; j < njobs; ++j) hj[j] = ElimJob{};
        size_t sc = 0;
        size_t lds_sets = 0, lds_trials = 0;
        for (int j = 0; j < njobs; ++j) {
            const qsf_elim_job& q = jobs[j];
            const int st = (q.n + 63) / 64;
            ElimJob& e = hj[j];
            u64* base = el_u64.p + u_off[j];
            int* ib = el_int.p + i_off[j];
            e.gen = base;
            e.check = base + (size_t)q.K * st;
            e.K = q.K; e.n = q.n; e.stride = st; e.kcheck = std::max(0, q.kcheck);
            e.every_nonzero = q.every_nonzero ? 1 : 0;
            e.best = ib;
            e.nsets = ib + 1;
            e.set_cap = want_sets ? q.set_cap : 0;
            e.sets = want_sets && q.sets ? base + (size_t)q.K * st + (size_t)e.kcheck * st : nullptr;
            e.ranks = want_sets ? ib + 2 : nullptr;
            e.nparts = q.nparts > 0 ? q.nparts : 0;
            e.part_off = e.nparts ? ib + 2 + e.set_cap : nullptr;
            e.part_cols = e.nparts ? ib + 2 + e.set_cap + e.nparts + 1 : nullptr;
            int* tail = ib + 2 + e.set_cap + (e.nparts ? e.nparts + 1 + q.part_off[q.nparts] : 0);
            e.order = q.order ? tail : nullptr;
            e.set_of_col = (want_sets && q.set_of_col) ? tail + (q.order ? q.n : 0) : nullptr;
            e.scratch = elim_scratch_words(q.K, st, 1) ? el_scratch.p + sc : nullptr;
            sc += elim_scratch_words(q.K, st, 1) * (size_t)std::max(1, batch);
            lds_sets = std::max(lds_sets, info_sets_lds_bytes(e));
            lds_trials = std::max(lds_trials, is_trials_lds_bytes(e));
            if (want_sets) el_stats.sets_bytes += (u64)q.K * st * 8;   // generator read once
        }
        {   // inputs: generator and detector rows; neighbouring jobs' ranges in one copy when
            // the output area between them is small (it is overwritten on the device)
            constexpr size_t kGap = 32768;   // u64
            size_t lo = 0, hi_end = 0;
            bool open = false;
            for (int j = 0; j < njobs; ++j) {
                const qsf_elim_job& q = jobs[j];
                const size_t w = (size_t)q.K * ((q.n + 63) / 64) + (size_t)std::max(0, q.kcheck) * ((q.n + 63) / 64);
                if (!w) continue;
                if (open && u_off[j] - hi_end <= kGap) {
                    hi_end = u_off[j] + w;
                    continue;
                }
                if (open) HIPCHK(hipMemcpyAsync(el_u64.p + lo, hu + lo, (hi_end - lo) * 8, hipMemcpyHostToDevice, stream));
                lo = u_off[j];
                hi_end = u_off[j] + w;
                open = true;
            }
            if (open) HIPCHK(hipMemcpyAsync(el_u64.p + lo, hu + lo, (hi_end - lo) * 8, hipMemcpyHostToDevice, stream));
        }
        HIPCHK(hipMemcpyAsync(el_int.p, hi, ni * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(el_jobs.p, hj, njobs * sizeof(ElimJob), hipMemcpyHostToDevice, stream));
        if (want_sets) {
            HIPCHK(hipEventRecord(ev0, stream));
            HIPCHK(launch_info_sets(el_jobs.p, njobs, lds_sets, stream));
            HIPCHK(hipEventRecord(ev1, stream));
        }
        if (want_bound) {
            if (!el_b0) HIPCHK(hipEventCreate(&el_b0));
            if (!el_b1) HIPCHK(hipEventCreate(&el_b1));
            HIPCHK(hipEventRecord(el_b0, stream));
            for (int t0 = 0; t0 < trials; t0 += batch)
                HIPCHK(launch_is_trials(el_jobs.p, njobs, lds_trials, seed, t0, std::min(batch, trials - t0), stream));
            HIPCHK(hipEventRecord(el_b1, stream));
        }
        if (want_sets) HIPCHK(hipMemcpyAsync(hu, el_u64.p, nu * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(hi, el_int.p, ni * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        float ms = 0;
        if (want_sets) { HIPCHK(hipEventElapsedTime(&ms, ev0, ev1)); el_stats.sets_ms = ms; }
        if (want_bound) {
            HIPCHK(hipEventElapsedTime(&ms, el_b0, el_b1));
            el_stats.bound_ms = ms;
            el_stats.trials = trials;
        }
        for (int j = 0; j < njobs; ++j) {
            qsf_elim_job& q = jobs[j];
            const int st = (q.n + 63) / 64;
            const int* h = hi + i_off[j];
            q.best = want_bound ? std::min(q.init, h[0]) : q.init;
            if (!want_sets) continue;
            q.nsets = h[1];
            if (q.set_of_col) {
                const int* tail = h + 2 + q.set_cap + (q.nparts > 0 ? q.nparts + 1 + q.part_off[q.nparts] : 0);
                std::memcpy(q.set_of_col, tail + (q.order ? q.n : 0), (size_t)q.n * 4);
            }
            const int keep = std::min(q.nsets, q.set_cap);
            std::memcpy(q.ranks, h + 2, (size_t)keep * 4);
            if (!q.sets) continue;
            std::memcpy(q.sets, hu + u_off[j] + (size_t)q.K * st + (size_t)std::max(0, q.kcheck) * st,
                        (size_t)keep * q.K * st * 8);
            el_stats.sets_bytes += (u64)keep * q.K * st * 8;   // systematic generators written
        }
        return 0;
    }

    // ---- min mode -------------------------------------------------------------------
    static bool tileable(const Solve& s, const qsf_enum_plan& p) {
        return p.K <= kMaxTiledK && s.RW <= kMaxRW && p.d <= 32;
    }

    int enumerate(const qsf_enum_plan& p, int floor, int* best_out) {
        stats = qsf_hip_stats{};
        Solve& s = solve_for(p);
        const u64 per_g = s.C(p.K, p.d);
        stats.codewords = sat_mul(per_g, (u64)p.num_gamma);
        if (per_g == 0 || p.num_gamma == 0) { *best_out = p.current_best; return 0; }
        if (stats.codewords > (1ull << 62)) throw HipError{4, "level has more than 2^62 codewords"};
        const bool tiled = tileable(s, p) && (policy == 2 || (policy == 0 && stats.codewords >= (1ull << 16)));
        if (tiled) {
            std::vector<TItem> items{TItem{&s, p, floor, choose_variant(s, p)}};
            std::vector<int> res;
            run_tiled(items, false, 0, 0, res);
            if (res[0] > floor) observe(s, p, items[0].V, item_hits[0]);   // ran to its end
            *best_out = std::min(res[0], p.current_best);
            return 0;
        }
        if (small_sweep(p, stats.codewords)) {
            std::vector<SItem> items{SItem{&s, &p}};
            std::vector<int> res;
            run_sweep_small(items, res);
            *best_out = std::min(res[0], p.current_best);
            return 0;
        }
        int* bp = dev_best();
        HIPCHK(hipMemcpyAsync(bp, &p.current_best, sizeof(int), hipMemcpyHostToDevice, stream));
        launch_sweep(s, p, false, 0, 0, bp, dev_counter());
        int result = 0;
        HIPCHK(hipMemcpyAsync(&result, bp, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.kernel_ms = ms;
        *best_out = std::min(result, p.current_best);
        return 0;
    }

    // qsf_hip_enumerate_floor on the scratch array only (see solve_noalloc).
    int enumerate_noalloc(const qsf_enum_plan& p, int floor, int* best_out) {
        if (!scratch_dev) throw HipError{3, "no-allocation scratch address unknown"};
        struct Reset {
            bool& f;
            ~Reset() { f = false; }
        } reset{noalloc};
        noalloc = true;
        try {
            return enumerate(p, floor, best_out);
        } catch (...) {
            // the failed call may have left the stream with work on scratch: drain it, and
            // forget the resident solve so the next call uploads its tables again
            (void)hipStreamSynchronize(stream);
            na_solve.reset();
            na_run_at = kScratchHead;
            throw;
        }
    }

    // Many problems' levels at once: every tileable level's tiles share one persistent
    // launch per kernel variant; levels the tiled kernel cannot take run on the row-sweep
    // kernel one after another.
    // groups (optional): plans with the same group index are levels of one problem and share
    // a running bound; skipped_out[j]: plan j's tiles skipped because that shared bound met
    // its floor through another plan (0 when groups is null).
    int enumerate_batch(const qsf_enum_plan* plans, const int* floors, int n, int* best_out,
                        const int* groups = nullptr, uint64_t* skipped_out = nullptr) {
        stats = qsf_hip_stats{};
        if (skipped_out) std::fill(skipped_out, skipped_out + n, (uint64_t)0);
        cache_cap = std::max(cache_cap, (size_t)n + 6);
        std::vector<Solve*> solves = solves_for(plans, n);   // cache_cap > n: none evicted
        std::vector<TItem> items;
        std::vector<SItem> small;
        std::vector<int> item_of(n, -1), small_of(n, -1), sweep;
        std::vector<int> phase(n, 0);   // a plan's position among its group's plans: its launch phase
        if (groups) {
            std::unordered_map<int, int> seen;
            for (int j = 0; j < n; ++j) phase[j] = seen[groups[j]]++;
        }
        for (int j = 0; j < n; ++j) {
            const qsf_enum_plan& p = plans[j];
            Solve& s = *solves[j];
            best_out[j] = p.current_best;
            const u64 per_g = s.C(p.K, p.d);
            const u64 cw = sat_mul(per_g, (u64)p.num_gamma);
            if (cw == 0) continue;
            if (cw > (1ull << 62)) throw HipError{4, "level has more than 2^62 codewords"};
            stats.codewords = sat_add(stats.codewords, cw);
            if (tileable(s, p) && policy != 1) {
                item_of[j] = (int)items.size();
                TItem it{&s, p, floors ? floors[j] : -1, choose_variant(s, p, true, matrix_min_work() >> 6),
                         groups ? groups[j] : -1, phase[j]};
                // a level after the first of its problem meets a bound the call cannot know;
                // below 2^26 codewords the choice costs more host time than it saves
                if (groups && phase[j] > 0 && cw >= (1ull << 26) && !std::getenv("QSA_TILED_VARIANT"))
                    candidate_segments(it);
                items.push_back(it);
            } else if (small_sweep(p, cw)) {
                small_of[j] = (int)small.size();
                small.push_back(SItem{&s, &plans[j]});
            } else {
                sweep.push_back(j);
            }
        }
        {   // a matrix variant whose levels together stay under the single-level minimum
            // goes back to the VALU kernel (a level counts once per variant among its
            // candidates)
            auto same_mm = [](const Variant& x, const Variant& y) { return x.mm == y.mm && x.xd == y.xd && x.rpl == y.rpl; };
            std::vector<std::pair<Variant, u64>> tot;
            auto add = [&](const Variant& v, u64 cw) {
                auto f = std::find_if(tot.begin(), tot.end(), [&](const std::pair<Variant, u64>& e) { return same_mm(e.first, v); });
                if (f == tot.end()) tot.push_back({v, cw}); else f->second = sat_add(f->second, cw);
            };
            auto small_total = [&](const Variant& v) {
                for (const auto& e : tot)
                    if (same_mm(e.first, v)) return e.second < matrix_min_work();
                return false;
            };
            for (const TItem& it : items) {
                const u64 cw = sat_mul(it.s->C(it.p.K, it.p.d), (u64)it.p.num_gamma);
                if (it.nseg == 0) {
                    if (it.V.mm != 0) add(it.V, cw);
                    continue;
                }
                for (int k = 0; k < it.nseg; ++k) {
                    bool seen = false;
                    for (int q = 0; q < k; ++q) seen = seen || same_variant(it.seg_v[q], it.seg_v[k]);
                    if (!seen && it.seg_v[k].mm != 0) add(it.seg_v[k], cw);
                }
            }
            for (TItem& it : items) {
                if (it.nseg == 0) {
                    if (it.V.mm != 0 && small_total(it.V)) it.V = choose_variant(*it.s, it.p, false);
                    continue;
                }
                for (int k = 0; k < it.nseg; ++k)
                    if (it.seg_v[k].mm != 0 && small_total(it.seg_v[k])) {
                        qsf_enum_plan pb = it.p;
                        pb.current_best = it.seg_max[k];
                        it.seg_v[k] = choose_variant(*it.s, pb, false);
                    }
                prune_segments(it);
            }
        }
        static const bool prof_items = std::getenv("QSA_PROFILE_ITEMS") != nullptr;   // diagnostics
        if (prof_items)
            for (const TItem& it : items)
                std::fprintf(stderr, "[prof item] solve %016llx d %d best %d thr %ld cw %llu phase %d variant %d,%d,%d,%d sf %.3f\n",
                             (unsigned long long)it.s->key, it.p.d, it.p.current_best, level_thr(*it.s, it.p),
                             (unsigned long long)sat_mul(it.s->C(it.p.K, it.p.d), (u64)it.p.num_gamma), it.phase,
                             it.V.fd, it.V.xd, it.V.rpl, it.V.mm, it.s->sf);
        double ms_total = 0;
        if (!items.empty()) {
            std::vector<int> res;
            run_tiled(items, false, 0, 0, res);
            ms_total += stats.kernel_ms;
            for (size_t i = 0; i < items.size(); ++i)
                if (res[i] > items[i].floor && item_skipped[i] == 0) observe(*items[i].s, items[i].p, items[i].V, item_hits[i]);
            if (skipped_out)
                for (int j = 0; j < n; ++j) skipped_out[j] = item_of[j] >= 0 ? item_skipped[item_of[j]] : 0;
            for (int j = 0; j < n; ++j)
                if (item_of[j] >= 0) best_out[j] = std::min(res[item_of[j]], plans[j].current_best);
        }
        if (!small.empty()) {
            std::vector<int> res;
            run_sweep_small(small, res);
            ms_total += stats.kernel_ms;
            for (int j = 0; j < n; ++j)
                if (small_of[j] >= 0) best_out[j] = std::min(res[small_of[j]], plans[j].current_best);
        }
        for (int j : sweep) {
            const qsf_enum_plan& p = plans[j];
            int* bp = dev_best();
            HIPCHK(hipMemcpyAsync(bp, &p.current_best, sizeof(int), hipMemcpyHostToDevice, stream));
            launch_sweep(*solves[j], p, false, 0, 0, bp, dev_counter());
            int result = 0;
            HIPCHK(hipMemcpyAsync(&result, bp, sizeof(int), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            ms_total += ms;
            best_out[j] = std::min(result, p.current_best);
        }
        stats.kernel_ms = ms_total;
        stats.kernel = !items.empty() ? 2 : !sweep.empty() ? 1 : !small.empty() ? 3 : 0;
        return 0;
    }

    // ---- small levels -----------------------------------------------------------------
    // Levels under 2^16 codewords (min mode, automatic policy) go to bz_sweep_small; the
    // row-sweep policy keeps them on bz_rowsweep, the second implementation of the tests.
    bool small_sweep(const qsf_enum_plan& p, u64 codewords) const {
        return policy == 0 && codewords < (1ull << 16) && p.d >= 1 && p.d <= kMaxD && p.stride <= 16;
    }

    struct SItem {
        Solve* s;
        const qsf_enum_plan* p;
    };

    // One bz_sweep_small launch for the levels of `items`: one upload of
    // [items x int bounds | items x SJob], one read-back of the bounds.
    void run_sweep_small(const std::vector<SItem>& items, std::vector<int>& res) {
        const int ni = (int)items.size();
        const size_t o_jobs = align16((size_t)ni * sizeof(int));
        const size_t total = align16(o_jobs + (size_t)ni * sizeof(SJob));
        unsigned char *host = nullptr, *dev = nullptr;
        stage_block(total, host, dev);
        int* hb = reinterpret_cast<int*>(host);
        SJob* hj = reinterpret_cast<SJob*>(host + o_jobs);
        u64 blocks = 0;
        for (int i = 0; i < ni; ++i) {
            const Solve& s = *items[i].s;
            const qsf_enum_plan& p = *items[i].p;
            hb[i] = p.current_best;
            SJob J{};
            J.gamma = s.gamma.p;
            J.check = s.check.p;
            J.binom = s.dbinom.p;
            J.best = reinterpret_cast<int*>(dev) + i;
            J.total = s.C(p.K, p.d);
            J.K = p.K; J.d = p.d; J.stride = p.stride; J.kcheck = p.kcheck; J.ngamma = p.num_gamma; J.bk = s.dbk;
            // runs of <= 32 combinations, >= 1024 threads per generator when the level has them
            J.chunk = (int)std::max<u64>(1, std::min<u64>(32, (J.total + 1023) / 1024));
            const u64 threads = (J.total + (u64)J.chunk - 1) / (u64)J.chunk;
            J.blocks_per_g = (int)((threads + 255) / 256);
            J.block_base = blocks;
            blocks += (u64)J.blocks_per_g * (u64)p.num_gamma;
            hj[i] = J;
        }
        HIPCHK(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, stream));
        SweepJobsArgs a{reinterpret_cast<const SJob*>(dev + o_jobs), ni};
        stats.kernel = 3;
        HIPCHK(hipEventRecord(ev0, stream));
        if (blocks > 0) {
            hipLaunchKernelGGL(bz_sweep_small, dim3((unsigned)blocks), dim3(256), 0, stream, a);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(ev1, stream));
        HIPCHK(hipMemcpyAsync(host, dev, (size_t)ni * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        res.assign(hb, hb + ni);
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.kernel_ms = ms;
    }

    // Kernel variant: FD residual dwords tested in the hot loop, XD held for the medium
    // test. The hot-loop test is chosen so that on fair-coin residual bits at most ~5% of
    // stream steps would reach the medium test; the medium test (2 FD dwords when
    // available) absorbs the extra light codewords of structured codes (stabilizer
    // elements), which would otherwise reach the slow path. thr = bound - (flagged rows
    // that any d-subset must contain).
    // mm: 0 VALU hot test (bz_tiled<fd, xd, rpl>), 1 int8 products (bz_mfma<xd>), 2 / 3 fp4
    // products of 32 x 32 / 16 x 16 codewords, 4 two chained 32 x 32 fp4 products
    // (bz_f4<xd, mm>); the matrix tests have fd = 0.
    // w: words per list subset mask (2 for K > 128, VALU kernel only).
    struct Variant {
        int fd, xd, rpl, mm;
        int w = 1;
    };
    static double binom_cdf_below(int bits, long thr) {   // P(Binomial(bits, 1/2) < thr)
        if (thr <= 0) return 0.0;
        if (thr > bits) return 1.0;
        double pmf = std::pow(0.5, bits), cdf = 0.0;
        for (long k = 0; k < thr; ++k) { cdf += pmf; pmf = pmf * (bits - k) / (k + 1); }
        return cdf;
    }
    // The matrix tests run 4-wave workgroups over tiles of 1024 (512) lane elements x 512
    // (256) stream elements: a launch of less than ~2^28 codewords cannot fill the chip,
    // and the VALU kernel's finer tiles win there. Above it the cost model decides. A
    // single level must hold that much work itself; in a batched run the levels of one
    // variant share a launch, so levels of >= 2^22 codewords qualify when their variant's
    // launch reaches 2^28 in total (enumerate_batch).
    static u64 matrix_min_work() {
        const char* mw = std::getenv("QSA_MATRIX_MIN_WORK");   // tuning/testing override
        return mw ? (u64)std::strtoull(mw, nullptr, 10) : (1ull << 28);
    }
    // Matrix-core tests (fd == 0): rpl = lane elements per wave / 64. min_work: smallest
    // level (codewords) a matrix test may take (default matrix_min_work()).
    static Variant choose_variant(const Solve& s, const qsf_enum_plan& p, bool allow_mfma = true,
                                  u64 min_work = ~0ull) {
        if (p.K <= kMaxMatrixK) return choose_variant_k128(s, p, allow_mfma, min_work);
        Variant v = choose_variant_k128(s, p, false, min_work);   // two-word masks: VALU kernel only
        v.w = 2;
        return v;
    }
    // The hot test of a variant: its unit (the matrix tests count groups of one wave's lane
    // blocks x 32 stream elements holding a value below the threshold: 8192 codewords, 4096
    // for the chained fp4 test; the VALU test stream steps of 64 x RPL codewords) and its
    // test bits (30 for int8, 60 for fp4 32x32x64, 124 for the 128-dimension fp4 tests,
    // 32 FD for VALU).
    static double test_unit(const Variant& V) {
        return V.mm == 5 ? 16384.0 : V.mm == 4 ? 4096.0 : V.mm > 0 ? 8192.0 : 64.0 * V.rpl;
    }
    static int test_bits(const Variant& V, int res) {
        return std::min(V.mm == 1 || V.mm == 5 ? 30 : V.mm == 2 ? 60 : V.mm >= 3 ? 124 : 32 * V.fd, res);
    }
    // Hot-test passes on a level of cw codewords with threshold thr if the residual bits
    // were fair coins (per-codeword pass probability p = P(Binomial(bits, 1/2) < thr)).
    static double fair_hits(const Variant& V, u64 cw, double p) {
        const double unit = test_unit(V);
        return (double)cw / unit * std::min(1.0, unit * p);
    }
    static int valu_index(const Variant& V) {   // position in tuning::kValuCost
        static const int fdxr[8][3] = {{1, 1, 32}, {1, 2, 16}, {1, 2, 24}, {1, 2, 32}, {2, 2, 16}, {2, 4, 12}, {3, 4, 12}, {4, 4, 12}};
        for (int k = 0; k < 8; ++k)
            if (fdxr[k][0] == V.fd && fdxr[k][1] == V.xd && fdxr[k][2] == V.rpl) return k;
        return 7;
    }
    // Predicted device seconds of a level on variant V: a launch constant, a cost per
    // codeword, and a cost per hot-test pass, the passes being the fair-coin expectation
    // scaled by the solve's structure factor; where a passing unit holds many passing
    // codewords (loose thresholds) the exact tests' work grows with the codewords, so the
    // count is max(passing units, passing codewords). Constants: tuning.hpp, measured by
    // qubitserf_amd/tune.py at about one passing codeword per passing unit.
    static double predict_s(const Solve& s, const Variant& V, u64 cw, long thr, int res) {
        const double p = binom_cdf_below(test_bits(V, res), thr);
        const double f = s.factor_at(p);
        const double hits = std::max(fair_hits(V, cw, p) * f, (double)cw * std::min(1.0, p * f));
        if (V.mm == 1) return tuning::kMatrixLaunch + (double)cw * tuning::kInt8Cost + hits * tuning::kInt8Hit;
        if (V.mm == 4) return tuning::kMatrixLaunch + (double)cw * tuning::kF4x2Cost + hits * tuning::kF4x2Hit;
        if (V.mm == 5) return tuning::kMatrixLaunch + (double)cw * tuning::kF4PkCost + hits * tuning::kF4PkHit;
        if (V.mm >= 2) return tuning::kMatrixLaunch + (double)cw * tuning::kF4Cost + hits * tuning::kF4Hit;
        const int k = valu_index(V);
        return tuning::kValuLaunch + (double)cw * tuning::kValuCost[k] + hits * tuning::kValuHit[k];
    }
    // The level's threshold on residual weight: bound minus the fewest flagged rows any
    // d-subset of an active generator contains.
    static long level_thr(const Solve& s, const qsf_enum_plan& p) {
        int min_off = p.d;
        for (int g = 0; g < p.num_gamma; ++g) min_off = std::min(min_off, std::max(0, p.d - (p.K - s.nflag[g])));
        return (long)p.current_best - min_off;
    }
    // After a tiled level ran to its end: the structure factor it measured, when the test
    // was in the tail (fewer than half of its units would pass on fair coins) and expected
    // at least 8 passes.
    static void observe(Solve& s, const qsf_enum_plan& p, const Variant& V, u64 hits) {
        if (V.mm == 3) return;
        const u64 cw = sat_mul(s.C(p.K, p.d), (u64)p.num_gamma);
        const double pr = binom_cdf_below(test_bits(V, std::max(1, s.maxres)), level_thr(s, p));
        const double e = fair_hits(V, cw, pr);
        if (e < 8.0 || test_unit(V) * pr > 0.5) return;
        s.sf = std::min(4096.0, std::max(1.0, (double)hits / e));
        s.sf_logp = std::log(pr);
    }
    static Variant choose_variant_k128(const Solve& s, const qsf_enum_plan& p, bool allow_mfma, u64 min_work) {
        static const Variant table[] = {{1, 1, 32, 0}, {1, 2, 16, 0}, {2, 2, 16, 0}, {2, 4, 12, 0}, {3, 4, 12, 0},
                                        {4, 4, 12, 0}, {1, 2, 24, 0}, {1, 2, 32, 0}, {0, 1, 4, 1},  {0, 2, 4, 1},
                                        {0, 4, 4, 1},  {0, 2, 4, 2},  {0, 4, 4, 2},  {0, 4, 2, 3},  {0, 4, 4, 4},
                                        {0, 4, 4, 5}};
        // the matrix tests compare their products with min(bound, 127); their padding
        // elements stay above any bound <= 100
        allow_mfma = allow_mfma && p.current_best <= 100;
        if (const char* env = std::getenv("QSA_TILED_VARIANT")) {   // tuning: "fd,xd,rpl[,mm]"
            int f = 0, x = 0, r = 0, m = -1;
            const int got = std::sscanf(env, "%d,%d,%d,%d", &f, &x, &r, &m);
            if (got >= 3)
                for (const Variant& v : table)
                    if (v.fd == f && v.xd == x && v.rpl == r && (got == 3 || v.mm == m) && (x <= s.RD || x == 4) &&
                        (v.mm == 0 || allow_mfma) && (v.mm < 2 || v.mm == 5 || s.RD >= (v.mm == 2 ? 2 : 3)) &&
                        (v.mm != 5 || p.current_best <= kPackMaxT))
                        return v;
        }
        if (const char* env = std::getenv("QSA_TILED_FD")) {   // tuning/testing override
            int f = std::atoi(env);
            if (f == 1) return s.RD == 1 ? table[0] : table[6];
            if (f == 2) return s.RD == 2 ? table[2] : table[3];
            if (f == 3) return table[4];
            if (f == 4) return table[5];
        }
        if (min_work == ~0ull) min_work = matrix_min_work();
        const u64 cw = sat_mul(s.C(p.K, p.d), (u64)p.num_gamma);
        allow_mfma = allow_mfma && cw >= min_work && !std::getenv("QSA_NO_MFMA");
        const long thr = level_thr(s, p);
        const int res = std::max(1, s.maxres);
        // Every applicable variant is priced by the measured cost model; the cheapest runs.
        // VALU variants test 32 FD of the RD residual dwords in the hot loop and hold XD.
        Variant best = table[0];
        double best_s = 1e300;
        auto consider = [&](const Variant& v) {
            const double t = predict_s(s, v, cw, thr, res);
            if (t < best_s) { best_s = t; best = v; }
        };
        if (s.RD == 1) {
            consider(table[0]);
        } else {
            consider(table[6]); consider(table[7]); consider(table[1]);
            consider(s.RD == 2 ? table[2] : table[3]);
            if (s.RD >= 3) { consider(table[4]); consider(table[5]); }
        }
        if (allow_mfma) {
            // int8 32x32x32 products over 30 test bits; fp4 32x32x64 products over 60, or two
            // of them chained over 124 (table[14]). The 16x16x128 fp4 form (124 bits,
            // table[13]) is not chosen automatically: its 256-element stream chunks pay the
            // per-chunk bound reload and barrier four times as often, and on the
            // quasi-cyclic evaluation codes it ran at half the VALU kernel's rate.
            consider(s.RD >= 3 ? table[10] : s.RD == 2 ? table[9] : table[8]);
            if (!std::getenv("QSA_NO_F4") && s.RD >= 2) consider(s.RD >= 3 ? table[12] : table[11]);
            if (!std::getenv("QSA_NO_F4") && s.RD >= 3) consider(table[14]);
            // the packed fp4 test: 30 test bits on two lane elements per product row (up to
            // bound kPackMaxT, where its 8-bit fields hold every value)
            // (QSA_NO_PACK=1 leaves it out: the int8 test's figures on the same levels)
            if (!std::getenv("QSA_NO_F4") && !std::getenv("QSA_NO_PACK") && p.current_best <= kPackMaxT)
                consider(table[15]);
        }
        return best;
    }

    // Decomposition of one level into products  F + subsets(L, lk) x subsets(Sr, sk).
    // The level's d-subsets of [0, K) split by how many rows fall in [0, h) and [h, K).
    // A product whose shorter list has fewer than kExpandBelow elements would rebuild the
    // 64 x RPL register tile for only a few stream steps; each element of that short list
    // is then made a fixed row set F and the long list's range is cut in two again,
    // which turns it into balanced products. Exact cover: every d-subset lies in exactly
    // one product.
    struct RectBuilder {
        const Solve& s;
        int g, RW, S;
        u64 lane_tile;
        bool matrix;   // a matrix-core test: orient each product by its padded work
        u64 flist_below = 0;   // > 0: products with a list this short take fixed-set lists
        std::vector<Rect> rv;
        std::vector<u32> fixed;
        static constexpr u64 kExpandBelow = 64;

        struct RowSet {   // fixed rows, rows 0..255
            u64 w[4];
        };

        void emit(const RowSet& f, int xb, int xm, int xk, int yb, int ym, int yk, u64 nx, u64 ny) {
            Rect r{};
            r.g = g;
            // the longer list on the lanes; with the MFMA test's 1024-element lane tiles, the
            // orientation that pads less (lane tiles x stream chunks)
            auto padded = [&](u64 nl, u64 ns) {
                return (double)((nl + lane_tile - 1) / lane_tile) * (double)lane_tile * (double)((ns + S - 1) / S) * (double)S;
            };
            const bool x_lanes = matrix ? padded(nx, ny) <= padded(ny, nx) : nx >= ny;
            r.lbase = x_lanes ? xb : yb; r.lm = x_lanes ? xm : ym; r.lk = x_lanes ? xk : yk;
            r.sbase = x_lanes ? yb : xb; r.sm = x_lanes ? ym : xm; r.sk = x_lanes ? yk : xk;
            r.n_lane = x_lanes ? nx : ny;
            r.n_stream = x_lanes ? ny : nx;
            r.lane_blocks = (r.n_lane + lane_tile - 1) / lane_tile;
            r.stream_chunks = (r.n_stream + S - 1) / S;
            r.fixed = -1;
            r.nfix = 0;
            for (int q = 0; q < 4; ++q) r.fmask[q] = f.w[q];
            if (f.w[0] | f.w[1] | f.w[2] | f.w[3]) {
                r.fixed = add_fixed(f);
                r.nfix = 1;
            }
            rv.push_back(r);
        }
        // One fixed-set entry: the XOR of its rows' residual + syndrome words, then its
        // flagged-row count. Returns its index.
        int add_fixed(const RowSet& f) {
            const int idx = (int)(fixed.size() / (RW + 1));
            const size_t at = fixed.size();
            fixed.resize(at + RW + 1, 0);
            int off = 0;
            for (int row = 0; row < s.K; ++row) {
                if (!((f.w[row >> 6] >> (row & 63)) & 1)) continue;
                const u32* src = s.htbl.data() + ((size_t)g * s.K + row) * RW;
                for (int w = 0; w < RW; ++w) fixed[at + w] ^= src[w];
                off += (int)((s.hflags[4 * g + (row >> 6)] >> (row & 63)) & 1);
            }
            fixed[at + RW] = (u32)off;
            return idx;
        }
        // The products F_e + subsets(x) x subsets(y) of every fixed set F_e of a list as one
        // rectangle for the VALU kernel: the longer list on the lanes, the stream every
        // (F_e, shorter-list subset) pair. common: the rows all F_e share.
        void emit_flist(const RowSet& common, const std::vector<RowSet>& fs, int xb, int xm, int xk, int yb, int ym,
                        int yk, u64 nx, u64 ny) {
            Rect r{};
            r.g = g;
            const bool x_lanes = nx >= ny;
            r.lbase = x_lanes ? xb : yb; r.lm = x_lanes ? xm : ym; r.lk = x_lanes ? xk : yk;
            r.sbase = x_lanes ? yb : xb; r.sm = x_lanes ? ym : xm; r.sk = x_lanes ? yk : xk;
            r.n_lane = x_lanes ? nx : ny;
            r.n_stream = (x_lanes ? ny : nx) * (u64)fs.size();
            r.lane_blocks = (r.n_lane + lane_tile - 1) / lane_tile;
            r.stream_chunks = (r.n_stream + S - 1) / S;
            for (int q = 0; q < 4; ++q) r.fmask[q] = common.w[q];
            r.fixed = -1;
            for (const RowSet& fe : fs) {
                const int idx = add_fixed(fe);
                if (r.fixed < 0) r.fixed = idx;
            }
            r.nfix = (int)fs.size();
            rv.push_back(r);
        }

        // Whether split() emits the product of xk-subsets of xm rows and yk-subsets of ym
        // rows (nx and ny subsets) whole at this depth, rather than expanding it.
        bool whole(u64 nx, u64 ny, int xm, int xk, int ym, int yk, int depth) const {
            const bool x_small = nx <= ny;
            const u64 ns = x_small ? nx : ny;
            const int bm = x_small ? ym : xm, bk = x_small ? yk : xk;
            // the matrix tests' 1024 x 512 tiles gain little from splitting below the first
            // level (K = 94, d = 9: 1.2e3 rectangles at 15% padding instead of 7.4e4 at 14%),
            // while every rectangle costs host time, upload and a longer tile search
            const int max_depth = matrix ? 1 : 3;
            // A product against a single subset (a level's ka = 0 or ka = d term: one list of
            // C(h, d) against the empty subset) is split down to 2^16 codewords: unsplit, the
            // VALU kernel tests each lane block against a one-element stream, building 768
            // lane elements per 768 codewords (C(28, 8) x 1: 3.1e6 codewords in 4048 such
            // blocks). Its halves are products of two lists, and splitting it makes no more
            // rectangles than halves. A short list of several subsets is left whole: its
            // split multiplies the rectangles by its length, each with a few hundred
            // codewords.
            const u64 floor_cw = ns == 1 ? (1ull << 16) : (1ull << 22);
            return ns >= kExpandBelow || bk == 0 || bm < 2 || depth >= max_depth || sat_mul(nx, ny) < floor_cw;
        }

        void split(const RowSet& f, int xb, int xm, int xk, int yb, int ym, int yk, int depth) {
            const u64 nx = s.C(xm, xk), ny = s.C(ym, yk);
            if (nx == 0 || ny == 0) return;
            if (whole(nx, ny, xm, xk, ym, yk, depth)) {
                emit(f, xb, xm, xk, yb, ym, yk, nx, ny);
                return;
            }
            // small side, enumerated: fewer than 64 subsets, so its range has fewer than 64
            // rows (or the subset is the whole range, of sk <= 32 rows); the big side split
            // in halves
            const bool x_small = nx <= ny;
            const u64 ns = x_small ? nx : ny;
            const int bb = x_small ? yb : xb, bm = x_small ? ym : xm, bk = x_small ? yk : xk;
            const int sb = x_small ? xb : yb, sk = x_small ? xk : yk;
            const int mid = bm / 2;
            std::vector<RowSet> fs(ns, f);   // f plus each subset of the small side, colex order
            {
                u64 m = sk == 0 ? 0 : ((1ull << sk) - 1);
                for (u64 e = 0; e < ns; ++e) {
                    for (u64 mm = m; mm; mm &= mm - 1) {
                        const int row = sb + __builtin_ctzll(mm);
                        fs[e].w[row >> 6] |= 1ull << (row & 63);
                    }
                    if (e + 1 < ns) {
                        const u64 c = m & (~m + 1ull), r = m + c;
                        m = (((r ^ m) >> 2) >> __builtin_ctzll(c)) | r;
                    }
                }
            }
            for (int j = std::max(0, bk - (bm - mid)); j <= std::min(bk, mid); ++j) {
                const u64 a2 = s.C(mid, j), b2 = s.C(bm - mid, bk - j);
                if (a2 == 0 || b2 == 0) continue;
                if (!whole(a2, b2, mid, j, bm - mid, bk - j, depth + 1)) {
                    for (const RowSet& fe : fs) split(fe, bb, mid, j, bb + mid, bm - mid, bk - j, depth + 1);
                } else if (flist_below > 0 && ns > 1 && std::min(a2, b2) < flist_below) {
                    // a child emitted whole whose shorter list is short: every fixed set of
                    // the small side in one rectangle
                    emit_flist(f, fs, bb, mid, j, bb + mid, bm - mid, bk - j, a2, b2);
                } else {
                    for (const RowSet& fe : fs) emit(fe, bb, mid, j, bb + mid, bm - mid, bk - j, a2, b2);
                }
            }
        }
    };

    struct TItem {
        Solve* s;
        qsf_enum_plan p;
        int floor;
        Variant V;
        int group = -1;   // items (levels) of one problem share a bound slot; -1: alone
        int phase = 0;    // launch phase: phases run in order, each after the previous one
        // Candidate variants of a level queued ahead of its bound (nseg >= 2, at least one a
        // matrix test): segment k's variant serves bounds <= seg_max[k]; V is the last.
        int nseg = 0;
        int seg_max[4] = {0, 0, 0, 0};
        Variant seg_v[4];
    };

    static bool same_variant(const Variant& x, const Variant& y) {
        return x.fd == y.fd && x.xd == y.xd && x.rpl == y.rpl && x.mm == y.mm && x.w == y.w;
    }
    // A level queued ahead of the bound it will meet: the variant the cost model picks at
    // each bound from the level's floor + 1 (at most kCandidateBounds below the current
    // bound) up to the current bound, as ascending segments; a bound below the lowest one
    // priced takes the lowest segment. Kept only when some segment is a matrix test (VALU
    // candidates alone measured slower: their variants differ little in cost, and each
    // candidate launch walks its passed-over tiles).
    static constexpr int kCandidateBounds = 8;
    void candidate_segments(TItem& it) const {
        const int hi = it.p.current_best;
        it.nseg = 0;
        if (hi >= QSF_WEIGHT_NONE || hi > 127) return;
        const int lo = std::max({1, it.floor + 1, hi - kCandidateBounds + 1});
        for (int b = lo; b <= hi; ++b) {
            qsf_enum_plan pb = it.p;
            pb.current_best = b;
            const Variant v = choose_variant(*it.s, pb, true, matrix_min_work() >> 6);
            if (it.nseg > 0 && same_variant(it.seg_v[it.nseg - 1], v)) {
                it.seg_max[it.nseg - 1] = b;
            } else if (it.nseg < 4) {
                it.seg_v[it.nseg] = v;
                it.seg_max[it.nseg++] = b;
            } else {   // more than four segments: the last one covers the rest
                it.seg_max[3] = b;
            }
        }
        prune_segments(it);
    }
    // Merges neighbouring segments of one variant; drops the candidates when none is a
    // matrix test or only one remains (the level then runs its top segment's variant).
    static void prune_segments(TItem& it) {
        int m = 0;
        bool matrix = false;
        for (int k = 0; k < it.nseg; ++k) {
            matrix = matrix || it.seg_v[k].mm > 0;
            if (m > 0 && same_variant(it.seg_v[m - 1], it.seg_v[k])) {
                it.seg_max[m - 1] = it.seg_max[k];
                continue;
            }
            it.seg_v[m] = it.seg_v[k];
            it.seg_max[m++] = it.seg_max[k];
        }
        if (m > 0) it.V = it.seg_v[m - 1];
        it.nseg = (m >= 2 && matrix) ? m : 0;
    }

    static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

    void pin_reserve(size_t bytes) {
        if (bytes <= pin_cap) return;
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_cap = 0;
        HIPCHK(hipHostMalloc((void**)&pin, bytes));
        pin_cap = bytes;
    }

    // One tiled run: the levels of `items` (job i = items[i]) decomposed into rectangles,
    // grouped by launch phase and kernel variant. One upload of a block laid out as
    //   [groups x 256-byte counter lines | items x int bounds | items x u64 passes |
    //    items x u64 skipped tiles | shared bounds | items x TJob | rects | fixed rows]
    // (each section from a 256-byte boundary), one persistent launch per group (tiles
    // pulled from the group's counter), one read-back of counters and bounds.
    // res[i] = the bound job i ended with.
    // shared_rows (collect mode): one device row counter for every launch of this run and
    // of earlier runs appending to the same output (levels 1..d of one collection);
    // nullptr: each variant counts its own rows from 0.
    void run_tiled(std::vector<TItem>& items, bool collect, int keep, u64 cap, std::vector<int>& res,
                   unsigned long long* shared_rows = nullptr) {
        const auto prof_t0 = std::chrono::steady_clock::now();   // host-time diagnostics (QSA_PROFILE_GROUPS)
        constexpr int NV = kVariants;
        static const Variant vtab[NV] = {
            {1, 1, 32, 0},    {1, 2, 16, 0},    {1, 2, 24, 0},    {1, 2, 32, 0},    {2, 2, 16, 0},    {2, 4, 12, 0},
            {3, 4, 12, 0},    {4, 4, 12, 0},    {0, 1, 4, 1},     {0, 2, 4, 1},     {0, 4, 4, 1},     {0, 2, 4, 2},
            {0, 4, 4, 2},     {0, 4, 2, 3},     {1, 1, 32, 0, 2}, {1, 2, 16, 0, 2}, {1, 2, 24, 0, 2}, {1, 2, 32, 0, 2},
            {2, 2, 16, 0, 2}, {2, 4, 12, 0, 2}, {3, 4, 12, 0, 2}, {4, 4, 12, 0, 2}, {0, 4, 4, 4}, {0, 4, 4, 5}};
        auto vidx = [&](const Variant& v) {
            for (int k = 0; k < NV; ++k)
                if (vtab[k].fd == v.fd && vtab[k].xd == v.xd && vtab[k].rpl == v.rpl && vtab[k].mm == v.mm &&
                    vtab[k].w == v.w)
                    return k;
            return v.w == 2 ? 21 : 7;
        };
        auto stream_len = [](const Variant& v) { return v.mm == 0 ? kStreamMax : v.mm == 1 ? kMfmaStream : f4_stream(v.mm); };
        auto lane_tile = [](const Variant& v) {
            return v.mm == 0 ? 64ull * v.rpl : v.mm == 1 ? (u64)kMfmaLane : (u64)f4_lane_tile(v.mm);
        };
        const int ni = (int)items.size();
        // Launch groups: the items of one (phase, variant) share a persistent launch; groups
        // launch phase by phase (the levels a problem may not need come in later phases and
        // are skipped once an earlier phase settles it), variants in order of first item.
        // A matrix-test level's thin products (a list shorter than thin_below, which would
        // leave the 1024-lane x 512-element tiles nearly empty) run on the level's VALU
        // variant instead, in a second group of the same phase: same job, same bound.
        std::vector<int> item_grp(ni), item_grp2(ni, -1), grp_v, grp_phase, grp_side;
        std::vector<std::vector<int>> cand_v(ni), cand_grp(ni);   // a level's candidate variants and their groups
        // (tuning override QSA_THIN_BELOW, 0: off; a variant forced by QSA_TILED_VARIANT
        // runs the whole level)
        const char* thin_env = std::getenv("QSA_THIN_BELOW");
        const u64 thin_below = std::getenv("QSA_TILED_VARIANT") ? 0
                               : thin_env ? (u64)std::strtoull(thin_env, nullptr, 10) : (u64)256;
        {
            std::vector<int> key;   // (phase * NV + variant) * 2 + side per group
            auto group_of = [&](int phase, int v, int side) {   // side: thin products only
                const int k = (phase * NV + v) * 2 + side;
                int g = (int)(std::find(key.begin(), key.end(), k) - key.begin());
                if (g == (int)key.size()) {
                    key.push_back(k);
                    grp_v.push_back(v);
                    grp_phase.push_back(phase);
                    grp_side.push_back(side);
                }
                return g;
            };
            std::vector<int> idx(ni);
            for (int i = 0; i < ni; ++i) idx[i] = i;
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return items[x].phase < items[y].phase; });
            for (int i : idx) {
                const TItem& it = items[i];
                if (!collect && it.nseg >= 2) {   // candidates: distinct variants in segment order
                    for (int k = 0; k < it.nseg; ++k) {
                        const int v = vidx(it.seg_v[k]);
                        if (std::find(cand_v[i].begin(), cand_v[i].end(), v) == cand_v[i].end()) cand_v[i].push_back(v);
                    }
                    for (int v : cand_v[i]) cand_grp[i].push_back(group_of(it.phase, v, 0));
                    item_grp[i] = cand_grp[i][0];
                } else {
                    item_grp[i] = group_of(it.phase, vidx(it.V), 0);
                }
                bool matrix = it.V.mm > 0;   // thin products leave a matrix test's launch
                for (int v : cand_v[i]) matrix = matrix || vtab[v].mm > 0;
                if (!collect && thin_below > 0 && matrix) {
                    // a candidate level's thin products: the VALU variant at its matrix
                    // candidates' highest bound (the bound at which it would leave them)
                    qsf_enum_plan pv = it.p;
                    for (int k = 0; k < it.nseg; ++k)
                        if (it.seg_v[k].mm > 0) pv.current_best = it.seg_max[k];
                    item_grp2[i] = group_of(it.phase, vidx(choose_variant(*it.s, pv, false)), 1);
                }
            }
        }
        // Job entries: item i's own (its only or first candidate's), one more per further
        // candidate, and one for its thin products; all share the level's bound, table and
        // skip count, each with its own pass counter.
        std::vector<std::vector<int>> cand_job(ni);
        int nj = ni;
        for (int i = 0; i < ni; ++i)
            for (size_t k = 0; k < cand_v[i].size(); ++k) cand_job[i].push_back(k == 0 ? i : nj++);
        std::vector<int> job2(ni, -1);
        for (int i = 0; i < ni; ++i)
            if (item_grp2[i] >= 0) job2[i] = nj++;
        const int NG = (int)grp_v.size();
        // stream chunk length per VALU variant: halved (down to 64) while the launch would
        // have fewer tiles than half the resident waves, so that a small or medium level
        // (one code per call) spreads over the chip instead of a few hundred waves; the
        // [64..160] random codes one per call: 241 -> 187 ms of device time (a target of 4
        // tiles per wave gave 200 ms and slowed the gross d = 7 level by 15%). Tuning
        // overrides: QSA_TILED_S (a fixed length), QSA_TILED_S_TILES (the target).
        std::vector<int> vS(NG);
        for (int g = 0; g < NG; ++g) vS[g] = stream_len(vtab[grp_v[g]]);
        {
            const char* env = std::getenv("QSA_TILED_S");
            std::vector<double> vcw(NG, 0.0);
            for (int i = 0; i < ni; ++i) {
                const double cw = (double)sat_mul(items[i].s->C(items[i].p.K, items[i].p.d), (u64)items[i].p.num_gamma);
                vcw[item_grp[i]] += cw;
                for (size_t k = 1; k < cand_grp[i].size(); ++k) vcw[cand_grp[i][k]] += cw;
            }
            const double resident = (double)std::max(1, cus) * 16;
            const char* per = std::getenv("QSA_TILED_S_TILES");   // tuning: tiles per resident wave
            const double want = per ? std::atof(per) : 0.5;
            for (int g = 0; g < NG; ++g) {
                const Variant& V = vtab[grp_v[g]];
                if (V.mm != 0) continue;
                if (grp_side[g]) {   // thin products: each short list in one chunk
                    vS[g] = (int)std::max<u64>(64, std::min<u64>(kStreamMax, thin_below));
                    continue;
                }
                if (env) {
                    vS[g] = std::max(64, std::min(kStreamMax, std::atoi(env)));
                } else {
                    int S = kStreamMax;
                    while (S > 64 && vcw[g] / (64.0 * V.rpl * S) < want * resident) S /= 2;
                    vS[g] = S;
                }
            }
        }
        std::vector<std::vector<Rect>> grect(NG);
        std::vector<u32> hfixed;
        std::vector<size_t> fixed_off(ni, 0);
        std::vector<int> glds(NG, 0);
        // each item's decomposition on its own (host threads for large batches), then
        // concatenated in item order
        std::vector<std::vector<Rect>> irect(ni), irect2(ni);
        std::vector<std::vector<std::vector<Rect>>> irectc(ni);   // per further candidate
        std::vector<std::vector<u32>> ifixed(ni);
        auto valu_geometry = [&](Rect v, int grp, int job) {   // a product on the VALU kernel
            const u64 lt = lane_tile(vtab[grp_v[grp]]);
            const u64 S2 = (u64)vS[grp];
            if (v.nfix <= 1 && v.n_lane < v.n_stream) {   // a fixed-set list stays on the stream
                std::swap(v.lbase, v.sbase); std::swap(v.lm, v.sm); std::swap(v.lk, v.sk);
                std::swap(v.n_lane, v.n_stream);
            }
            v.lane_blocks = (v.n_lane + lt - 1) / lt;
            v.stream_chunks = (v.n_stream + S2 - 1) / S2;
            v.job = job;
            return v;
        };
        auto decompose = [&](int i) {
            const TItem& it = items[i];
            const Solve& s = *it.s;
            const int K = it.p.K, d = it.p.d, h = s.h;
            // a level with candidates is decomposed once, for its first matrix candidate
            int dk = 0;
            for (size_t k = 0; k < cand_v[i].size(); ++k)
                if (vtab[cand_v[i][k]].mm > 0) { dk = (int)k; break; }
            const int dgrp = cand_v[i].empty() ? item_grp[i] : cand_grp[i][dk];
            const Variant& DV = vtab[grp_v[dgrp]];
            RectBuilder B{s, 0, s.RW, vS[dgrp], lane_tile(DV), DV.mm > 0};
            B.flist_below = collect ? 0 : thin_below;   // fixed-set lists: VALU kernel, min mode
            for (int g = 0; g < it.p.num_gamma; ++g) {
                B.g = g;
                for (int ka = std::max(0, d - (K - h)); ka <= std::min(d, h); ++ka)
                    B.split(RectBuilder::RowSet{}, 0, h, ka, h, K - h, d - ka, 0);
            }
            for (Rect& r : B.rv) r.job = i;
            if (item_grp2[i] >= 0) {   // thin products: VALU geometry, longer list on the lanes
                std::vector<Rect> fat;
                for (Rect& r : B.rv) {
                    if (r.nfix <= 1 && std::min(r.n_lane, r.n_stream) >= thin_below) fat.push_back(r);
                    else irect2[i].push_back(valu_geometry(r, item_grp2[i], job2[i]));
                }
                B.rv.swap(fat);
            }
            // the full products once per candidate: the matrix tests share the matrix
            // geometry, a VALU candidate takes them in its own
            irectc[i].resize(cand_v[i].size());
            for (size_t k = 0; k < cand_v[i].size(); ++k) {
                std::vector<Rect>& out = irectc[i][k];
                out.reserve(B.rv.size());
                for (const Rect& r : B.rv) {
                    if (vtab[cand_v[i][k]].mm > 0) {
                        Rect c = r;
                        c.job = cand_job[i][k];
                        out.push_back(c);
                    } else {
                        out.push_back(valu_geometry(r, cand_grp[i][k], cand_job[i][k]));
                    }
                }
            }
            irect[i] = cand_v[i].empty() ? std::move(B.rv) : std::vector<Rect>{};
            ifixed[i] = std::move(B.fixed);
        };
        const auto prof_t1 = std::chrono::steady_clock::now();
        HostPool::get().run((size_t)ni, host_workers((size_t)ni, 16), [&](size_t i) { decompose((int)i); });
        for (int i = 0; i < ni; ++i) {
            const TItem& it = items[i];
            const Solve& s = *it.s;
            const int vi = item_grp[i];
            grect[vi].insert(grect[vi].end(), irect[i].begin(), irect[i].end());
            fixed_off[i] = hfixed.size();
            hfixed.insert(hfixed.end(), ifixed[i].begin(), ifixed[i].end());
            if (it.p.K * s.RW <= kLdsFullRows) glds[vi] = std::max(glds[vi], it.p.K * s.RW);
            if (!irect2[i].empty()) {
                const int v2 = item_grp2[i];
                grect[v2].insert(grect[v2].end(), irect2[i].begin(), irect2[i].end());
                if (it.p.K * s.RW <= kLdsFullRows) glds[v2] = std::max(glds[v2], it.p.K * s.RW);
            }
            for (size_t k = 0; k < cand_v[i].size(); ++k) {
                const int vk = cand_grp[i][k];
                grect[vk].insert(grect[vk].end(), irectc[i][k].begin(), irectc[i][k].end());
                if (it.p.K * s.RW <= kLdsFullRows) glds[vk] = std::max(glds[vk], it.p.K * s.RW);
            }
        }
        item_thin.assign(ni, 0);
        for (int i = 0; i < ni; ++i) item_thin[i] = irect2[i].empty() ? 0 : 1;
        // tiles: a lane block x G stream chunks; G large enough to amortise the register
        // tile, small enough to leave >= 8 tiles per resident workgroup for load balance
        // (16 one-wave workgroups per CU for bz_tiled, 4 four-wave ones for bz_mfma)
        const u64 waves = (u64)std::max(1, cus) * 16;
        const u64 mfma_groups = (u64)std::max(1, cus) * 4;
        std::vector<int> gG(NG, 1);
        std::vector<u64> gtiles(NG, 0);
        size_t nrect = 0;
        for (int v = 0; v < NG; ++v) {
            u64 chunks = 0;
            for (const Rect& r : grect[v]) chunks += r.lane_blocks * r.stream_chunks;
            const u64 resident = vtab[grp_v[v]].mm > 0 ? mfma_groups : waves;
            gG[v] = (int)std::max<u64>(1, std::min<u64>(16, chunks / (8 * resident)));
            u64 tiles = 0;
            for (Rect& r : grect[v]) {
                r.stream_groups = (r.stream_chunks + gG[v] - 1) / gG[v];
                r.tile_base = tiles;
                tiles += r.lane_blocks * r.stream_groups;
            }
            gtiles[v] = tiles;
            nrect += grect[v].size();
        }
        static const bool prof_rects = std::getenv("QSA_PROFILE_RECTS") != nullptr;   // diagnostics
        if (prof_rects)
            for (int v = 0; v < NG; ++v) {
                const Variant& V = vtab[grp_v[v]];
                const double lt = (double)lane_tile(V), cap_tile = lt * vS[v] * gG[v];
                u64 thin = 0, lane_thin = 0;
                double real = 0, fill = 0;
                for (const Rect& r : grect[v]) {
                    const u64 t = r.lane_blocks * r.stream_groups;
                    if (r.n_stream < 128) thin += t;
                    if ((double)r.n_lane < lt) lane_thin += t;
                    real += (double)r.n_lane * (double)r.n_stream;
                    fill += (double)t * cap_tile;
                }
                std::fprintf(stderr, "[prof rects] phase %d variant %d: %zu rects, %llu tiles (%llu with < 128 stream, %llu with a partial lane tile), fill %.3f, %.4g codewords\n",
                             grp_phase[v], grp_v[v], grect[v].size(), (unsigned long long)gtiles[v], (unsigned long long)thin,
                             (unsigned long long)lane_thin, fill > 0 ? real / fill : 0.0, real);
            }
        int ngroups = 0, nchoice = 0;
        std::vector<int> choice_of(ni, -1);   // a candidate level's choice slot
        for (int i = 0; i < ni; ++i) {
            ngroups = std::max(ngroups, items[i].group + 1);
            if (!cand_v[i].empty()) choice_of[i] = nchoice++;
        }
        // The sections that the kernels update with atomics sit on cache lines of their own:
        // each group's tile counters (every workgroup increments next_tile at every tile),
        // the bounds (read by every wave at every stream chunk), the pass and skip counts
        // and the shared bounds. A tile counter on the line of the bounds it is polled beside
        // slowed the gross d = 10 level from 31.5 to 38.8 ms.
        constexpr size_t kLine = 256, kCntStride = kLine / sizeof(u64);
        auto line_up = [](size_t x) { return (x + kLine - 1) & ~(kLine - 1); };
        const size_t o_cnt = 0, o_best = (size_t)NG * kLine;
        const size_t o_hits = line_up(o_best + (size_t)ni * sizeof(int));
        const size_t o_skip = line_up(o_hits + (size_t)nj * sizeof(u64));
        const size_t o_shared = line_up(o_skip + (size_t)ni * sizeof(u64));
        const size_t o_choice = line_up(o_shared + (size_t)ngroups * sizeof(int));
        const size_t o_jobs = line_up(o_choice + (size_t)nchoice * sizeof(int));
        const size_t o_rects = align16(o_jobs + (size_t)nj * sizeof(TJob));
        const size_t o_fixed = align16(o_rects + nrect * sizeof(Rect));
        // per group: the rectangles' first tiles (u64 each) and the tile index (u32 per
        // 2^shift tiles, the shift keeping it under kTileIndex entries)
        constexpr u64 kTileIndex = 1u << 16;
        std::vector<int> gshift(NG, 0);
        std::vector<u64> gnix(NG, 0), gtb_off(NG, 0), gix_off(NG, 0);
        size_t ntb = 0, nix = 0;
        for (int v = 0; v < NG; ++v) {
            while ((gtiles[v] >> gshift[v]) > kTileIndex) ++gshift[v];
            gnix[v] = gtiles[v] ? ((gtiles[v] - 1) >> gshift[v]) + 1 : 0;
            gtb_off[v] = ntb;
            gix_off[v] = nix;
            ntb += grect[v].size();
            nix += gnix[v];
        }
        const size_t o_tb = align16(o_fixed + hfixed.size() * sizeof(u32) + 4);
        const size_t o_tix = align16(o_tb + ntb * sizeof(u64));
        const size_t total = align16(o_tix + nix * sizeof(u32) + 4);
        unsigned char *host = nullptr, *dev = nullptr;
        stage_block(total, host, dev);
        std::memset(host, 0, o_jobs);
        int* hb = reinterpret_cast<int*>(host + o_best);
        int* hs = reinterpret_cast<int*>(host + o_shared);
        for (int g = 0; g < ngroups; ++g) hs[g] = QSF_WEIGHT_NONE;
        for (const TItem& it : items)
            if (it.group >= 0) hs[it.group] = std::min(hs[it.group], it.p.current_best);
        int* hch = reinterpret_cast<int*>(host + o_choice);
        for (int c = 0; c < nchoice; ++c) hch[c] = -1;
        TJob* hj = reinterpret_cast<TJob*>(host + o_jobs);
        for (int i = 0; i < ni; ++i) {
            const TItem& it = items[i];
            const Solve& s = *it.s;
            hb[i] = collect ? keep + 1 : it.p.current_best;
            TJob J{};
            J.tbl = s.tbl.p;
            J.flags = s.flags.p;
            J.fixed = reinterpret_cast<const u32*>(dev + o_fixed) + fixed_off[i];
            J.gamma = s.gamma.p;
            J.best = reinterpret_cast<int*>(dev + o_best) + i;
            J.hits = reinterpret_cast<unsigned long long*>(dev + o_hits) + i;
            J.skipped = reinterpret_cast<unsigned long long*>(dev + o_skip) + i;
            J.shared = it.group >= 0 ? reinterpret_cast<int*>(dev + o_shared) + it.group : J.best;
            J.K = it.p.K; J.RD = s.RD; J.SD = s.SD; J.RW = s.RW; J.stride = it.p.stride;
            J.floor = it.floor;
            hj[i] = J;
            if (job2[i] >= 0) {   // thin products: always run
                TJob J2 = J;
                J2.hits = reinterpret_cast<unsigned long long*>(dev + o_hits) + job2[i];
                hj[job2[i]] = J2;
            }
            for (size_t k = 0; k < cand_v[i].size(); ++k) {
                TJob Jc = J;
                Jc.hits = reinterpret_cast<unsigned long long*>(dev + o_hits) + cand_job[i][k];
                Jc.choice = reinterpret_cast<int*>(dev + o_choice) + choice_of[i];
                Jc.variant = cand_v[i][k];
                Jc.nseg = it.nseg;
                for (int q = 0; q < 4; ++q) {
                    Jc.seg_max[q] = q < it.nseg ? it.seg_max[q] : 0;
                    Jc.seg_var[q] = q < it.nseg ? vidx(it.seg_v[q]) : cand_v[i][k];
                }
                hj[cand_job[i][k]] = Jc;
            }
        }
        std::vector<size_t> grect_off(NG, 0);
        {
            size_t at = o_rects;
            for (int v = 0; v < NG; ++v) {
                grect_off[v] = at;
                if (!grect[v].empty()) std::memcpy(host + at, grect[v].data(), grect[v].size() * sizeof(Rect));
                at += grect[v].size() * sizeof(Rect);
            }
        }
        if (!hfixed.empty()) std::memcpy(host + o_fixed, hfixed.data(), hfixed.size() * sizeof(u32));
        {
            u64* tb = reinterpret_cast<u64*>(host + o_tb);
            u32* tix = reinterpret_cast<u32*>(host + o_tix);
            for (int v = 0; v < NG; ++v) {
                const std::vector<Rect>& R = grect[v];
                for (size_t r = 0; r < R.size(); ++r) tb[gtb_off[v] + r] = R[r].tile_base;
                size_t r = 0;   // the rectangle holding tile q << shift, by a forward sweep
                for (u64 q = 0; q < gnix[v]; ++q) {
                    const u64 t = q << gshift[v];
                    while (r + 1 < R.size() && R[r + 1].tile_base <= t) ++r;
                    tix[gix_off[v] + q] = (u32)r;
                }
            }
        }
        const auto prof_t2 = std::chrono::steady_clock::now();
        HIPCHK(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, stream));

        auto* cnt = reinterpret_cast<unsigned long long*>(dev + o_cnt);
        stats.kernel = 2;
        stats.tiles = 0;
        std::vector<int> used(NG, 0);
        std::vector<u64> grp_side_cw(NG, 0);
        for (int v = 0; v < NG; ++v)
            if (grp_side[v])
                for (const Rect& r : grect[v]) grp_side_cw[v] += r.n_lane * r.n_stream;
        while (gev_dyn.size() < (size_t)(2 * NG)) {   // start/stop events per launch group
            hipEvent_t e = nullptr;
            HIPCHK(hipEventCreate(&e));
            gev_dyn.push_back(e);
        }
        double stats_cw = -1;
        for (int gi = 0; gi < NG; ++gi) {   // groups are numbered in launch order
            const int v = grp_v[gi];
            if (gtiles[gi] == 0) continue;
            used[gi] = 1;
            const Variant& V = vtab[v];
            TiledArgs a{};
            a.jobs = reinterpret_cast<const TJob*>(dev + o_jobs);
            a.njobs = ni;   // levels (distinct bounds): a second job entry shares its level's bound
            a.rects = reinterpret_cast<const Rect*>(dev + grect_off[gi]);
            a.nrects = (int)grect[gi].size();
            a.total_tiles = gtiles[gi];
            a.tile_base = reinterpret_cast<const u64*>(dev + o_tb) + gtb_off[gi];
            a.tile_ix = reinterpret_cast<const unsigned*>(dev + o_tix) + gix_off[gi];
            a.n_ix = (unsigned)gnix[gi];
            a.ix_shift = gshift[gi];
            a.binom = gbinom;
            a.bk = kMaxD + 1;
            a.S = vS[gi];
            a.G = gG[gi];
            a.lds_full = glds[gi];
            a.next_tile = cnt + kCntStride * gi;
            a.slow_hits = cnt + kCntStride * gi + 1;
            a.collect = collect ? 1 : 0;
            a.keep = keep;
            a.out_rows = shared_rows ? shared_rows : cnt + kCntStride * gi + 3;
            a.out = out.p;
            a.out_cap = cap;
            const u64 grid = std::min<u64>(gtiles[gi], V.mm > 0 ? mfma_groups : waves);
            const dim3 gd((unsigned)grid), bd(V.mm > 0 ? 64 * kMfmaWaves : 64);
            const size_t lds = (size_t)glds[gi] * sizeof(u32);
            HIPCHK(hipEventRecord(gev_dyn[2 * gi], stream));
            switch (v) {
                case 0: hipLaunchKernelGGL((bz_tiled<1, 1, 32, 1>), gd, bd, lds, stream, a); break;
                case 1: hipLaunchKernelGGL((bz_tiled<1, 2, 16, 1>), gd, bd, lds, stream, a); break;
                case 2: hipLaunchKernelGGL((bz_tiled<1, 2, 24, 1>), gd, bd, lds, stream, a); break;
                case 3: hipLaunchKernelGGL((bz_tiled<1, 2, 32, 1>), gd, bd, lds, stream, a); break;
                case 4: hipLaunchKernelGGL((bz_tiled<2, 2, 16, 1>), gd, bd, lds, stream, a); break;
                case 5: hipLaunchKernelGGL((bz_tiled<2, 4, 12, 1>), gd, bd, lds, stream, a); break;
                case 6: hipLaunchKernelGGL((bz_tiled<3, 4, 12, 1>), gd, bd, lds, stream, a); break;
                case 7: hipLaunchKernelGGL((bz_tiled<4, 4, 12, 1>), gd, bd, lds, stream, a); break;
                case 8: hipLaunchKernelGGL((bz_mfma<1>), gd, bd, lds, stream, a); break;
                case 9: hipLaunchKernelGGL((bz_mfma<2>), gd, bd, lds, stream, a); break;
                case 10: hipLaunchKernelGGL((bz_mfma<4>), gd, bd, lds, stream, a); break;
                case 11: hipLaunchKernelGGL((bz_f4<2, 2>), gd, bd, lds, stream, a); break;
                case 12: hipLaunchKernelGGL((bz_f4<4, 2>), gd, bd, lds, stream, a); break;
                case 13: hipLaunchKernelGGL((bz_f4<4, 3>), gd, bd, lds, stream, a); break;
                case 14: hipLaunchKernelGGL((bz_tiled<1, 1, 32, 2>), gd, bd, lds, stream, a); break;
                case 15: hipLaunchKernelGGL((bz_tiled<1, 2, 16, 2>), gd, bd, lds, stream, a); break;
                case 16: hipLaunchKernelGGL((bz_tiled<1, 2, 24, 2>), gd, bd, lds, stream, a); break;
                case 17: hipLaunchKernelGGL((bz_tiled<1, 2, 32, 2>), gd, bd, lds, stream, a); break;
                case 18: hipLaunchKernelGGL((bz_tiled<2, 2, 16, 2>), gd, bd, lds, stream, a); break;
                case 19: hipLaunchKernelGGL((bz_tiled<2, 4, 12, 2>), gd, bd, lds, stream, a); break;
                case 20: hipLaunchKernelGGL((bz_tiled<3, 4, 12, 2>), gd, bd, lds, stream, a); break;
                case 21: hipLaunchKernelGGL((bz_tiled<4, 4, 12, 2>), gd, bd, lds, stream, a); break;
                case 22: hipLaunchKernelGGL((bz_f4<4, 4>), gd, bd, lds, stream, a); break;
                default: hipLaunchKernelGGL((bz_f4<4, 5>), gd, bd, lds, stream, a); break;
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(gev_dyn[2 * gi + 1], stream));
            stats.tiles = (int)std::min<u64>((u64)stats.tiles + gtiles[gi], 0x7fffffff);
            // the run's kernel in the stats: the variant chosen for the most codewords (a
            // matrix-test level counts for its matrix variant, thin products included)
            double gcw = 0;
            for (int i = 0; i < ni; ++i)
                if (item_grp[i] == gi) gcw += (double)sat_mul(items[i].s->C(items[i].p.K, items[i].p.d), (u64)items[i].p.num_gamma);
            if (gcw > stats_cw) {
                stats_cw = gcw;
                stats.fd = V.fd;
                stats.xd = V.xd;
                stats.rpl = V.rpl;
                stats.mfma = V.mm;
                // VALU lane-ops per codeword of the hot loop; the matrix tests spend 0.5 (v_min3
                // over their products) besides the products themselves (32 int8 multiply-adds per
                // codeword, or 64 / 128 fp4 ones)
                stats.ops_per_codeword = V.mm > 0 ? 0.5 : (2.0 * V.fd * V.rpl + V.rpl / 2.0 + 3.0) / V.rpl;
            }
        }
        const auto prof_t3 = std::chrono::steady_clock::now();
        HIPCHK(hipMemcpyAsync(host, dev, o_jobs, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (std::getenv("QSA_PROFILE_GROUPS")) {
            auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
                return std::chrono::duration<double, std::milli>(b - a).count();
            };
            std::fprintf(stderr, "[prof run] %d items, %zu rects, %zu bytes: prepare %.3f ms, decompose+block %.3f ms, launch %.3f ms, wait %.3f ms\n",
                         ni, nrect, total, ms(prof_t0, prof_t1), ms(prof_t1, prof_t2), ms(prof_t2, prof_t3),
                         ms(prof_t3, std::chrono::steady_clock::now()));
        }
        res.assign(hb, hb + ni);
        const unsigned long long* hh = reinterpret_cast<const unsigned long long*>(host + o_hits);
        item_hits.assign(hh, hh + ni);
        const unsigned long long* hk = reinterpret_cast<const unsigned long long*>(host + o_skip);
        item_skipped.assign(hk, hk + ni);
        for (int i = 0; i < ni; ++i) {   // a candidate level: the variant picked and its passes
            if (cand_v[i].empty()) continue;
            const int c = hch[choice_of[i]];
            for (size_t k = 0; k < cand_v[i].size(); ++k)
                if (cand_v[i][k] == c) {
                    items[i].V = vtab[c];
                    item_hits[i] = hh[cand_job[i][k]];
                }
            items[i].nseg = 0;
        }
        const unsigned long long* hc = reinterpret_cast<const unsigned long long*>(host + o_cnt);
        double ms_total = 0;
        stats.slow_hits = stats.medium_hits = 0;
        stats.side_ms = 0;
        stats.side_codewords = 0;
        stats.side_ops_per_codeword = 0;
        double side_ops = 0;
        out_rows_last = 0;
        static const bool prof_groups = std::getenv("QSA_PROFILE_GROUPS") != nullptr;   // diagnostics
        for (int gi = 0; gi < NG; ++gi) {
            if (!used[gi]) continue;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, gev_dyn[2 * gi], gev_dyn[2 * gi + 1]));
            ms_total += ms;
            if (grp_side[gi]) {   // a launch of thin products only: time and codewords beside the matrix test
                const Variant& V = vtab[grp_v[gi]];
                stats.side_ms += ms;
                stats.side_codewords += grp_side_cw[gi];
                side_ops += (2.0 * V.fd * V.rpl + V.rpl / 2.0 + 3.0) / V.rpl * (double)grp_side_cw[gi];
            }
            if (prof_groups) {
                u64 cw = 0;
                int nitems = 0;
                for (int i = 0; i < ni; ++i)
                    if (item_grp[i] == gi) { cw += sat_mul(items[i].s->C(items[i].p.K, items[i].p.d), (u64)items[i].p.num_gamma); ++nitems; }
                std::fprintf(stderr, "[prof group] phase %d variant %d%s: %d levels, %llu codewords, %llu tiles, S %d G %d, %.3f ms, exact %llu medium %llu\n",
                             grp_phase[gi], grp_v[gi], grp_side[gi] ? " (thin products)" : "", nitems, (unsigned long long)cw, (unsigned long long)gtiles[gi], vS[gi], gG[gi], ms,
                             hc[kCntStride * gi + 1], hc[kCntStride * gi + 2]);
            }
            stats.slow_hits += hc[kCntStride * gi + 1];
            stats.medium_hits += hc[kCntStride * gi + 2];
            if (vtab[grp_v[gi]].mm >= 2) {   // the fp4 kernels count launches that met another output layout here
                if (hc[kCntStride * gi + 3] != 0)
                    throw HipError{3, "fp4 matrix products: output layout or block scales differ from the ones the hot "
                                      "test assumes, or the packed test met a bound above its limit"};
            } else {
                out_rows_last += hc[kCntStride * gi + 3];
            }
        }
        if (stats.side_codewords) stats.side_ops_per_codeword = side_ops / (double)stats.side_codewords;
        stats.kernel_ms = ms_total;
    }
    unsigned long long out_rows_last = 0;   // collect mode: rows the last tiled run produced
    DevBuf<u64> sorted;                      // collect_upto: the deduplicated rows
    std::vector<u64> item_hits;             // the last tiled run's hot-test passes per item
    std::vector<u64> item_skipped;          // its tiles skipped through a shared bound, per item
    std::vector<char> item_thin;            // a matrix-test item whose thin products ran on the VALU kernel
    std::vector<hipEvent_t> gev_dyn;        // start/stop events of a tiled run's launch groups

    void launch_sweep(Solve& s, const qsf_enum_plan& p, bool collect, int keep, u64 cap, int* bp,
                      unsigned long long* cntp) {
        if (p.stride > 16) throw HipError{4, "row-sweep kernel supports n <= 1024"};
        if (p.d > kMaxD) throw HipError{4, "weight level above 64"};
        SweepArgs a{};
        a.gamma = s.gamma.p; a.check = s.check.p; a.kcheck = p.kcheck; a.stride = p.stride;
        a.K = p.K; a.d = p.d; a.ngamma = p.num_gamma; a.binom = s.dbinom.p; a.bk = s.dbk;
        a.total = s.C(p.K, p.d);
        const u64 target_threads = (u64)std::max(1, cus) * 2048;
        u64 chunk = std::max<u64>(64, (a.total * (u64)p.num_gamma + target_threads - 1) / target_threads);
        a.chunk = chunk;
        a.best = bp; a.collect = collect ? 1 : 0; a.keep = keep;
        a.out_rows = cntp + 3; a.out = out.p; a.out_cap = cap;
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

    // ---- collect mode, levels 1..d at once -------------------------------------------
    // Every level's rows are appended on the device to one buffer (one shared counter, one
    // atomic per wave and append), then sorted and deduplicated on the device; one copy back.
    int collect_upto(const qsf_enum_plan& p0, int keep, uint64_t** hits, int64_t* rows) {
        stats = qsf_hip_stats{};
        *hits = nullptr;
        *rows = 0;
        Solve& s = solve_for(p0);
        const int st = p0.stride;
        for (int d = 1; d <= p0.d; ++d) stats.codewords = sat_add(stats.codewords, sat_mul(s.C(p0.K, d), (u64)p0.num_gamma));
        if (keep <= 0 || p0.num_gamma == 0 || stats.codewords == 0) return 0;
        if (stats.codewords > (1ull << 62)) throw HipError{4, "collection has more than 2^62 codewords"};
        int* bp = dev_best();
        unsigned long long* cntp = dev_counter();
        u64 cap = std::max<u64>(1ull << 16, collect_cap);
        for (;;) {
            out.reserve(cap * (u64)st);
            HIPCHK(hipMemsetAsync(cntp, 0, 4 * sizeof(unsigned long long), stream));
            double ms_total = 0;
            // the tiled levels as the items of one run (a launch per variant, not per level)
            std::vector<TItem> items;
            for (int d = 1; d <= p0.d; ++d) {
                qsf_enum_plan p = p0;
                p.d = d;
                const u64 cw = sat_mul(s.C(p.K, d), (u64)p.num_gamma);
                if (cw == 0) continue;
                if (tileable(s, p) && (policy == 2 || (policy == 0 && cw >= (1ull << 16)))) {
                    qsf_enum_plan pv = p;
                    pv.current_best = keep + 1;
                    items.push_back(TItem{&s, p, -1, choose_variant(s, pv, false)});
                } else {
                    launch_sweep(s, p, true, keep, cap, bp, cntp);
                    HIPCHK(hipStreamSynchronize(stream));
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
                    ms_total += ms;
                }
            }
            if (!items.empty()) {
                std::vector<int> res;
                run_tiled(items, true, keep, cap, res, cntp + 3);
                ms_total += stats.kernel_ms;
            }
            unsigned long long got = 0;
            HIPCHK(hipMemcpyAsync(&got, cntp + 3, sizeof(got), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            stats.kernel_ms = ms_total;
            if (got > cap) { cap = got; continue; }   // overflow: run again with room for all
            collect_cap = std::max<u64>(collect_cap, got);
            u64 kept = 0;
            if (got) {
                sorted.reserve(got * (u64)st);
                std::string why;
                HIPCHK(hipEventRecord(ev0, stream));
                if (int rc = sort_dedup_rows(out.p, got, st, sorted.p, &kept, stream, why)) throw HipError{rc, why};
                HIPCHK(hipEventRecord(ev1, stream));
                HIPCHK(hipEventSynchronize(ev1));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
                sort_ms = ms;
            }
            u64* h = (u64*)std::malloc(std::max<u64>(1, kept) * (u64)st * 8);
            if (!h) throw HipError{3, "host allocation failed"};
            if (kept) HIPCHK(hipMemcpy(h, sorted.p, kept * (u64)st * 8, hipMemcpyDeviceToHost));
            appended_last = got;
            *hits = h;
            *rows = (int64_t)kept;
            return 0;
        }
    }
    u64 collect_cap = 0;           // largest collection so far (the next one's first capacity)
    u64 appended_last = 0;         // rows the last collect_upto appended before deduplication
    double sort_ms = 0;            // its device sort + dedup time

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
        int* bp = dev_best();
        unsigned long long* cntp = dev_counter();
        const bool tiled = tileable(s, p) && (policy == 2 || (policy == 0 && stats.codewords >= (1ull << 16)));
        u64 cap = 1 << 16;
        for (;;) {
            out.reserve(cap * (u64)p.stride);
            unsigned long long got = 0;
            if (tiled) {
                qsf_enum_plan pv = p;
                pv.current_best = keep + 1;
                std::vector<TItem> items{TItem{&s, p, -1, choose_variant(s, pv, false)}};
                std::vector<int> res;
                run_tiled(items, true, keep, cap, res);
                got = out_rows_last;
            } else {
                HIPCHK(hipMemsetAsync(cntp, 0, 4 * sizeof(unsigned long long), stream));
                launch_sweep(s, p, true, keep, cap, bp, cntp);
                HIPCHK(hipMemcpyAsync(&got, cntp + 3, sizeof(got), hipMemcpyDeviceToHost, stream));
                HIPCHK(hipStreamSynchronize(stream));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
                stats.kernel_ms = ms;
            }
            if (got > cap) { cap = got; continue; }   // overflow: rerun with room for all
            u64* h = (u64*)std::malloc(std::max<u64>(1, got) * p.stride * 8);
            if (!h) throw HipError{3, "host allocation failed"};
            if (got) HIPCHK(hipMemcpy(h, out.p, got * p.stride * 8, hipMemcpyDeviceToHost));
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

bool is_permutation(const int* order, int n) {   // null counts as the identity
    if (!order) return true;
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; ++i) {
        if (order[i] < 0 || order[i] >= n || seen[order[i]]) return false;
        seen[order[i]] = 1;
    }
    return true;
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

int qsf_hip_enumerate_floor_noalloc(const qsf_enum_plan* plan, int floor, int* best_out) {
    std::string why;
    if (!best_out || !qsa::plan_ok(plan, why)) {
        engine().err = "qsf_hip_enumerate_floor_noalloc: " + (best_out ? why : std::string("null best_out"));
        return 1;
    }
    return run_guarded("qsf_hip_enumerate_floor_noalloc",
                       [&](qsa::Engine& e) { return e.enumerate_noalloc(*plan, floor, best_out); });
}

size_t qsf_hip_noalloc_scratch_bytes(void) { return kScratchBytes; }

uint64_t qsf_hip_solve_bytes(uint64_t key) {
    auto& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    for (const auto& s : e.cache)
        if (s->key == key && key != 0) return (uint64_t)s->arena.n * sizeof(qsa::u64);
    return 0;
}

int qsf_hip_mem_info(size_t* free_bytes, size_t* total_bytes) {
    return run_guarded("qsf_hip_mem_info", [&](qsa::Engine&) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) != hipSuccess) throw qsa::HipError{3, "hipMemGetInfo failed"};
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
        return 0;
    });
}

int qsf_hip_plan_variant(const qsf_enum_plan* plan, double factor, double at_p, int* out) {
    std::string why;
    if (!out || !qsa::plan_ok(plan, why)) {
        engine().err = "qsf_hip_plan_variant: " + (out ? why : std::string("null out"));
        return 1;
    }
    auto& e = engine();
    std::lock_guard<std::mutex> lk(e.mu);
    try {
        qsa::Solve s;
        e.build_host(s, *plan);   // host tables only: no device needed
        if (factor > 0.0 && at_p > 0.0) {
            s.sf = factor;
            s.sf_logp = std::log(at_p);
        }
        const qsa::Engine::Variant v = qsa::Engine::choose_variant(s, *plan);
        out[0] = v.fd; out[1] = v.xd; out[2] = v.rpl; out[3] = v.mm;
        return 0;
    } catch (const std::exception& x) {
        e.err = std::string("qsf_hip_plan_variant: ") + x.what();
        return 3;
    }
}

int qsf_hip_enumerate_batch(const qsf_enum_plan* plans, const int* floors, int n, int* best_out) {
    if (n < 0 || (n > 0 && (!plans || !best_out))) { engine().err = "qsf_hip_enumerate_batch: bad arguments"; return 1; }
    for (int j = 0; j < n; ++j) {
        std::string why;
        if (!qsa::plan_ok(plans + j, why)) {
            engine().err = "qsf_hip_enumerate_batch: plan " + std::to_string(j) + ": " + why;
            return 1;
        }
    }
    if (n == 0) return 0;
    return run_guarded("qsf_hip_enumerate_batch", [&](qsa::Engine& e) { return e.enumerate_batch(plans, floors, n, best_out); });
}

int qsf_hip_enumerate_levels(const qsf_enum_plan* plans, const int* floors, const int* groups, int n, int* best_out,
                             uint64_t* skipped_out) {
    if (n < 0 || (n > 0 && (!plans || !best_out || !groups))) {
        engine().err = "qsf_hip_enumerate_levels: bad arguments";
        return 1;
    }
    for (int j = 0; j < n; ++j) {
        std::string why;
        if (!qsa::plan_ok(plans + j, why)) {
            engine().err = "qsf_hip_enumerate_levels: plan " + std::to_string(j) + ": " + why;
            return 1;
        }
        if (groups[j] < 0 || groups[j] >= n) {
            engine().err = "qsf_hip_enumerate_levels: group index out of range";
            return 1;
        }
    }
    if (n == 0) return 0;
    return run_guarded("qsf_hip_enumerate_levels", [&](qsa::Engine& e) {
        return e.enumerate_batch(plans, floors, n, best_out, groups, skipped_out);
    });
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

int qsf_hip_eliminate(qsf_elim_job* jobs, int njobs, int trials, uint64_t seed, int what) {
    if (!jobs || njobs < 0 || trials < 0) { engine().err = "qsf_hip_eliminate: bad arguments"; return 1; }
    for (int j = 0; j < njobs; ++j) {
        const qsf_elim_job& q = jobs[j];
        const bool sets = (what & QSF_ELIM_SETS) != 0;
        if (q.K < 0 || q.n <= 0 || q.n > qsa::kElimMaxN || (q.K > 0 && !q.gen) || q.kcheck < 0 ||
            (q.kcheck > 0 && !q.check) || (sets && (q.set_cap < 1 || !q.ranks)) ||
            (q.nparts > 0 && (!q.part_off || !q.part_cols)) || !qsa::is_permutation(q.order, q.n)) {
            engine().err = "qsf_hip_eliminate: bad job " + std::to_string(j) + " (n must be in [1, 4096])";
            return 1;
        }
    }
    if (njobs == 0) return 0;
    return run_guarded("qsf_hip_eliminate", [&](qsa::Engine& e) { return e.eliminate(jobs, njobs, trials, seed, what); });
}

int qsf_hip_is_bound(const uint64_t* gen, int K, int n, const uint64_t* check, int kcheck, int every_nonzero,
                     int trials, uint64_t seed, int init, int* best_out) {
    if (!best_out) { engine().err = "qsf_hip_is_bound: null best_out"; return 1; }
    qsf_elim_job q{};
    q.gen = gen; q.K = K; q.n = n; q.check = check; q.kcheck = kcheck; q.every_nonzero = every_nonzero;
    q.init = init;
    const int rc = qsf_hip_eliminate(&q, 1, trials, seed, QSF_ELIM_BOUND);
    if (rc == 0) *best_out = q.best;
    return rc;
}

void qsf_hip_last_elim_stats(qsf_hip_elim_stats* out) {
    if (out) *out = engine().el_stats;
}

int qsf_hip_collect_upto(const qsf_enum_plan* plan, int keep_weight, uint64_t** hits, int64_t* rows) {
    std::string why;
    if (!hits || !rows || !qsa::plan_ok(plan, why)) {
        engine().err = "qsf_hip_collect_upto: " + (hits && rows ? why : std::string("null output"));
        return 1;
    }
    return run_guarded("qsf_hip_collect_upto", [&](qsa::Engine& e) { return e.collect_upto(*plan, keep_weight, hits, rows); });
}

void qsf_hip_last_collect_stats(uint64_t* appended, double* sort_ms) {
    auto& e = engine();
    if (appended) *appended = e.appended_last;
    if (sort_ms) *sort_ms = e.sort_ms;
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
