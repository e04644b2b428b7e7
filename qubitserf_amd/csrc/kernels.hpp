// Device-side definitions shared by the enumeration kernels and their host launcher.
//
// The engine answers one question per weight level d (the reference's Backend::enumerate,
// cpp/qsf/backend.hpp:57-58): the lightest logical codeword that is the XOR of d rows of
// one of the active generators. Two kernels do it.
//
//  * bz_tiled<FD, RPL>  — the hot kernel. The K rows are cut into A = [0, h) and
//    B = [h, K). A d-subset is an a-subset of A joined with a (d-a)-subset of B, so level d
//    is the union over a of the complete products  subsets(A, a) x subsets(B, d-a).
//    Each product is tiled like a GEMM: 64 lanes x RPL register-resident codewords of the
//    larger list, times a chunk of S codewords of the smaller list streamed from LDS with
//    a broadcast read. One candidate costs FD x (v_xor_b32 + v_bcnt_u32_b32) plus half a
//    v_min3: no loads, no divergence, no per-candidate unranking.
//  * bz_rowsweep — the plain kernel: every thread owns a run of lexicographic combinations
//    and XORs full rows (the reference's two-level sweep, enum_core.hpp:109-194). Used for
//    small levels and K > 256.
//
// Row masks. A list's subsets are bit masks over its row range: one 64-bit word for ranges
// of up to 64 rows (K <= 128: both halves fit), two words (W = 2) for up to 128 rows, which
// the VALU kernel uses for 128 < K <= 256. Fixed row sets and the flagged rows of a
// generator are 4-word masks over rows 0..255.
//
// Codeword representation in the tiled kernel ("residual form"). For generator g, a
// column that holds a single 1, at row i, is a unit column of row i; one unit column per
// row is set aside (row i is then "flagged"). A codeword's weight on the set-aside columns
// is exactly the number of flagged rows XORed into it, so
//     weight = popcount(XOR of residual parts) + #flagged rows chosen,
// where the residual part of a row is the row without the set-aside columns, packed into
// RD dwords. For a full-rank systematic generator every row is flagged and the residual
// is the n-K redundancy bits. Logical test: the detector rows enter only through the
// per-row syndrome (bit c = <check_c, row>), XORed like the residual (SD dwords).
#pragma once
#include <cstdint>

namespace qsa {

using u64 = std::uint64_t;
using u32 = std::uint32_t;

constexpr int kMaxTiledK = 256;     // VALU kernel: two-word masks over each half when K > 128
constexpr int kMaxMatrixK = 128;    // matrix-core kernels: one-word masks over each half
constexpr int kMaxD = 64;           // binomial table columns
constexpr int kStreamMax = 512;     // LDS stream chunk capacity (elements)
constexpr int kMfmaBlocks = 8;      // bz_mfma: 32-element lane blocks per wave
constexpr int kMfmaWaves = 4;       // bz_mfma: waves per workgroup, sharing each stream chunk
constexpr int kMfmaLane = 32 * kMfmaBlocks * kMfmaWaves;   // bz_mfma: lane elements per tile
constexpr int kMfmaStream = 512;    // bz_mfma: stream chunk length (two elements per thread)
// bz_f4<XD, MM>: 32 x 32 products (MM = 2), 16 x 16 products (MM = 3), two chained
// 32 x 32 products (MM = 4), or 32 x 32 products over two lane elements per row (MM = 5)
// per lane block
constexpr int f4_rows(int mm) { return mm == 3 ? 16 : 32; }
constexpr int f4_blocks(int mm) { return mm == 4 ? 4 : kMfmaBlocks; }                  // lane blocks per wave
constexpr int f4_per_row(int mm) { return mm == 5 ? 2 : 1; }                           // lane elements per product row
constexpr int f4_lane_tile(int mm) { return f4_rows(mm) * f4_blocks(mm) * f4_per_row(mm) * kMfmaWaves; }   // lane elements per tile
constexpr int f4_stream(int mm) { return mm == 2 || mm == 5 ? 512 : 256; }             // stream chunk length
constexpr int kPackMaxT = 30;       // bz_f4<xd, 5>: the largest bound the packed test runs at
constexpr int kMaxRW = 64;          // residual + syndrome dwords per row in the tiled kernel
constexpr int kLdsFullRows = 1024;  // dwords: the whole K x RW table is kept in LDS up to this size
constexpr int kBigOffset = (1 << 30) + (1 << 29);  // masks lanes past the end of a list

// One complete product  F + subsets(L, lk) x subsets(Sr, sk)  of one generator: L and Sr
// are disjoint row ranges of at most 64 rows (128 in the two-word kernels), F a fixed set
// of rows outside both (its codeword is XORed into every stream element). The register
// tile holds subsets of L, the stream subsets of Sr.
struct Rect {
    int job;               // index into TiledArgs::jobs
    int g;
    int lbase, lm, lk;     // lane list: lk-subsets of rows [lbase, lbase + lm)
    int sbase, sm, sk;     // stream list: sk-subsets of rows [sbase, sbase + sm)
    int fixed;             // index into TiledArgs::fixed, or -1 for no fixed rows
    // nfix > 1 (VALU kernel, min mode): a list of nfix fixed row sets, entries fixed ..
    // fixed + nfix - 1; the stream is then every (fixed set, sk-subset) pair, fixed set
    // major: element e = f C(sm, sk) + s. fmask holds the rows all of them share.
    int nfix;
    u64 fmask[4];          // the fixed rows as a 256-bit row mask
    u64 n_lane, n_stream;  // list lengths C(lm, lk), C(sm, sk) (x nfix on the stream side)
    u64 lane_blocks;       // ceil(n_lane / (64 * RPL))
    u64 stream_chunks;     // ceil(n_stream / S)
    u64 stream_groups;     // ceil(stream_chunks / G): a tile is one lane block x G chunks
    u64 tile_base;         // index of this rectangle's first tile
};

// One problem's level inside a tiled launch (a launch may hold many problems' levels).
struct TJob {
    const u32* tbl;          // [g][row][RW]: RD residual dwords, then SD syndrome dwords
    const u64* flags;        // [g][4]: flagged rows 0..255 as a 256-bit mask
    const u32* fixed;        // [f][RW + 1]: residual+syndrome of a fixed row set, then its flagged count
    const u64* gamma;        // [g][row][stride] full rows (collect mode)
    int* best;               // running bound; starts at current_best
    // The bound shared by the levels of one problem run in the same launch (== best when
    // the job is alone): a tile is skipped once it is <= floor, and counted in `skipped`.
    int* shared;
    unsigned long long* skipped;
    unsigned long long* hits;   // hot-test passes (medium-test entries) on this job, for the tuner
    int K, RD, SD, RW, stride;
    int floor;               // stop once best <= floor (floor < 0: never)
    // A level queued ahead of the bound it will meet (several levels of a problem in one
    // call) may carry candidate kernel variants for its full products: one job entry and
    // launch per candidate over the same products, and the first wave that reaches the
    // level picks the candidate for the bound it meets there, once, into *choice (-1: not
    // yet); the other candidates' tiles are passed over. choice == nullptr: no choice.
    // Candidate k serves bounds <= seg_max[k] (ascending); the last one any bound.
    int* choice;
    int variant;             // this job's candidate (the launch's kernel variant)
    int nseg;
    int seg_max[4];
    int seg_var[4];
};

// The candidate of a multi-variant level for the bound the level meets: picked by the
// first caller, the same for every later one.
__device__ inline int pick_variant(const TJob* J, int bound) {
    int c = __hip_atomic_load(J->choice, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c >= 0) return c;
    int want = J->seg_var[J->nseg - 1];
    for (int k = 0; k + 1 < J->nseg; ++k)
        if (bound <= J->seg_max[k]) { want = J->seg_var[k]; break; }
    const int old = atomicCAS(J->choice, -1, want);
    return old < 0 ? want : old;
}

struct TiledArgs {
    const TJob* jobs;
    int njobs;               // levels in the launch; with one, the launch ends once it is settled
    const Rect* rects;       // sorted by tile_base
    int nrects;
    u64 total_tiles;
    // Tile -> rectangle lookup: tile_base[r] = rects[r].tile_base (a compact copy), and
    // tile_ix[q] = the rectangle holding tile q << ix_shift (n_ix entries), so the
    // rectangle of tile t lies in [tile_ix[t >> ix_shift], tile_ix[(t >> ix_shift) + 1]].
    const u64* tile_base;
    const unsigned* tile_ix;
    unsigned n_ix;
    int ix_shift;
    const u64* binom;        // C(n, k) at binom[n * bk + k], n <= kMaxTiledK, k < bk, saturating
    int bk;
    int S;                   // stream chunk length (<= kStreamMax)
    int G;                   // stream chunks per tile (the register tile is built once per tile)
    int lds_full;            // dwords of dynamic LDS for a job's whole table (0: none)
    unsigned long long* next_tile;
    unsigned long long* slow_hits;   // [0] slow-path entries, [1] medium-test entries (diagnostics)
    // collect mode (one job): every codeword of weight in [1, keep] is written out in full
    int collect, keep;
    unsigned long long* out_rows;
    u64* out;
    u64 out_cap;             // rows
};

// The rectangle holding tile t (see TiledArgs::tile_ix).
__device__ inline int find_rect(const TiledArgs& a, u64 t) {
    const u64 q = t >> a.ix_shift;
    int lo = (int)a.tile_ix[q];
    int hi = q + 1 < a.n_ix ? (int)a.tile_ix[q + 1] : a.nrects - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.tile_base[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}


struct SweepArgs {
    const u64* gamma;        // [g][row][stride] full rows
    const u64* check;        // [c][stride]
    int kcheck, stride, K, d, ngamma;
    const u64* binom;
    int bk;
    u64 total;               // C(K, d)
    u64 chunk;               // combinations per thread
    int* best;               // min mode
    int collect, keep;       // collect mode: weight in [1, keep]
    unsigned long long* out_rows;
    u64* out;
    u64 out_cap;             // rows
};

// bz_sweep_small: many small levels (min mode) in one launch. Workgroup b serves one
// (level, generator) pair: the pair's job is the last one with block_base <= b, the
// generator (b - block_base) / blocks_per_g.
struct SJob {
    const u64* gamma;        // [g][row][stride] full rows
    const u64* check;        // [c][stride]
    const u64* binom;        // the solve's C(n, k) at binom[n * bk + k]
    int* best;               // running bound; starts at current_best
    u64 total;               // C(K, d)
    u64 block_base;          // first workgroup of this job
    int K, d, stride, kcheck, ngamma, bk;
    int blocks_per_g;        // workgroups per active generator
    int chunk;               // lexicographic combinations per thread
};

struct SweepJobsArgs {
    const SJob* jobs;        // sorted by block_base
    int njobs;
};

constexpr int kSweepRowsLds = 2048;   // u64: a generator's rows are staged in LDS up to this size
constexpr int kSweepBinLds = 1024;    // u64: binomials C(n < K, k < d) staged in LDS up to this count

}  // namespace qsa
