/* MI355X (gfx950) Brouwer–Zimmermann enumeration engine — plugin-level C ABI.
 *
 * These entry points replace the reference's GPU plugin, `qsf::Backend` as implemented by
 * `CudaBackend` (cpp/qsf/backend.hpp:53-61; cpp/qsf/cuda/backend_cuda.cu:506-645). A
 * host that already builds EnumPlans (the reference's bz.cpp / codeaut/bz.cpp drivers)
 * calls them through a ~40-line `HipBackend : qsf::Backend` shim that forwards each virtual
 * to the entry point of the same name.
 *
 * Plain pointers and sizes only. Every call is synchronous with respect to the caller:
 * it returns after the device work has finished. Calls from several host threads are
 * serialised inside the library. Nothing here falls back to the CPU: when the device is
 * missing or a HIP call fails the entry point returns a nonzero code and
 * qsf_hip_last_error() says why.
 */
#ifndef QSF_HIP_H
#define QSF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSF_WEIGHT_NONE (1 << 30) /* "no codeword" sentinel, cpp/qsf/backend.hpp:19 */

/* One weight level of work. Field for field the layout of qsf::EnumPlan
 * (cpp/qsf/backend.hpp:21-42), so a reference plan can be passed by address.
 *   gamma : num_gamma_total x K x stride u64, row-major (generator, row, word); bit j of a
 *           row is bit (j & 63) of word (j >> 6). Only the first num_gamma generators are
 *           enumerated at this level.
 *   check : kcheck x stride u64 logical-detector rows; kcheck == 0 => every codeword counts.
 *   binom : accepted for layout compatibility and ignored (the engine keeps its own table).
 *   buffers_key : unique per solve; device copies of gamma/check are cached under it and
 *           reused across the levels of that solve (never keyed on host pointers). */
typedef struct {
    int n;
    int stride;
    int K;
    int d;
    int num_gamma;
    int num_gamma_total;
    const uint64_t* gamma;
    int kcheck;
    const uint64_t* check;
    const uint64_t* binom;
    int binom_maxN;
    int binom_maxK;
    int current_best;
    uint64_t buffers_key;
} qsf_enum_plan;

/* 1 when a gfx950 device is present and the kernels are loadable (replaces
 * CudaBackend::available, backend_cuda.cu:485-503). */
int qsf_hip_available(void);

/* Min mode (replaces CudaBackend::enumerate, backend_cuda.cu:506-557): *best_out =
 * min(plan->current_best, lightest logical codeword that is the XOR of d rows of one of the
 * first num_gamma generators). Exact; any value >= current_best means "none lighter".
 * Returns 0 on success; 1 bad arguments; 2 no device; 3 HIP error; 4 level too large
 * (more than 2^62 codewords). */
int qsf_hip_enumerate(const qsf_enum_plan* plan, int* best_out);

/* Like qsf_hip_enumerate, but the level may stop as soon as a logical of weight <= floor
 * has been found: the caller knows no codeword of this solve is lighter than `floor`
 * (the BZ lower bound after level d-1), so such a codeword is the answer. floor < 0
 * disables it. The result is the same number qsf_hip_enumerate would return whenever that
 * number is <= floor or no codeword <= floor exists. */
int qsf_hip_enumerate_floor(const qsf_enum_plan* plan, int floor, int* best_out);

/* qsf_hip_enumerate_floor without allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, result and return codes as
 * qsf_hip_enumerate_floor, and the same kernels, but every byte of device memory it uses
 * comes from one static device array of qsf_hip_noalloc_scratch_bytes() bytes (16 MiB) that
 * the library holds on its GPU until the process ends: the level's bound and counters, the
 * solve's tables (uploaded once per buffers_key, from pageable host memory) and the run's
 * block. It calls no hipMalloc, hipFree or hipHostMalloc; host-side tables use the C++ heap.
 * A level whose tables and run block do not fit returns 4 with the reason in
 * qsf_hip_last_error(). The array keeps the last call's contents; every call uploads what
 * must start at zero. Calls are serialised by the library's lock, so two calls that overlap
 * from different host threads run one after the other on the array (two copies of the
 * library in one process would each have their own). */
int qsf_hip_enumerate_floor_noalloc(const qsf_enum_plan* plan, int floor, int* best_out);
size_t qsf_hip_noalloc_scratch_bytes(void);

/* Device memory accounting (no reference counterpart; for admission control):
 * qsf_hip_solve_bytes: bytes of the device arena holding the solve cached under `key`
 * (its generators, detector rows, binomials, flags and residual table), 0 if none is cached;
 * qsf_hip_mem_info: the device's free and total memory (hipMemGetInfo). */
uint64_t qsf_hip_solve_bytes(uint64_t key);
int qsf_hip_mem_info(size_t* free_bytes, size_t* total_bytes);

/* Batched min mode (no reference counterpart: the reference runs one level of one problem
 * per call). Plan j is one level of its own problem (distinct buffers_key per problem);
 * best_out[j] is what qsf_hip_enumerate_floor(&plans[j], floors[j], ...) would return
 * (floors may be NULL: no early stop). Every tileable level's tiles share one persistent
 * launch per kernel variant; qsf_hip_last_stats then sums codewords and device time. */
int qsf_hip_enumerate_batch(const qsf_enum_plan* plans, const int* floors, int n, int* best_out);

/* Several levels of each of many problems in one batched run (no reference counterpart;
 * the batched solver's lookahead rounds). Plans with equal groups[j] (0 <= groups[j] < n) are
 * levels of one problem (same buffers_key): each enumerates on its own bound, starting at
 * its current_best, so best_out[j] is what qsf_hip_enumerate_floor(&plans[j], floors[j], ...)
 * returns; the group also keeps a shared bound (the minimum found by any of its plans), and
 * a tile of plan j is skipped once that shared bound is <= floors[j] (the level cannot change
 * the outcome: an earlier level of the same launch settled the problem). skipped_out[j]
 * (may be NULL) counts plan j's tiles skipped that way; best_out[j] is exact when it is 0. */
int qsf_hip_enumerate_levels(const qsf_enum_plan* plans, const int* floors, const int* groups, int n, int* best_out,
                             uint64_t* skipped_out);

/* Collect mode (replaces CudaBackend::collect, backend_cuda.cu:561-645): every codeword of
 * Hamming weight in [1, keep_weight] that is the XOR of d rows of one of the first
 * num_gamma generators, `stride` u64 words each, duplicates allowed. On success *hits is a
 * buffer owned by the library (free with qsf_hip_free) holding *rows codewords. Complete
 * or an error: device-buffer overflow is retried with a larger buffer, never truncated. */
int qsf_hip_collect(const qsf_enum_plan* plan, int keep_weight, uint64_t** hits, int64_t* rows);
void qsf_hip_free(void* p);

/* Collect mode over levels 1..plan->d at once (no reference counterpart; what
 * cpp/codeaut/bz.cpp:78-113 does with one Backend::collect per level, a host sort and a
 * dedup): every codeword of Hamming weight in [1, keep_weight] that is the XOR of at most
 * plan->d rows of one of the first num_gamma generators, each once, sorted
 * lexicographically by u64 word (word 0 first, unsigned; the order of the reference's
 * RowLess, codeaut/bz.cpp:41-51). The levels append to one device buffer (one atomic per
 * wave and append); the rows are sorted and deduplicated on the device and copied back once.
 * *hits (free with qsf_hip_free) holds *rows rows of `stride` words. qsf_hip_last_stats:
 * codewords of all levels and their enumeration time; qsf_hip_last_collect_stats: rows
 * appended before deduplication and the device sort time. Returns as qsf_hip_collect. */
int qsf_hip_collect_upto(const qsf_enum_plan* plan, int keep_weight, uint64_t** hits, int64_t* rows);
void qsf_hip_last_collect_stats(uint64_t* appended, double* sort_ms);

/* Elimination phase on the device, many problems per call (one wave per problem, or per
 * problem and trial). Each job describes one BZ problem: `gen` K rows of stride =
 * (n+63)/64 u64 words spanning the code searched, `check` kcheck detector rows (a row is a
 * logical when it has odd overlap with one of them) or every_nonzero = 1.
 *
 * QSF_ELIM_SETS: the information sets in systematic form (replaces gamma_sequence_greedy,
 *   cpp/distfind/css.cpp:177-203, over restricted_rref, gf2.cpp:42-58). Without a
 *   partition: the greedy peel — eliminate over the live columns in order (pivot = first
 *   row at/below the rank with a 1), the pivots form a set and the matrix its generator,
 *   continue from that matrix over the live columns that were neither pivots nor zero
 *   columns, stop when a round removes nothing; the live columns start as 0..n-1, or as
 *   `order` when given (a proposal's column order). With a partition (nparts > 0; part p is
 *   part_cols[part_off[p] .. part_off[p+1])): set p is gen made systematic over part p
 *   (css.cpp:358-361). Outputs: nsets, and the first min(nsets, set_cap) generators
 *   (set_cap x K x stride u64, in `sets`; sets may be NULL for ranks only) and ranks (in
 *   `ranks`), caller-owned.
 * QSF_ELIM_BOUND: the random information-set upper bound (replaces random_is_seed,
 *   cpp/distfind/bz.cpp:24-70, run before level 1 at bz.cpp:128 and bz.cpp:252): best =
 *   min(init, lightest nonzero logical row over `trials` systematic forms). Trial t orders
 *   the columns by Fisher–Yates over the splitmix64 stream seeded with
 *   seed ^ (0xd1b54a32d192ed03 * (t + 1)), swapping position i with floor(hi32(x) (i+1) / 2^32),
 *   so the result is a pure function of the arguments.
 * n <= 4096. Return codes as qsf_hip_enumerate. */
#define QSF_ELIM_SETS 1
#define QSF_ELIM_BOUND 2
typedef struct {
    const uint64_t* gen;
    int K;
    int n;
    const uint64_t* check;
    int kcheck;
    int every_nonzero;
    const int* part_cols;
    const int* part_off;
    int nparts;
    int init;
    int best;            /* out */
    int nsets;           /* out */
    uint64_t* sets;      /* out, set_cap x K x stride */
    int set_cap;
    int* ranks;          /* out, set_cap */
    const int* order;    /* optional: the peel's initial column order, a permutation of [0, n) */
    int* set_of_col;     /* optional out (QSF_ELIM_SETS, n ints): index of the set (in peel or
                            part order) whose pivot column c became, -1 for none */
} qsf_elim_job;
int qsf_hip_eliminate(qsf_elim_job* jobs, int njobs, int trials, uint64_t seed, int what);

/* One job, QSF_ELIM_BOUND only. */
int qsf_hip_is_bound(const uint64_t* gen, int K, int n, const uint64_t* check, int kcheck, int every_nonzero,
                     int trials, uint64_t seed, int init, int* best_out);

/* Last qsf_hip_eliminate call: device milliseconds of the information-set kernel and of
 * the random-IS kernels, and the information-set kernel's algorithmic HBM bytes
 * (generators read once + systematic generators written). */
typedef struct {
    double sets_ms;
    double bound_ms;
    uint64_t sets_bytes;
    int jobs;
    int trials;
} qsf_hip_elim_stats;
void qsf_hip_last_elim_stats(qsf_hip_elim_stats* out);

/* Meet-in-the-middle test of one weight (replaces has_codeword_of_weight,
 * cpp/distfind/mitm.cpp:241-314): *hit = 1 iff some word x of Hamming weight d has a zero
 * syndrome and, unless count_all, a nonzero detector fingerprint. syn[j] holds column j of
 * the code's parity checks (bit r = row r; at most 63 rows), log[j] column j of the detector
 * rows (at most 64). Columns [0, nl) form the left half and [nl, n) the right, at most 64
 * each. Returns 0; 1 bad arguments; 3 HIP error; 4 a weight split too large for the device
 * table (more than 2^27 parts on the smaller side). */
int qsf_hip_mitm_weight(const uint64_t* syn, const uint64_t* log, int n, int nl, int d, int count_all, int* hit);

/* The symplectic form (replaces has_codeword_of_weight_sym, cpp/distfind/mitm.cpp:131-237):
 * *hit = 1 iff some Pauli operator touching exactly d qubits has a zero syndrome and a
 * nonzero detector fingerprint. syn / log hold 2q columns in [z | x] order as above (a
 * qubit's Z reads column j, X column q + j, Y both). Qubits [0, ql) form the left half and
 * [ql, q) the right, at most 64 each; a side of weight k has C(m, k) 3^k parts. Returns as
 * qsf_hip_mitm_weight. */
int qsf_hip_mitm_weight_sym(const uint64_t* syn, const uint64_t* log, int q, int ql, int d, int* hit);
const char* qsf_hip_mitm_last_error(void);

/* Statistics of the last enumerate/collect call on this thread's library instance:
 * codewords examined (C(K,d) x num_gamma, the reference's `work`, bz.cpp:151-154), device
 * time in milliseconds of the enumeration kernel, and which kernel ran
 * (0 = none, 1 = row-sweep kernel, 2 = tiled split kernel, 3 = small-level row sweep:
 * levels under 2^16 codewords, many per launch). */
typedef struct {
    uint64_t codewords;
    double kernel_ms;
    int kernel;
    int tiles;
    int fd;        /* tiled kernel: residual dwords in the first (fast) test */
    uint64_t slow_hits;   /* tiled kernel: stream steps that entered the exact slow path */
    int xd;               /* tiled kernel: residual dwords held per codeword (medium test) */
    uint64_t medium_hits; /* tiled kernel: stream steps that passed the hot-loop test */
    int rpl;              /* tiled kernel: codewords held per lane */
    double ops_per_codeword; /* tiled kernel: hot-loop int32 VALU lane-ops per codeword,
                                (2 FD RPL + RPL/2 + 3) / RPL: FD xor+bcnt pairs per codeword,
                                a v_min3 per two codewords, 3 per stream step */
    int mfma;             /* tiled kernel: matrix-core hot test of the last launch: 0 none (VALU),
                             1 int8 32x32x32 products, 2 fp4 32x32x64, 3 fp4 16x16x128,
                             4 two chained fp4 32x32x64, 5 packed fp4 32x32x64 (two lane
                             elements per row, bounds <= 30) */
    double side_ms;       /* tiled kernel: of kernel_ms, the VALU launches that took a matrix-test
                             level's thin products (a list too short for the matrix tiles) */
    uint64_t side_codewords; /* tiled kernel: of codewords, those thin products' */
    double side_ops_per_codeword; /* tiled kernel: hot-loop lane-ops per codeword of those launches */
} qsf_hip_stats;
void qsf_hip_last_stats(qsf_hip_stats* out);

/* The hot-test variant the automatic choice would run a level on (diagnostics and tests;
 * host work only, no device needed): out[0..3] = fd, xd, rpl, mm as in qsf_hip_stats. The
 * choice prices every applicable variant with a cost model whose constants were measured
 * on an MI355X (qubitserf_amd/tune.py -> csrc/tuning.hpp) and whose hot-test pass rates
 * are fair-coin rates times the solve's structure factor; factor > 0 with at_p > 0 sets
 * that factor as if a level with fair per-codeword pass probability at_p had measured it
 * (otherwise the prior, 1). Returns 0, or 1 for bad arguments. */
int qsf_hip_plan_variant(const qsf_enum_plan* plan, double factor, double at_p, int* out);

/* Kernel choice for later calls (process-wide): 0 automatic (default), 1 always the
 * row-sweep kernel, 2 the tiled kernels wherever they apply (K <= 256, d <= 32). For tests
 * and tuning; results never depend on it. Returns the previous policy. */
int qsf_hip_set_kernel_policy(int policy);

/* Human-readable reason for the last nonzero return code. */
const char* qsf_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QSF_HIP_H */
