/* distfind C ABI of the gfx950 engine — the symbols the reference's ctypes layer binds
 * (python/qubitserf/distfind/_native.py:39-107), with the reference's signatures and
 * struct layouts (cpp/distfind/capi.h:12-111).
 *
 * What this engine answers: methods "bz" and "mitm" (classical, CSS and non-CSS codes)
 * with backend "gpu" or "auto", on the device. Return codes: 0 ok; 1 null pointer
 * (capi.cpp:35, :71); 2 odd column count for a symplectic matrix (capi.cpp:171); 3 "cc" on
 * a non-CSS code (capi.cpp:214); and, new here, 4 method not provided by this engine
 * ("cc", "mitm" beyond 63 parity checks / 64 detectors / n = 128 (128 qubits for a
 * non-CSS code); the operator-weight and subsystem entry points), 5 backend not provided ("cpu", or no gfx950 device), 6 device
 * error (distfind_last_error() has the text). There is no silent
 * CPU fallback: `DistfindResult.backend` is always "hip" on success.
 */
#ifndef QSA_DISTFIND_CAPI_H
#define QSA_DISTFIND_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cpp/distfind/capi.h:12-19 */
typedef struct {
    int distance;     /* -1 when undefined (no logicals) */
    int lower_bound;
    int proven;
    int levels;
    double seconds;
    char backend[16];
} DistfindResult;

/* cpp/distfind/capi.h:38-45 */
typedef struct {
    int z_weight;
    int x_weight;
    int proven;
    double seconds;
    char backend[16];
} DistfindOpResult;

/* which: 'Z', 'X' or 'M' (= min over both, interleaved). cpp/distfind/capi.h:24-29 */
int distfind_css_distance(const uint8_t* Hx, int hx_rows, int hx_cols,
                          const uint8_t* Hz, int hz_rows, int hz_cols,
                          const char* method, char which, const char* backend,
                          int threads, int verbose, DistfindResult* out);

/* cpp/distfind/capi.h:32-36 */
int distfind_classical_distance(const uint8_t* H, int rows, int cols,
                                const char* method, const char* backend,
                                int threads, int verbose, DistfindResult* out);

/* S is rows x 2q in [z|x] order; CSS input is split and routed to the CSS solver, a
 * non-CSS code runs BZ on the weight-doubling isometric image (cpp/distfind/capi.h:81-85). */
int distfind_stabilizer_distance(const uint8_t* S, int rows, int cols,
                                 const char* method, char which, const char* backend,
                                 int threads, int verbose, DistfindResult* out);

/* Exported so the reference's binding loads; these return 4 (cpp/distfind/capi.h:52-108). */
int distfind_operator_weight(const uint8_t* Gx, int gx_rows, int gx_cols,
                             const uint8_t* Gz, int gz_rows, int gz_cols,
                             const uint8_t* z_op, const uint8_t* x_op, int n,
                             const char* method, const char* backend,
                             int threads, int verbose, DistfindOpResult* out);
int distfind_subsystem_distance(const uint8_t* Gx, int gx_rows, int gx_cols,
                                const uint8_t* Gz, int gz_rows, int gz_cols,
                                const char* method, char which, const char* backend,
                                int threads, int verbose, DistfindResult* out);
int distfind_subsystem_stabilizer_distance(const uint8_t* G, int rows, int cols,
                                           const char* method, char which, const char* backend,
                                           int threads, int verbose, DistfindResult* out);
int distfind_stabilizer_operator_weight(const uint8_t* G, int rows, int cols,
                                        const uint8_t* op, int op_len,
                                        const char* method, const char* backend,
                                        int threads, int verbose, DistfindResult* out);

/* 1 if usable: "gpu" / "auto" need a gfx950 device; "cpu" is 0 (cpp/distfind/capi.cpp:290-299). */
int distfind_backend_available(const char* name);
const char* distfind_version(void);

/* Text for the last nonzero return code of any distfind_* call (new; no reference counterpart). */
const char* distfind_last_error(void);

/* One enumerated weight level of the last distfind_* solve on the calling thread (new; the
 * reference prints the same fields under DISTFIND_PROFILE, cpp/distfind/bz.cpp:147-155). */
typedef struct {
    int32_t side;          /* 'S' single subproblem, 'Z' / 'X' interleaved CSS sides */
    int32_t d;             /* weight level */
    int32_t active;        /* information sets enumerated at this level */
    int32_t ngamma;        /* information sets in total */
    int32_t K;             /* rows per generator */
    int32_t inner, outer;  /* upper and lower bound after the level */
    int32_t kernel;        /* 0 none, 1 row sweep, 2 tiled, 3 small-level row sweep */
    int32_t fd, xd, rpl;   /* tiled kernel variant (hot-test dwords, held dwords, codewords/lane) */
    int32_t mfma;          /* matrix-core hot test: 0 none, 1 int8 32x32x32, 2 fp4 32x32x64, 3 fp4 16x16x128,
                              4 two chained fp4 32x32x64, 5 packed fp4 32x32x64 */
    uint64_t codewords;    /* C(K, d) x active */
    uint64_t medium_hits;  /* stream steps past the hot-loop test */
    uint64_t slow_hits;    /* stream steps into the exact slow path */
    double ms;             /* wall time of the level */
    double kernel_ms;      /* device time of the enumeration kernel (HIP events) */
    double ops_per_codeword; /* hot-loop int32 VALU lane-ops per codeword of the variant */
    double side_ms;        /* of kernel_ms: the VALU launches beside a matrix test (its thin products) */
    uint64_t side_codewords; /* of codewords: those thin products' */
    double side_ops_per_codeword; /* hot-loop int32 VALU lane-ops per codeword of those launches */
} DistfindLevel;

/* Copies up to `cap` levels into `out` (may be NULL); returns the number recorded. */
int distfind_last_trace(DistfindLevel* out, int cap);

/* ---- batched solves (new; no reference counterpart) ---------------------------------
 * Many codes in one call: every problem is stepped level by level exactly as the single
 * entry points step it (same distance, lower_bound, proven, levels), and each round
 * enumerates one level of every unfinished problem in one batched device call
 * (qsf_hip_enumerate_batch); the information sets and upper bounds of all problems come
 * from three batched device eliminations. */
typedef struct {
    int kind;              /* 0 classical: A = H; 1 CSS: A = Hx, B = Hz */
    const uint8_t* A;
    int ar, ac;
    const uint8_t* B;
    int br, bc;
    char which;            /* CSS: 'M' (min of both sides, interleaved), 'Z' or 'X' */
    const int* order;      /* optional, single-side problems: column order (permutation of
                              [0, n)) for the greedy information-set peel; NULL = 0..n-1 */
    const float* scores;   /* optional, single-side problems: a proposal's score per column */
    int candidates;        /* with scores: column orders sampled from the scores (descending
                              score + Gumbel noise); the peel with the least predicted BZ work
                              replaces the natural one when strictly better */
    uint64_t proposal_seed;
    int64_t work_cap;      /* > 0: this problem's own work cap (else the call's work_cap) */
} DistfindBatchItem;

typedef struct {
    int rounds;            /* batched enumeration calls */
    int capped;            /* problems stopped by work_cap (proven = 0) */
    uint64_t codewords;    /* enumerated in total */
    double device_ms;      /* enumeration kernels */
    double setup_s;        /* problem build, information sets and upper bounds */
    double seconds;        /* whole call */
} DistfindBatchStats;

/* work_cap > 0: a problem stops before a level would take its enumerated codewords past
 * work_cap; it then reports its bounds so far with proven = 0. is_trials: random
 * information-set trials for each problem's upper bound (<= 0: 256, the reference's count,
 * bz.cpp:128). Return codes as distfind_css_distance; out has `count` entries. */
int distfind_batch_distance(const DistfindBatchItem* items, int count, int64_t work_cap, int is_trials, int verbose,
                            DistfindResult* out);
void distfind_last_batch_stats(DistfindBatchStats* out);

/* Continuous batching (a session): problems join between rounds and leave when done.
 * distfind_session_new returns NULL when the engine is unusable (distfind_last_error).
 * add: builds the items' problems and their information sets and bounds on the device,
 * writes one id per item. step: one round — one level of every unfinished problem in one
 * batched device call; *active = problems still unfinished. poll: up to cap finished
 * problems (id, result; result.seconds = time since its add), returns the count. Results
 * equal the single-call results. A session is not thread-safe. */
void* distfind_session_new(int64_t work_cap, int is_trials);
int distfind_session_add(void* session, const DistfindBatchItem* items, int count, int64_t* ids);
int distfind_session_step(void* session, int* active);
int distfind_session_poll(void* session, int64_t* ids, DistfindResult* out, int cap);
int distfind_session_active(void* session);

/* Device bytes the session's unfinished problems hold now (the arenas of their cached
 * solves, measured with qsf_hip_solve_bytes); -1 for a NULL session. */
int64_t distfind_session_device_bytes(void* session);
void distfind_session_stats(void* session, DistfindBatchStats* out);
void distfind_session_free(void* session);

/* Level schedule of later solves (process-wide; no reference counterpart). 0 (default) =
 * the reference's: a generator of rank r < K is enumerated only at levels d with
 * (d+1) - (K-r) > 0 (cpp/distfind/bz.cpp:142-144) but counted in the lower bound at every
 * level (bz.cpp:159-161), which can certify a distance above the true one when a lighter
 * codeword lies in that generator's skipped levels 1..K-r-1. 1 = sound: the generators that
 * contribute before the bound can reach the upper bound are enumerated from level 1 and
 * only they are counted. The initial setting is QSA_SOUND_BOUNDS from the environment.
 * Returns the previous setting. Solves already in flight (sessions) keep theirs. */
int distfind_set_sound_bounds(int on);

/* Host logic exposed for tests (no device needed): the Edmonds matroid partition of the
 * columns of G (rows x cols uint8) into independent classes (cpp/distfind/css.cpp:274-328),
 * classes ordered by size as the reference sorts them (css.cpp:354-355), empty classes
 * dropped; part_of_col[c] = class of column c. Return 0, 1 bad arguments, 6 failed check. */
int distfind_matroid_parts(const uint8_t* G, int rows, int cols, int* part_of_col);

/* ---- text formats of the distfind tool (cpp/distfind/cli/distfind_cli.cpp:100-188) ----
 * 0/1 matrix: one row per line, "0"/"1" tokens separated by whitespace, blank lines
 * skipped; out gets rows x cols bytes (row-major, at most cap; NULL to size only).
 * Pauli CSS stabilisers: one per line (I . _ space = identity, X, Z), a blank line ends
 * the input; Y or mixed X/Z lines are rejected; all-identity lines are dropped; hx / hz get
 * the X-type / Z-type rows (n bytes each, at most cap bytes per matrix).
 * Return 0, or 2 with distfind_last_error() saying what is wrong (1 for NULL outputs). */
int distfind_parse_matrix_text(const char* text, uint8_t* out, int64_t cap, int* rows, int* cols);
int distfind_parse_pauli_css(const char* text, uint8_t* hx, uint8_t* hz, int64_t cap, int* hx_rows, int* hz_rows,
                             int* n);

#ifdef __cplusplus
}
#endif

#endif
