/* codeaut low-weight collection C ABI of the gfx950 engine — the codeaut_bz_* symbols of
 * the reference (cpp/codeaut/capi.h:57-69), same signatures and handle semantics.
 *
 * codeaut_bz_collect enumerates, for every subset size sw = 1..p, all XORs of sw rows of the
 * m x n generator G with Hamming weight in [1, keep_weight] (on the device: collect mode
 * of the enumeration engine), then sorts and deduplicates them (cpp/codeaut/bz.cpp:58-117).
 * backend: 0 = cpu (not provided here: the handle reports ok = 0), 1 = gpu, 2 = auto.
 * threads is accepted and ignored. */
#ifndef QSA_CODEAUT_CAPI_H
#define QSA_CODEAUT_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void* codeaut_bz_collect(const uint8_t* G, int32_t m, int32_t n, int32_t p, int32_t keep_weight,
                         int32_t backend, int32_t threads);
int32_t codeaut_bz_ok(void* h);                       /* 1 if the collection is complete */
int64_t codeaut_bz_count(void* h);                    /* distinct codewords collected */
int64_t codeaut_bz_combos(void* h);                   /* subset combinations evaluated */
void codeaut_bz_copy_rows(void* h, uint8_t* buf);     /* count x n 0/1 bytes */
const char* codeaut_bz_backend(void* h);              /* "hip" */
void codeaut_bz_free(void* h);

/* "gpu" / "auto" -> 1 with a gfx950 device; "cpu" -> 0 (cpp/codeaut/capi.cpp:11-21). */
int32_t codeaut_backend_available(const char* name);
const char* codeaut_version(void);

/* Leon's code-automorphism ABI (cpp/codeaut/capi.h:20-48), which the reference's codeaut
 * binding attaches without a guard (python/qubitserf/codeaut/_native.py:31-61). Leon is not
 * part of this engine: qaut_leon_run returns a handle whose qaut_leon_ok is 0 (dim -1, every
 * count 0, the copy functions write nothing); free it with qaut_leon_free. */
void* qaut_leon_run(const uint8_t* G, int32_t m, int32_t n, int32_t max_dim, int32_t use_invariant);
int32_t qaut_leon_ok(void* h);
int32_t qaut_leon_dim(void* h);
int32_t qaut_leon_n(void* h);
int32_t qaut_leon_num_codewords(void* h);
int32_t qaut_leon_num_classes(void* h);
int32_t qaut_leon_num_gens(void* h);
int32_t qaut_leon_num_factors(void* h);
void qaut_leon_copy_gens(void* h, int32_t* buf);
void qaut_leon_copy_factors(void* h, int64_t* buf);
void qaut_leon_copy_weights(void* h, int32_t* buf);
void qaut_leon_free(void* h);

#ifdef __cplusplus
}
#endif

#endif
