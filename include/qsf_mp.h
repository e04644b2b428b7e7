/* Device kernels of the learned information-set proposal (no reference counterpart: the
 * reference builds its information sets greedily, cpp/distfind/css.cpp:177-203). A
 * three-layer message-passing scorer on the Tanner graph, BF16 operands with FP32
 * accumulation. All tensor pointers are device pointers; `stream` is a hipStream_t (NULL =
 * the null stream). Kernel calls are asynchronous on that stream. Return 0 on success, 1 bad
 * arguments, 3 launch failure. */
#ifndef QSF_MP_H
#define QSF_MP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device memory for the scorer's tensors (hipMalloc / hipMemcpy / hipMemset on the null
 * stream). kind: 1 host to device, 2 device to host, 3 device to device. */
void* qsf_dev_alloc(size_t bytes);
void qsf_dev_free(void* p);
int qsf_dev_copy(void* dst, const void* src, size_t bytes, int kind);
int qsf_dev_zero(void* dst, size_t bytes);
int qsf_dev_sync(void);

/* C[b] (M x N) = op(A[b]) (M x K) * op(B[b]) (K x N), b < batch, FP32 accumulation.
 * A, B: BF16 (uint16 bit patterns). op(A)[i][k] = transA ? A[k lda + i] : A[i lda + k];
 * op(B)[k][j] = transB ? B[j ldb + k] : B[k ldb + j]. C: FP32, or BF16 (out_bf16 = 1,
 * round to nearest even). accumulate = 1 adds the old C; bias (FP32, N values, may be
 * NULL) is added per column; relu = 1 clamps at 0. Batch strides sA, sB, sC in elements. */
typedef struct {
    const void* A;
    const void* B;
    void* C;
    const float* bias;
    int batch, M, N, K, lda, ldb, ldc;
    long long sA, sB, sC;
    int transA, transB, out_bf16, accumulate, relu;
} qsf_mp_gemm_desc;
int qsf_mp_gemm(const qsf_mp_gemm_desc* d, void* stream);

/* The scorer's inputs for `batch` codes from their 0/1 parity checks H (batch x m x n bytes,
 * code b's hm[b] x hn[b] in the top-left corner, zeros elsewhere): Ac (batch x m x n) and Av
 * (batch x n x m), H and H^T scaled by the inverse row and column degrees, and the column
 * features X (batch x n x 16: mask, degree, mean check degree, m/n, noise[j][0..3], zeros),
 * all BF16 patterns; inv[k - 1] is the pattern of 1 / k for k up to max(m, n), noise is
 * n x 4 floats (every code's first hn[b] rows). Returns 1 for bad arguments or m + n >
 * 12288. */
int qsf_mp_pack(const uint8_t* H, const int* hm, const int* hn, int batch, int m, int n, const uint16_t* inv,
                const float* noise, uint16_t* Ac, uint16_t* Av, uint16_t* X, void* stream);

int qsf_mp_to_bf16(const float* x, uint16_t* y, long long n, void* stream);
/* dy[i] = 0 where the BF16 activation y[i] <= 0 */
int qsf_mp_relu_bwd(const uint16_t* y, float* dy, long long n, void* stream);
/* out[j] (+)= sum_r x[r cols + j] */
int qsf_mp_colsum(const float* x, float* out, long long rows, int cols, int accumulate, void* stream);
/* *loss += sum_i mask_i BCE(z_i, t_i) inv_count; dz_i = mask_i (sigmoid(z_i) - t_i) inv_count */
int qsf_mp_bce(const float* z, const float* t, const float* mask, float* dz, float* loss, long long n, float inv_count,
               void* stream);
/* Adam step `step` (1-based) on FP32 weights w with moments m, v; wb (may be NULL) gets BF16 copies */
int qsf_mp_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, long long n, float lr, float b1, float b2,
                float eps, int step, void* stream);

#ifdef __cplusplus
}
#endif

#endif
