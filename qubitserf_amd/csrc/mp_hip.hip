// Kernels of the learned information-set proposal: a Tanner-graph message-passing scorer
// with BF16 operands and FP32 accumulation (MFMA), trained and run on the device.
//
// Everything here is dense and small (codes of n <= 256 columns, hidden width 64), so the
// network is a chain of batched GEMMs (adjacency products and 64-wide linear layers) plus
// elementwise kernels. GEMM: one wave per 16 x 16 output tile, v_mfma_f32_16x16x16_bf16
// steps over K; operands read straight from HBM/L2 (they are a few hundred KB).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/qsf_mp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct GemmArgs {
    const uint16_t* A;   // bf16
    const uint16_t* B;   // bf16
    void* C;             // f32 or bf16
    const float* bias;   // per output column, or null
    int M, N, K, lda, ldb, ldc;
    long long sA, sB, sC;   // batch strides (elements)
    int transA, transB, out_bf16, accumulate, relu;
};

// C[b] (M x N) = op(A[b]) (M x K) * op(B[b]) (K x N) [+ C] [+ bias] [relu]
// op(A)[i][k] = transA ? A[k * lda + i] : A[i * lda + k]; op(B)[k][j] likewise.
__global__ __launch_bounds__(64) void gemm_bf16(GemmArgs g) {
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16, b = blockIdx.z;
    const uint16_t* A = g.A + (size_t)b * g.sA;
    const uint16_t* B = g.B + (size_t)b * g.sB;
    const int r = lane & 15, q = lane >> 4;   // operand row/column within the tile, k group
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4 a, bb;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int k = k0 + 4 * q + t;
            const int i = i0 + r, j = j0 + r;
            uint16_t av = 0, bv = 0;
            if (i < g.M && k < g.K) av = g.transA ? A[(size_t)k * g.lda + i] : A[(size_t)i * g.lda + k];
            if (j < g.N && k < g.K) bv = g.transB ? B[(size_t)j * g.ldb + k] : B[(size_t)k * g.ldb + j];
            a[t] = (short)av;
            bb[t] = (short)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
    }
    // lane holds D[4q + t][r]
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int i = i0 + 4 * q + t, j = j0 + r;
        if (i >= g.M || j >= g.N) continue;
        const size_t at = (size_t)b * g.sC + (size_t)i * g.ldc + j;
        float v = acc[t];
        if (g.accumulate) v += g.out_bf16 ? bf2f(((const uint16_t*)g.C)[at]) : ((const float*)g.C)[at];
        if (g.bias) v += g.bias[j];
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.out_bf16) ((uint16_t*)g.C)[at] = f2bf(v);
        else ((float*)g.C)[at] = v;
    }
}

__global__ void to_bf16(const float* x, uint16_t* y, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// dy *= (y > 0): the ReLU derivative taken at the stored (bf16) activation
__global__ void relu_bwd(const uint16_t* y, float* dy, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (!(bf2f(y[i]) > 0.f)) dy[i] = 0.f;
}

// out[j] (+)= sum over rows of x[row][j]   (bias gradients); one thread per column
__global__ void colsum(const float* x, float* out, long long rows, int cols, int accumulate) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (long long r = 0; r < rows; ++r) s += x[r * cols + j];
    out[j] = accumulate ? out[j] + s : s;
}

// mean binary cross-entropy with logits over masked entries; dz = mask (sigmoid(z) - t) / count
__global__ void bce(const float* z, const float* t, const float* mask, float* dz, float* loss, long long n, float inv_count) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float m = mask[i];
        const float p = 1.f / (1.f + __expf(-z[i]));
        dz[i] = m * (p - t[i]) * inv_count;
        if (m != 0.f) {
            // log(1 + e^z) - t z, computed without overflow
            const float l = fmaxf(z[i], 0.f) - z[i] * t[i] + __logf(1.f + __expf(-fabsf(z[i])));
            atomicAdd(loss, m * l * inv_count);
        }
    }
}

// Adam on FP32 master weights; writes the BF16 copy the GEMMs read
__global__ void adam(float* w, const float* g, float* m, float* v, uint16_t* wb, long long n, float lr, float b1,
                     float b2, float eps, float c1, float c2) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

// One workgroup per code: the scorer's operators and input features from the code's 0/1
// parity check H (hm[b] x hn[b] of the padded m x n), as bfloat16 bit patterns:
//   Ac[b][i][j] = H[i][j] ? inv[max(dc_i, 1) - 1] : 0       (dc_i: ones in row i)
//   Av[b][j][i] = H[i][j] ? inv[max(dv_j, 1) - 1] : 0       (dv_j: ones in column j)
//   X[b][j][f], f < 16: mask_j = (j < hn), dv_j / max(1, max dv),
//     (sum_i dc_i H[i][j]) / max(dv_j, 1) / max(1, max dc), (hm / hn) mask_j,
//     noise[j][0..3] mask_j, then zeros
// with inv[k - 1] the pattern of 1 / k. The same values, bit for bit, as the host packing
// (qubitserf_amd/proposal/model.py:pack_batch_bits): the sums are exact integers and every
// division is one correctly rounded f32 division, in the same order.
__global__ void pack_ops(const uint8_t* H, const int* hm, const int* hn, int m, int n, const uint16_t* inv,
                         const float* noise, uint16_t* Ac, uint16_t* Av, uint16_t* X) {
    extern __shared__ int deg[];   // dc[0..m), dv[m..m+n)
    __shared__ int mx[2];
    const int b = blockIdx.x, t = threadIdx.x;
    const uint8_t* h = H + (size_t)b * m * n;
    if (t == 0) mx[0] = mx[1] = 0;
    for (int i = t; i < m; i += blockDim.x) {
        int c = 0;
        for (int j = 0; j < n; ++j) c += h[(size_t)i * n + j];
        deg[i] = c;
    }
    for (int j = t; j < n; j += blockDim.x) {
        int c = 0;
        for (int i = 0; i < m; ++i) c += h[(size_t)i * n + j];
        deg[m + j] = c;
    }
    __syncthreads();
    for (int i = t; i < m; i += blockDim.x) atomicMax(&mx[0], deg[i]);
    for (int j = t; j < n; j += blockDim.x) atomicMax(&mx[1], deg[m + j]);
    __syncthreads();
    const float mdc = fmaxf(1.f, (float)mx[0]), mdv = fmaxf(1.f, (float)mx[1]);
    const float ratio = (float)((double)hm[b] / (double)hn[b]);
    for (int e = t; e < m * n; e += blockDim.x) {
        const int i = e / n, j = e - i * n;
        const bool one = h[e] != 0;
        Ac[(size_t)b * m * n + e] = one ? inv[max(deg[i], 1) - 1] : (uint16_t)0;
        Av[(size_t)b * n * m + (size_t)j * m + i] = one ? inv[max(deg[m + j], 1) - 1] : (uint16_t)0;
    }
    for (int j = t; j < n; j += blockDim.x) {
        int s = 0;
        for (int i = 0; i < m; ++i) s += h[(size_t)i * n + j] ? deg[i] : 0;
        const float dv = (float)deg[m + j];
        const float mask = j < hn[b] ? 1.f : 0.f;
        float f[16] = {};
        f[0] = mask;
        f[1] = dv / mdv;
        f[2] = ((float)s / fmaxf(dv, 1.f)) / mdc;
        f[3] = ratio * mask;
        if (j < hn[b])
            for (int k = 0; k < 4; ++k) f[4 + k] = noise[j * 4 + k];
        uint16_t* x = X + ((size_t)b * n + j) * 16;
        for (int k = 0; k < 16; ++k) x[k] = f2bf(f[k]);
    }
}

unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

extern "C" {

void* qsf_dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

void qsf_dev_free(void* p) {
    if (p) (void)hipFree(p);
}

int qsf_dev_copy(void* dst, const void* src, size_t bytes, int kind) {
    if (!bytes) return 0;
    if (!dst || !src || kind < 1 || kind > 3) return 1;
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? 0 : 3;
}

int qsf_dev_zero(void* dst, size_t bytes) {
    if (!bytes) return 0;
    if (!dst) return 1;
    return hipMemset(dst, 0, bytes) == hipSuccess ? 0 : 3;
}

int qsf_dev_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : 3; }

int qsf_mp_gemm(const qsf_mp_gemm_desc* d, void* stream) {
    if (!d || !d->A || !d->B || !d->C || d->M < 0 || d->N < 0 || d->K < 0 || d->batch < 0) return 1;
    if (d->M == 0 || d->N == 0 || d->batch == 0) return 0;
    GemmArgs g{(const uint16_t*)d->A, (const uint16_t*)d->B, d->C, d->bias, d->M, d->N, d->K, d->lda, d->ldb, d->ldc,
               d->sA, d->sB, d->sC, d->transA, d->transB, d->out_bf16, d->accumulate, d->relu};
    dim3 grid((unsigned)((d->M + 15) / 16), (unsigned)((d->N + 15) / 16), (unsigned)d->batch);
    hipLaunchKernelGGL(gemm_bf16, grid, dim3(64), 0, (hipStream_t)stream, g);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int qsf_mp_pack(const uint8_t* H, const int* hm, const int* hn, int batch, int m, int n, const uint16_t* inv,
                const float* noise, uint16_t* Ac, uint16_t* Av, uint16_t* X, void* stream) {
    if (batch < 0 || m < 0 || n < 0 || (batch > 0 && (!H || !hm || !hn || !inv || !noise || !Ac || !Av || !X))) return 1;
    if (batch == 0 || n == 0) return 0;
    const size_t lds = (size_t)(m + n) * sizeof(int);
    if (lds > 48 * 1024) return 1;
    hipLaunchKernelGGL(pack_ops, dim3((unsigned)batch), dim3(256), lds, (hipStream_t)stream, H, hm, hn, m, n, inv, noise,
                       Ac, Av, X);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int qsf_mp_to_bf16(const float* x, uint16_t* y, long long n, void* stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(to_bf16, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, n);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int qsf_mp_relu_bwd(const uint16_t* y, float* dy, long long n, void* stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(relu_bwd, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, y, dy, n);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int qsf_mp_colsum(const float* x, float* out, long long rows, int cols, int accumulate, void* stream) {
    if (cols <= 0) return 0;
    hipLaunchKernelGGL(colsum, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, (hipStream_t)stream, x, out, rows, cols,
                       accumulate);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int qsf_mp_bce(const float* z, const float* t, const float* mask, float* dz, float* loss, long long n, float inv_count,
               void* stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(bce, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, z, t, mask, dz, loss, n, inv_count);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

int qsf_mp_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, long long n, float lr, float b1, float b2,
                float eps, int step, void* stream) {
    if (n <= 0) return 0;
    float c1 = 1.f, c2 = 1.f;
    for (int i = 0; i < step; ++i) { c1 *= b1; c2 *= b2; }
    hipLaunchKernelGGL(adam, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, wb, n, lr, b1, b2, eps,
                       1.f - c1, 1.f - c2);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // extern "C"
