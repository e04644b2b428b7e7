// Calibration probes for the measurements (a separate executable, not part of the library).
//
//   qsa_calib valu            integer-VALU peak: every lane runs 8 independent chains of
//                             x = popcount(x ^ k) + x, i.e. one v_xor_b32 + one
//                             v_bcnt_u32_b32 per step, over 2048 x 256 threads. Prints the
//                             achieved int32 lane-ops/s.
//   qsa_calib copy <MiB>      a plain device copy of <MiB> MiB (one read + one write of each
//                             byte); used to scale the FETCH_SIZE / WRITE_SIZE counters.
//   qsa_calib mfma            int8 matrix-core peak: each wave chains 8 independent
//                             v_mfma_i32_32x32x32_i8 accumulators; prints int8 ops/s
//                             (2 per multiply-add).
//   qsa_calib mfma_max        the int8 hot-test pattern: per step 8 products (zero
//                             accumulator) reduced with v_max3; prints products/s.
//   qsa_calib mfma_issue      the hot test's products alone: per step 8 independent
//                             products (zero accumulator), one value of each consumed;
//                             prints products/s (the matrix cores' rate for this shape).
//   qsa_calib mfma_valu       8 such products plus 64 v_min3 that do not read them:
//                             shows how far VALU work overlaps the products on a SIMD.
//   qsa_calib f4              fp4 matrix-core rate: 8 independent chains of
//                             v_mfma_scale_f32_32x32x64_f8f6f4 (fp4 operands, unit scales);
//                             prints fp4 ops/s (2 per multiply-add).
//   qsa_calib f4_max          the fp4 hot-test pattern: per step 8 such products (zero
//                             accumulator) reduced with v_min3_i32 over their bit patterns;
//                             prints products/s (1024 codewords each).
//   An optional last argument sets the waves per CU of the mfma probes (default 12).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

__global__ __launch_bounds__(256) void valu_chain(unsigned* out, unsigned key, int iters) {
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned x0 = t, x1 = t * 3u, x2 = t * 5u, x3 = t * 7u, x4 = t * 11u, x5 = t * 13u, x6 = t * 17u, x7 = t * 19u;
    for (int i = 0; i < iters; ++i) {
        x0 = __popc(x0 ^ key) + x0;
        x1 = __popc(x1 ^ key) + x1;
        x2 = __popc(x2 ^ key) + x2;
        x3 = __popc(x3 ^ key) + x3;
        x4 = __popc(x4 ^ key) + x4;
        x5 = __popc(x5 ^ key) + x5;
        x6 = __popc(x6 ^ key) + x6;
        x7 = __popc(x7 ^ key) + x7;
    }
    out[t] = x0 ^ x1 ^ x2 ^ x3 ^ x4 ^ x5 ^ x6 ^ x7;
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64) void mfma_chain(int* out, int iters) {
    const int l = threadIdx.x;
    v4i a[8], b = {l, l * 3, l * 5, l * 7};
    v16i c[8];
    for (int k = 0; k < 8; ++k) {
        a[k] = v4i{l + k, l ^ k, l * k, l - k};
        c[k] = v16i{};
    }
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) c[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[k], b, c[k], 0, 0, 0);
    int r = 0;
    for (int k = 0; k < 8; ++k)
        for (int v = 0; v < 16; ++v) r ^= c[k][v];
    out[blockIdx.x * 64 + l] = r;
}

__global__ __launch_bounds__(64) void mfma_max(int* out, int iters) {
    const int l = threadIdx.x;
    v4i a[8], b = {l, l * 3, l * 5, l * 7};
    for (int k = 0; k < 8; ++k) a[k] = v4i{l + k, l ^ k, l * k, l - k};
    int hit = 0;
    for (int i = 0; i < iters; ++i) {
        b[0] += 1;
        int mx = -(1 << 30);
        const v16i z = {};
#pragma unroll
        for (int k0 = 0; k0 < 8; k0 += 4) {
            const v16i p0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[k0 + 0], b, z, 0, 0, 0);
            const v16i p1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[k0 + 1], b, z, 0, 0, 0);
            const v16i p2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[k0 + 2], b, z, 0, 0, 0);
            const v16i p3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[k0 + 3], b, z, 0, 0, 0);
            auto red = [](const v16i& c) {
                int m = max(max(c[0], c[1]), c[2]);
#pragma unroll
                for (int v = 3; v < 15; v += 2) m = max(max(m, c[v]), c[v + 1]);
                return max(m, c[15]);
            };
            mx = max(mx, max(max(red(p0), red(p1)), max(red(p2), red(p3))));
        }
        hit += __ballot(mx > (1 << 29)) != 0ull;
    }
    out[blockIdx.x * 64 + l] = hit;
}

// V = 0: products only; V = 1: products plus 64 independent v_min3 per step
template <int V>
__global__ __launch_bounds__(64) void mfma_issue(int* out, int iters) {
    const int l = threadIdx.x;
    v4i a[8], b = {l, l * 3, l * 5, l * 7};
    for (int k = 0; k < 8; ++k) a[k] = v4i{l + k, l ^ k, l * k, l - k};
    v16i w;
    for (int v = 0; v < 16; ++v) w[v] = l * v + 7 * v;
    int hit = 0;
    const v16i z = {};
    for (int i = 0; i < iters; ++i) {
        b[0] += 1;
        int mn = 1 << 30;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const v16i p = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[k], b, z, 0, 0, 0);
            mn = min(mn, p[k]);
            if (V == 1) {
                asm volatile("" : "+v"(w));   // fresh values: the min3 chain is not hoisted
#pragma unroll
                for (int v = 0; v < 16; v += 2) mn = min(min(mn, w[v]), w[v + 1]);
            }
        }
        hit += __ballot(mn < -(1 << 29)) != 0ull;
    }
    out[blockIdx.x * 64 + l] = hit;
}

typedef int v8i_ __attribute__((ext_vector_type(8)));
typedef float v16f_ __attribute__((ext_vector_type(16)));
__device__ __forceinline__ v16f_ f4mm(v8i_ a, v8i_ b, v16f_ c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 4, 4, 0, 127, 0, 127);
}

__global__ __launch_bounds__(64) void f4_chain(float* out, int iters) {
    const int l = threadIdx.x;
    v8i_ a[8], b = {l * 0x12345, l * 0x2233, l * 5, l * 7, 0, 0, 0, 0};
    v16f_ c[8];
    for (int k = 0; k < 8; ++k) {
        a[k] = v8i_{(l + k) * 0x1111, (l ^ k) * 0x2222, l * k, l - k, 0, 0, 0, 0};
        c[k] = v16f_{};
    }
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < 8; ++k) c[k] = f4mm(a[k], b, c[k]);
    float r = 0;
    for (int k = 0; k < 8; ++k)
        for (int v = 0; v < 16; ++v) r += c[k][v];
    out[blockIdx.x * 64 + l] = r;
}

__global__ __launch_bounds__(64) void f4_max(int* out, int iters) {
    const int l = threadIdx.x;
    v8i_ a[8], b = {0x22222222, 0x22222222 ^ (l << 3), 0x2a2a2a2a, 0x22a222a2, 0, 0, 0, 0};
    for (int k = 0; k < 8; ++k) a[k] = v8i_{0x22222222 ^ (k << 7), 0x2a2a2a2a, 0x22222222 ^ (l << 11), 0x22222222, 0, 0, 0, 0};
    int hit = 0;
    for (int i = 0; i < iters; ++i) {
        b[0] ^= (i & 7) << 3;
        int mn = 0x7fffffff;
        const v16f_ z = {};
#pragma unroll
        for (int k0 = 0; k0 < 8; k0 += 4) {
            const v16f_ p0 = f4mm(a[k0 + 0], b, z);
            const v16f_ p1 = f4mm(a[k0 + 1], b, z);
            const v16f_ p2 = f4mm(a[k0 + 2], b, z);
            const v16f_ p3 = f4mm(a[k0 + 3], b, z);
            auto red = [](const v16f_& c) {
                int m = min(min(__float_as_int(c[0]), __float_as_int(c[1])), __float_as_int(c[2]));
#pragma unroll
                for (int v = 3; v < 15; v += 2) m = min(min(m, __float_as_int(c[v])), __float_as_int(c[v + 1]));
                return min(m, __float_as_int(c[15]));
            };
            mn = min(mn, min(min(red(p0), red(p1)), min(red(p2), red(p3))));
        }
        hit += __ballot(mn < -1000) != 0ull;
    }
    out[blockIdx.x * 64 + l] = hit;
}

__global__ __launch_bounds__(256) void copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: qsa_calib valu | copy <MiB>\n");
        return 2;
    }
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    if (std::strcmp(argv[1], "valu") == 0) {
        const int blocks = 2048 * 4, threads = 256, iters = 20000;
        unsigned* out;
        CHECK(hipMalloc(&out, (size_t)blocks * threads * sizeof(unsigned)));
        hipLaunchKernelGGL(valu_chain, dim3(blocks), dim3(threads), 0, 0, out, 0x9e3779b9u, 100);  // warm up
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(valu_chain, dim3(blocks), dim3(threads), 0, 0, out, 0x9e3779b9u, iters);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double ops = (double)blocks * threads * iters * 8 * 2;
        std::printf("{\"probe\": \"valu\", \"lane_ops\": %.6e, \"ms\": %.3f, \"tops\": %.4f}\n", ops, ms,
                    ops / (ms * 1e-3) / 1e12);
        CHECK(hipFree(out));
        return 0;
    }
    if (std::strcmp(argv[1], "copy") == 0 && argc >= 3) {
        const size_t mib = (size_t)std::atol(argv[2]);
        const size_t bytes = mib << 20, n = bytes / sizeof(uint4);
        uint4 *a, *b;
        CHECK(hipMalloc(&a, bytes));
        CHECK(hipMalloc(&b, bytes));
        CHECK(hipMemset(a, 1, bytes));
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(copy_u4, dim3(4096), dim3(256), 0, 0, a, b, n);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("{\"probe\": \"copy\", \"bytes\": %zu, \"ms\": %.3f, \"GBps\": %.1f}\n", bytes, ms,
                    2.0 * bytes / (ms * 1e-3) / 1e9);
        CHECK(hipFree(a));
        CHECK(hipFree(b));
        return 0;
    }
    const bool m32c = std::strcmp(argv[1], "mfma") == 0, m32r = std::strcmp(argv[1], "mfma_max") == 0;
    const bool m32i = std::strcmp(argv[1], "mfma_issue") == 0, m32v = std::strcmp(argv[1], "mfma_valu") == 0;
    const bool f4c = std::strcmp(argv[1], "f4") == 0, f4r = std::strcmp(argv[1], "f4_max") == 0;
    if (f4c || f4r) {
        const int per_cu = argc >= 3 ? std::atoi(argv[2]) : 12;
        const int blocks = 256 * per_cu, iters = 2000;
        void* out;
        CHECK(hipMalloc(&out, (size_t)blocks * 64 * sizeof(float)));
        auto launch = [&](int it) {
            if (f4c) hipLaunchKernelGGL(f4_chain, dim3(blocks), dim3(64), 0, 0, (float*)out, it);
            else hipLaunchKernelGGL(f4_max, dim3(blocks), dim3(64), 0, 0, (int*)out, it);
        };
        launch(10);
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, 0));
        launch(iters);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double mfmas = (double)blocks * iters * 8;
        std::printf("{\"probe\": \"%s\", \"waves_per_cu\": %d, \"ms\": %.3f, \"fp4_ops_per_s\": %.4e, \"products_per_s\": %.4e}\n",
                    argv[1], per_cu, ms, mfmas * 65536.0 * 2 / (ms * 1e-3), mfmas * 1024 / (ms * 1e-3));
        CHECK(hipFree(out));
        return 0;
    }
    if (m32c || m32r || m32i || m32v) {
        const int per_cu = argc >= 3 ? std::atoi(argv[2]) : 12;
        const int blocks = 256 * per_cu, iters = m32c ? 4000 : 2000;
        int* out;
        CHECK(hipMalloc(&out, (size_t)blocks * 64 * sizeof(int)));
        auto launch = [&](int it) {
            if (m32c) hipLaunchKernelGGL(mfma_chain, dim3(blocks), dim3(64), 0, 0, out, it);
            else if (m32r) hipLaunchKernelGGL(mfma_max, dim3(blocks), dim3(64), 0, 0, out, it);
            else if (m32i) hipLaunchKernelGGL(mfma_issue<0>, dim3(blocks), dim3(64), 0, 0, out, it);
            else hipLaunchKernelGGL(mfma_issue<1>, dim3(blocks), dim3(64), 0, 0, out, it);
        };
        launch(10);
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0, 0));
        launch(iters);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double mfmas = (double)blocks * iters * 8;
        const double ops_each = 65536.0;
        if (m32c)
            std::printf("{\"probe\": \"%s\", \"waves_per_cu\": %d, \"int8_ops\": %.6e, \"ms\": %.3f, \"tops\": %.1f}\n", argv[1],
                        per_cu, mfmas * ops_each, ms, mfmas * ops_each / (ms * 1e-3) / 1e12);
        else {
            const double cw = mfmas * 1024.0;
            std::printf("{\"probe\": \"%s\", \"waves_per_cu\": %d, \"products\": %.6e, \"ms\": %.3f, \"products_per_s\": %.4e}\n",
                        argv[1], per_cu, cw, ms, cw / (ms * 1e-3));
        }
        CHECK(hipFree(out));
        return 0;
    }
    std::fprintf(stderr, "unknown probe\n");
    return 2;
}
