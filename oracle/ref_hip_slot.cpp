// Test infrastructure: the reference built with this repository's HIP plugin
// (qubitserf_amd/plugin/backend_hip.cpp) in its accelerator slot. The reference's
// backend_cpu.cpp is compiled with -DQSF_CUDA, which leaves cuda_backend() to the
// accelerator translation unit (cpp/qsf/backend_cpu.cpp:80-82); here that unit is the HIP
// plugin instead of cpp/qsf/cuda/backend_cuda.cu. The reference's own drivers, C ABI and
// select_backend then run unchanged, with every "gpu"/"auto" level going through
// HipBackend. Linked into oracle/_ref/libqubitserf_ref_hip.so by oracle/ref.mk.
//
// ref_hip_plugin_enumerate / ref_hip_plugin_collect call the plugin's two virtuals on one
// level directly (not through select_backend), so a test can check the fallback contract
// on a machine without a device.
#include <cstdint>
#include <cstring>

#include "qsf/backend.hpp"
#include "qsf/combinatorics.hpp"

namespace qsf {
Backend* hip_backend();
Backend* cuda_backend() { return hip_backend(); }
}  // namespace qsf

namespace {
qsf::EnumPlan make_plan(const qsf::BinomTable& bt, int n, int stride, int K, int d, int ng, const uint64_t* gamma,
                        int kcheck, const uint64_t* check, int current_best) {
    qsf::EnumPlan p;
    p.n = n; p.stride = stride; p.K = K; p.d = d;
    p.num_gamma = ng; p.num_gamma_total = ng;
    p.gamma = gamma; p.kcheck = kcheck; p.check = check;
    p.binom = bt.c.data(); p.binom_maxN = bt.maxN; p.binom_maxK = bt.maxK;
    p.current_best = current_best;
    p.buffers_key = qsf::next_buffers_key();
    return p;
}
}  // namespace

extern "C" int ref_hip_plugin_enumerate(int n, int stride, int K, int d, int ng, const uint64_t* gamma, int kcheck,
                                        const uint64_t* check, int current_best) {
    qsf::BinomTable bt(K, K);
    return qsf::hip_backend()->enumerate(make_plan(bt, n, stride, K, d, ng, gamma, kcheck, check, current_best));
}

// Rows of the level's collection (duplicates allowed); the first `cap` rows are copied to out.
extern "C" int64_t ref_hip_plugin_collect(int n, int stride, int K, int d, const uint64_t* gamma, int keep,
                                          uint64_t* out, int64_t cap, int* ok) {
    qsf::BinomTable bt(K, K);
    qsf::CollectResult r =
        qsf::hip_backend()->collect(make_plan(bt, n, stride, K, d, 1, gamma, 0, nullptr, qsf::WEIGHT_NONE), keep);
    const int64_t rows = (int64_t)(r.hits.size() / (size_t)stride);
    if (out) std::memcpy(out, r.hits.data(), (size_t)(rows < cap ? rows : cap) * stride * sizeof(uint64_t));
    if (ok) *ok = r.ok ? 1 : 0;
    return rows;
}

extern "C" const char* ref_hip_plugin_name(void) {
    static std::string s = qsf::hip_backend()->name();
    return s.c_str();
}
