// The gfx950 engine as the reference's GPU plugin: a qsf::Backend (cpp/qsf/backend.hpp:53-61)
// over this library's plugin-level C ABI (include/qsf_hip.h). It is source for the
// reference's tree (it includes the reference's own qsf/backend.hpp), in the place of the
// CUDA translation unit cpp/qsf/cuda/backend_cuda.cu, and links against
// libqubitserf_amd.so:
//
//   g++ -std=c++17 -I<reference>/cpp -I<this repo>/include -c backend_hip.cpp
//
// It keeps the reference backend contract that CudaBackend keeps (backend_cuda.cu:506-645):
// a level the device does not run (no device, a level the engine refuses, a HIP error)
// is computed by the reference's own CPU backend, so enumerate() and collect() never throw
// and collect() is always complete. The reference's drivers rely on both: its extern "C"
// entry points have no try/catch (cpp/distfind/capi.cpp), and cpp/codeaut/bz.cpp:93-115
// reports a collection as complete without reading CollectResult.ok.
//
// Selection: hip_backend() is the factory. A maintainer adds one line to select_backend
// (cpp/qsf/backend_cpu.cpp:84-93), before the CUDA and Metal probes:
//     if (Backend* b = hip_backend(); ok(b)) return b;
// or, with no edit to the reference's sources, builds backend_cpu.cpp with -DQSF_CUDA and
// defines cuda_backend() as hip_backend() (the accelerator slot the CUDA TU fills).
#include <cstddef>
#include <cstdint>

#include "qsf/backend.hpp"
#include "qsf_hip.h"

namespace qsf {

// qsf_hip_enumerate takes the reference's plan by address: the two layouts must agree
// field for field (cpp/qsf/backend.hpp:21-42 against include/qsf_hip.h).
static_assert(sizeof(EnumPlan) == sizeof(qsf_enum_plan), "EnumPlan size");
static_assert(alignof(EnumPlan) == alignof(qsf_enum_plan), "EnumPlan alignment");
#define QSF_HIP_SAME_FIELD(f)                                                            \
    static_assert(offsetof(EnumPlan, f) == offsetof(qsf_enum_plan, f), "EnumPlan::" #f); \
    static_assert(sizeof(EnumPlan::f) == sizeof(qsf_enum_plan::f), "EnumPlan::" #f " size")
QSF_HIP_SAME_FIELD(n);
QSF_HIP_SAME_FIELD(stride);
QSF_HIP_SAME_FIELD(K);
QSF_HIP_SAME_FIELD(d);
QSF_HIP_SAME_FIELD(num_gamma);
QSF_HIP_SAME_FIELD(num_gamma_total);
QSF_HIP_SAME_FIELD(gamma);
QSF_HIP_SAME_FIELD(kcheck);
QSF_HIP_SAME_FIELD(check);
QSF_HIP_SAME_FIELD(binom);
QSF_HIP_SAME_FIELD(binom_maxN);
QSF_HIP_SAME_FIELD(binom_maxK);
QSF_HIP_SAME_FIELD(current_best);
QSF_HIP_SAME_FIELD(buffers_key);
#undef QSF_HIP_SAME_FIELD
static_assert(WEIGHT_NONE == QSF_WEIGHT_NONE, "no-codeword sentinel");

namespace {

const qsf_enum_plan* as_plan(const EnumPlan& p) { return reinterpret_cast<const qsf_enum_plan*>(&p); }

struct HipBackend : Backend {
    std::string name() const override { return "hip"; }

    bool available() const override { return qsf_hip_available() == 1; }

    int enumerate(const EnumPlan& plan) override {
        int best = plan.current_best;
        if (qsf_hip_enumerate(as_plan(plan), &best) == 0) return best;
        // no device, a level the engine refuses (n > 1024, more than 2^62 codewords) or a
        // HIP error: the reference's CPU enumerator answers, as in backend_cuda.cu:507-520
        return cpu_backend()->enumerate(plan);
    }

    CollectResult collect(const EnumPlan& plan, int keep_weight) override {
        if (keep_weight <= 0) return CollectResult{};   // nothing qualifies (backend_cuda.cu:562-564)
        uint64_t* hits = nullptr;
        int64_t rows = 0;
        if (qsf_hip_collect(as_plan(plan), keep_weight, &hits, &rows) != 0)
            return cpu_backend()->collect(plan, keep_weight);   // complete either way
        CollectResult r;
        try {
            r.hits.assign(hits, hits + (size_t)rows * (size_t)plan.stride);
        } catch (...) {
            qsf_hip_free(hits);
            return cpu_backend()->collect(plan, keep_weight);
        }
        qsf_hip_free(hits);
        return r;
    }
};

}  // namespace

Backend* hip_backend() {
    static HipBackend inst;
    return &inst;
}

}  // namespace qsf
