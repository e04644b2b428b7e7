// Test infrastructure: a C entry point into the reference's own CPU enumerator
// (qsf::cpu_backend()->enumerate, cpp/qsf/backend_cpu.cpp:33-47), so a single weight
// level can be timed and checked without running a whole solve. Linked into
// oracle/_ref/libqubitserf_ref.so by oracle/ref.mk together with the reference sources.
#include <cstdint>
#include "qsf/backend.hpp"
#include "qsf/combinatorics.hpp"

extern "C" int ref_enumerate_plan(int n, int stride, int K, int d, int ng, const uint64_t* gamma, int kcheck,
                                  const uint64_t* check, int current_best, int threads) {
    qsf::BinomTable bt(K, K);
    qsf::EnumPlan p;
    p.n = n; p.stride = stride; p.K = K; p.d = d;
    p.num_gamma = ng; p.num_gamma_total = ng;
    p.gamma = gamma; p.kcheck = kcheck; p.check = check;
    p.binom = bt.c.data(); p.binom_maxN = bt.maxN; p.binom_maxK = bt.maxK;
    p.current_best = current_best;
    p.buffers_key = qsf::next_buffers_key();
    qsf::set_cpu_threads(threads);
    return qsf::cpu_backend()->enumerate(p);
}
