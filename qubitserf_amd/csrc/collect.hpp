// Device-side sort and deduplication of collected rows (collect.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace qsa {
int sort_dedup_rows(const uint64_t* rows, uint64_t n, int stride, uint64_t* out, uint64_t* kept, hipStream_t stream,
                    std::string& err);
}
