// Host check of the engine's worker pool (pool.hpp): every index of every job runs exactly
// once, for job sizes and worker counts from 0 and 1 up to more workers than items; the
// first exception by index is rethrown after every index has run; two caller threads
// submitting jobs at once are served one job at a time; a child made by fork after the
// pool started runs jobs on threads of its own. Built and run by
// tests/test_cpu_pool.py; prints the first failure and exits 1.
#include <sys/wait.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "pool.hpp"

namespace {
int fails = 0;
void fail(const char* what, long a, long b) {
    if (fails++ < 5) std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
}

void check_counts() {
    for (std::size_t n : {0, 1, 2, 5, 16, 17, 100, 1000})
        for (std::size_t w : {0, 1, 2, 4, 16, 40}) {
            for (int rep = 0; rep < 20; ++rep) {
                std::vector<std::atomic<int>> hits(n);
                for (auto& h : hits) h = 0;
                qsa::HostPool::get().run(n, w, [&](std::size_t i) { hits[i]++; });
                for (std::size_t i = 0; i < n; ++i)
                    if (hits[i] != 1) { fail("index not run exactly once", (long)n, (long)i); return; }
            }
        }
}

void check_exceptions() {
    for (std::size_t w : {1, 3, 16}) {
        std::vector<std::atomic<int>> hits(64);
        for (auto& h : hits) h = 0;
        try {
            qsa::HostPool::get().run(64, w, [&](std::size_t i) {
                hits[i]++;
                if (i == 7 || i == 30) throw std::runtime_error(i == 7 ? "seven" : "thirty");
            });
            fail("no exception", (long)w, 0);
        } catch (const std::runtime_error& e) {
            if (std::string(e.what()) != "seven") fail("not the first exception by index", (long)w, 0);
        }
        for (int i = 0; i < 64; ++i)
            if (hits[i] != 1) fail("index skipped after an exception", (long)w, i);
    }
}

void check_concurrent_callers() {
    std::atomic<long> total{0};
    auto caller = [&](int seed) {
        for (int j = 0; j < 200; ++j)
            qsa::HostPool::get().run(50 + (seed + j) % 30, 16, [&](std::size_t i) { total += (long)i + 1; });
    };
    std::thread a(caller, 0), b(caller, 1);
    a.join();
    b.join();
    long want = 0;
    for (int seed = 0; seed < 2; ++seed)
        for (int j = 0; j < 200; ++j) {
            const long n = 50 + (seed + j) % 30;
            want += n * (n + 1) / 2;
        }
    if (total != want) fail("concurrent callers", total, want);
}

void check_fork() {
    qsa::HostPool::get().run(100, 8, [](std::size_t) {});   // the parent's threads exist
    const pid_t pid = fork();
    if (pid == 0) {
        std::atomic<long> sum{0};
        qsa::HostPool::get().run(1000, 8, [&](std::size_t i) { sum += (long)i; });
        _exit(sum == 999L * 1000 / 2 ? 0 : 3);
    }
    int status = 0;
    waitpid(pid, &status, 0);
    if (!WIFEXITED(status) || WEXITSTATUS(status) != 0) fail("forked child", status, 0);
}
}  // namespace

int main() {
    check_counts();
    check_exceptions();
    check_concurrent_callers();
    check_fork();
    if (fails) {
        std::printf("%d failures\n", fails);
        return 1;
    }
    std::printf("pool ok\n");
    return 0;
}
