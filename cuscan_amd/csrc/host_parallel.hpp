// Host thread helpers for the graph and initialisation steps.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

namespace umaphip {

inline int64_t host_workers() {
    const int64_t hw = std::max<unsigned>(1, std::thread::hardware_concurrency());
    return std::min<int64_t>(hw, 16);
}

// body(i) for i in [0, n), grains handed out dynamically to at most 16
// threads; runs inline when one grain covers everything.
template <class F>
void parallel_for(int64_t n, F &&body, int64_t grain = 1024) {
    const int64_t workers = std::min<int64_t>(host_workers(), (n + grain - 1) / grain);
    if (workers <= 1) {
        for (int64_t i = 0; i < n; ++i) body(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int64_t w = 0; w < workers; ++w)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t start = next.fetch_add(grain);
                if (start >= n) break;
                const int64_t stop = std::min(n, start + grain);
                for (int64_t i = start; i < stop; ++i) body(i);
            }
        });
    for (auto &t : pool) t.join();
}

// body(worker, begin, end) over `workers` contiguous blocks of [0, n), so a
// caller can keep one partial result per worker and reduce them in worker
// order (a fixed, thread-count-dependent summation order).
template <class F>
void parallel_blocks(int64_t n, int64_t workers, F &&body) {
    if (workers <= 1) {
        body(0, 0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (int64_t w = 0; w < workers; ++w) {
        const int64_t b = n * w / workers, e = n * (w + 1) / workers;
        pool.emplace_back([&body, w, b, e] { body(w, b, e); });
    }
    for (auto &t : pool) t.join();
}

}  // namespace umaphip
