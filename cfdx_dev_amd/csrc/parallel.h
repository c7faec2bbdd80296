// Minimal static-partition parallel loop over std::thread.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace cfdxhip {

inline uint32_t default_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(hc == 0 ? 1u : hc, 32u));
}

// f(begin, end, thread_index) over contiguous chunks of [0, n).
template <class F>
void parallel_chunks(uint64_t n, uint32_t threads, F&& f) {
    if (threads == 0) threads = default_threads();
    if (n < 4096 || threads == 1) {
        f(uint64_t(0), n, 0u);
        return;
    }
    threads = uint32_t(std::min<uint64_t>(threads, n / 1024 + 1));
    std::vector<std::thread> pool;
    pool.reserve(threads);
    const uint64_t chunk = (n + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
        const uint64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&f, b, e, t]() { f(b, e, t); });
    }
    for (auto& th : pool) th.join();
}

template <class F>
void parallel_for(uint64_t n, uint32_t threads, F&& f) {
    parallel_chunks(n, threads, [&f](uint64_t b, uint64_t e, uint32_t) {
        for (uint64_t i = b; i < e; ++i) f(i);
    });
}

}  // namespace cfdxhip
