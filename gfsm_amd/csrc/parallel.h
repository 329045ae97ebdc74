// parallel.h -- a minimal static-partition parallel_for over std::thread.
#pragma once

#include <algorithm>
#include <thread>
#include <vector>

namespace gfsm {

inline unsigned hw_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    return n ? std::min(n, 64u) : 1u;
}

// f(begin, end) on contiguous blocks of [0, n).
template <class F>
void parallel_for(uint64_t n, F &&f, unsigned threads)
{
    if (!threads) threads = hw_threads();
    if (n < 4096 || threads == 1) {
        if (n) f(uint64_t(0), n);
        return;
    }
    threads = (unsigned)std::min<uint64_t>(threads, n);
    std::vector<std::thread> ts;
    ts.reserve(threads);
    for (unsigned t = 0; t < threads; t++) {
        uint64_t b = n * t / threads, e = n * (t + 1) / threads;
        ts.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto &t : ts) t.join();
}

}  // namespace gfsm
