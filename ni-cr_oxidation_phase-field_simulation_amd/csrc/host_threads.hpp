// host_threads.hpp -- host work split over up to 16 std::threads.
//
// parallel_chunks(n, min_chunk, fn) calls fn(begin, end) over contiguous
// chunks of [0, n), one chunk per thread (one chunk and no thread for small n),
// and rethrows the first exception a chunk raised. Callers use it where every
// item is independent, or where each thread owns whole reductions, so the
// results do not depend on the split.
#pragma once

#include <algorithm>
#include <cstddef>
#include <exception>
#include <thread>
#include <vector>

namespace nicr_threads {

inline int host_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(16u, hw)));
}

template <class Fn>
void parallel_chunks(std::size_t n, std::size_t min_chunk, Fn fn) {
    const std::size_t nt = std::min<std::size_t>(host_threads(), std::max<std::size_t>(1, n / min_chunk));
    if (nt <= 1) {
        fn(std::size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> err(nt);
    for (std::size_t t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            try {
                fn(n * t / nt, n * (t + 1) / nt);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    for (auto& th : pool) th.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace nicr_threads
