// A fixed pool of host threads over an index range: the host loops of adds
// and training (detail::parallel_for, src/parallel.hpp:19-64).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace qh {

// set_num_threads / num_threads (src/threads.cpp): host threads for adds and training.
void set_num_threads(size_t n);
size_t num_threads();

// Runs fn(i) for i in [0, n) on num_threads() threads (detail::parallel_for, src/parallel.hpp:19-64).
template <class F>
void parallel_for(size_t n, F&& fn) {
    const size_t t = std::min(n, num_threads());
    if (t <= 1) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex m;
    auto work = [&] {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n) return;
            try {
                fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!err) err = std::current_exception();
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    for (size_t i = 1; i < t; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

} // namespace qh
