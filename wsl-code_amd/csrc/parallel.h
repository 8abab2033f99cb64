// Host thread pool helper (std::thread, static partition).
#pragma once

#include <algorithm>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace vemd {

template <class F>
void parallel_for(int64_t n, F f) {
    int nt = host_threads();
    if (n <= 0) return;
    if (nt <= 1 || n < 64) {
        f((int64_t)0, n);
        return;
    }
    nt = (int)std::min<int64_t>(nt, n);
    std::vector<std::thread> th;
    std::exception_ptr err;
    std::mutex mu;
    int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&, b, e]() {
            try {
                f(b, e);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                if (!err) err = std::current_exception();
            }
        });
    }
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace vemd
