// Host-thread helpers for the host-side passes over large DAGs (export,
// import of foreign formats): a thread count, chunked loops and a merge sort.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

namespace svdag {

// Host threads for the loops and sorts (at most 16).
inline uint32_t export_threads()
{
    const uint32_t hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hc ? hc : 1u));
}

// fn(begin, end, thread) over [0, n) in contiguous chunks, one per thread.
template <typename Fn>
inline void parallel_chunks(size_t n, uint32_t threads, Fn fn)
{
    if (n < 65536 || threads <= 1) {
        fn((size_t)0, n, 0u);
        return;
    }
    std::vector<std::thread> pool;
    const size_t step = (n + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
        const size_t b = std::min(n, (size_t)t * step), e = std::min(n, b + step);
        pool.emplace_back([=, &fn] { fn(b, e, t); });
    }
    for (auto& th : pool)
        th.join();
}

// Sorts v by `less`: sorted chunks, then rounds of pairwise merges, each
// round's merges on their own threads.
template <typename T, typename Less>
inline void parallel_sort(std::vector<T>& v, uint32_t threads, Less less)
{
    const size_t n = v.size();
    if (n < 65536 || threads <= 1) {
        std::sort(v.begin(), v.end(), less);
        return;
    }
    const size_t step = (n + threads - 1) / threads;
    std::vector<size_t> bounds;
    for (size_t b = 0; b < n; b += step)
        bounds.push_back(b);
    bounds.push_back(n);
    {
        std::vector<std::thread> pool;
        for (size_t c = 0; c + 1 < bounds.size(); ++c) {
            const size_t a = bounds[c], e = bounds[c + 1];
            pool.emplace_back([&v, a, e, &less] { std::sort(v.begin() + (ptrdiff_t)a, v.begin() + (ptrdiff_t)e, less); });
        }
        for (auto& th : pool)
            th.join();
    }
    while (bounds.size() > 2) {
        std::vector<size_t> next;
        std::vector<std::thread> pool;
        for (size_t c = 0; c + 1 < bounds.size(); c += 2) {
            next.push_back(bounds[c]);
            if (c + 2 < bounds.size()) {
                const size_t a = bounds[c], m = bounds[c + 1], e = bounds[c + 2];
                pool.emplace_back([&v, a, m, e, &less] {
                    std::inplace_merge(v.begin() + (ptrdiff_t)a, v.begin() + (ptrdiff_t)m, v.begin() + (ptrdiff_t)e, less);
                });
            }
        }
        next.push_back(n);
        for (auto& th : pool)
            th.join();
        bounds.swap(next);
    }
}

} // namespace svdag
