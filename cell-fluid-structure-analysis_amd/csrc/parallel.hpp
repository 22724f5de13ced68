// Host-side data parallelism for the set-up and geometry loops: a range split
// into one contiguous chunk per thread (at most 16 threads).  Every loop that
// uses it writes disjoint outputs per index, so results do not depend on the
// thread count.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace cfsa {

inline int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h, 16u));
}

// f(thread, begin, end) over [0, n); runs inline below min_parallel items.
template <class F>
void parallel_chunks(int64_t n, F &&f, int64_t min_parallel = 4096) {
  const int nt = host_threads();
  if (nt == 1 || n < min_parallel) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t a = std::min(n, t * chunk), b = std::min(n, a + chunk);
    th.emplace_back([&f, t, a, b] { f(t, a, b); });
  }
  for (auto &x : th) x.join();
}

// Sorts v by comp on up to host_threads() threads: contiguous chunks sorted in
// parallel, then merged pairwise.  comp must be a strict total order (no two
// distinct elements compare equal), so the result is the one std::sort gives.
template <class T, class C>
void parallel_sort(std::vector<T> &v, C comp) {
  const int64_t n = (int64_t)v.size();
  const int nt = host_threads();
  if (nt == 1 || n < (int64_t)1 << 16) {
    std::sort(v.begin(), v.end(), comp);
    return;
  }
  std::vector<int64_t> b((size_t)nt + 1);
  for (int t = 0; t <= nt; ++t) b[(size_t)t] = n * t / nt;
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&v, &b, &comp, t] {
        std::sort(v.begin() + b[(size_t)t], v.begin() + b[(size_t)t + 1], comp);
      });
    for (auto &x : th) x.join();
  }
  std::vector<T> tmp(v.size());
  std::vector<T> *src = &v, *dst = &tmp;
  for (int width = 1; width < nt; width *= 2) {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t += 2 * width) {
      const int64_t lo = b[(size_t)t], mid = b[(size_t)std::min(t + width, nt)],
                    hi = b[(size_t)std::min(t + 2 * width, nt)];
      th.emplace_back([src, dst, lo, mid, hi, &comp] {
        std::merge(src->begin() + lo, src->begin() + mid, src->begin() + mid,
                   src->begin() + hi, dst->begin() + lo, comp);
      });
    }
    for (auto &x : th) x.join();
    std::swap(src, dst);
  }
  if (src != &v) v.swap(tmp);
}

}  // namespace cfsa
