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

}  // namespace cfsa
