// Host-side data parallelism for the set-up and geometry loops: a range split
// into one contiguous chunk per thread (at most 16 threads).  Every loop that
// uses it writes disjoint outputs per index, so results do not depend on the
// thread count.  An exception thrown on a worker thread (std::bad_alloc, say)
// is rethrown on the calling thread after all workers have joined.
#pragma once

#include <algorithm>
#include <cstdint>
#include <exception>
#include <thread>
#include <vector>

namespace cfsa {

inline int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h, 16u));
}

// Runs the tasks task(0) .. task(count - 1) on one thread each and rethrows
// the first exception any of them raised.  If a thread cannot be started
// (std::system_error), the tasks not yet started run on the calling thread,
// after which every started thread is joined as usual.
template <class F>
void run_threads(int count, F &&task) {
  std::vector<std::exception_ptr> err((size_t)count);
  auto run = [&task, &err](int t) {
    try {
      task(t);
    } catch (...) {
      err[(size_t)t] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  th.reserve((size_t)count);
  int started = 0;
  try {
    for (; started < count; ++started) th.emplace_back(run, started);
  } catch (...) {
    // no thread for task `started` and the rest: run them here
    for (int t = started; t < count; ++t) run(t);
  }
  for (auto &x : th) x.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

// f(thread, begin, end) over [0, n); runs inline below min_parallel items.
template <class F>
void parallel_chunks(int64_t n, F &&f, int64_t min_parallel = 4096) {
  const int nt = host_threads();
  if (nt == 1 || n < min_parallel) {
    f(0, (int64_t)0, n);
    return;
  }
  const int64_t chunk = (n + nt - 1) / nt;
  run_threads(nt, [&](int t) {
    const int64_t a = std::min(n, t * chunk), b = std::min(n, a + chunk);
    f(t, a, b);
  });
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
  run_threads(nt, [&](int t) {
    std::sort(v.begin() + b[(size_t)t], v.begin() + b[(size_t)t + 1], comp);
  });
  std::vector<T> tmp(v.size());
  std::vector<T> *src = &v, *dst = &tmp;
  for (int width = 1; width < nt; width *= 2) {
    const int pairs = (nt + 2 * width - 1) / (2 * width);
    run_threads(pairs, [&, width](int q) {
      const int t = q * 2 * width;
      const int64_t lo = b[(size_t)t], mid = b[(size_t)std::min(t + width, nt)],
                    hi = b[(size_t)std::min(t + 2 * width, nt)];
      std::merge(src->begin() + lo, src->begin() + mid, src->begin() + mid, src->begin() + hi,
                 dst->begin() + lo, comp);
    });
    std::swap(src, dst);
  }
  if (src != &v) v.swap(tmp);
}

}  // namespace cfsa
