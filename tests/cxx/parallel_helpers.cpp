// Host-parallel helper checks (csrc/parallel.hpp): a worker's exception comes
// back on the calling thread after every worker has joined, every task runs
// exactly once, and the chunked loop and sort give the sequential results.
#include <atomic>
#include <cstdio>
#include <algorithm>
#include <numeric>
#include <string>
#include <stdexcept>
#include <vector>

#include "parallel.hpp"

int main() {
  using namespace cfsa;
  // every task runs once; an exception in task 3 is rethrown after the join
  std::vector<std::atomic<int>> ran(8);
  for (auto &r : ran) r = 0;
  bool caught = false;
  try {
    run_threads(8, [&](int t) {
      ran[(size_t)t]++;
      if (t == 3) throw std::runtime_error("task 3 failed");
    });
  } catch (const std::runtime_error &e) {
    caught = std::string(e.what()) == "task 3 failed";
  }
  if (!caught) return 1;
  for (auto &r : ran)
    if (r != 1) return 2;
  // chunked loop covers [0, n) once
  const int64_t n = 100000;
  std::vector<int> hit((size_t)n, 0);
  parallel_chunks(n, [&](int, int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) hit[(size_t)i]++;
  });
  for (int h : hit)
    if (h != 1) return 3;
  // a throwing chunk comes back too
  caught = false;
  try {
    parallel_chunks(n, [&](int, int64_t a, int64_t) {
      if (a == 0) throw std::logic_error("first chunk");
    });
  } catch (const std::logic_error &) {
    caught = true;
  }
  if (!caught) return 4;
  // parallel sort equals std::sort
  std::vector<int64_t> v((size_t)n), w;
  for (int64_t i = 0; i < n; ++i) v[(size_t)i] = (i * 7919) % 100003;
  w = v;
  parallel_sort(v, [](int64_t a, int64_t b) { return a < b; });
  std::sort(w.begin(), w.end());
  if (v != w) return 5;
  std::puts("parallel helpers ok");
  return 0;
}
