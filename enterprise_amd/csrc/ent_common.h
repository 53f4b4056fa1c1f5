// Shared host helpers: thread-local error text, a small parallel-for, and the
// counter-based RNG used by every seeded generator in the library.
#pragma once

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "ent_graph.h"

namespace ent {

// Records the message for ent_last_error() and returns code.
int set_error(int code, const char *fmt, ...);
void clear_error();

int host_threads();

// Runs fn(begin, end) over [0, n) in chunks of `grain` on host_threads()
// threads, handing out chunks dynamically.
template <class F>
void parallel_for(int64_t n, int64_t grain, F &&fn) {
  if (n <= 0) return;
  const int nt = host_threads();
  if (nt <= 1 || n <= grain) {
    fn(int64_t(0), n);
    return;
  }
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      int64_t b = next.fetch_add(grain, std::memory_order_relaxed);
      if (b >= n) break;
      int64_t e = b + grain < n ? b + grain : n;
      fn(b, e);
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  for (int i = 1; i < nt; ++i) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
}

// SplitMix64 (Steele, Lea, Flood 2014): a bijective 64-bit mixer. Used both as
// a sequential stream (state += golden gamma) and as a hash of (seed, index).
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t splitmix_next(uint64_t &state) {
  state += 0x9E3779B97F4A7C15ull;
  return mix64(state);
}
static inline uint64_t hash2(uint64_t seed, uint64_t idx) {
  return mix64(mix64(seed ^ 0x6A09E667F3BCC909ull) + idx * 0x9E3779B97F4A7C15ull);
}

}  // namespace ent
