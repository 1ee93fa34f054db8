// Internal helpers shared by the host and device sources of libcqmm.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

namespace cq {

// Records the message returned by cqmm_last_error() and returns `status`.
int set_error(int status, const std::string &msg);
int set_errorf(int status, const char *fmt, ...);

// Counter-based random numbers: a 64-bit mix of (key, counter). The same
// (key, counter) gives the same number whatever order blocks are visited in.
static inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static inline uint64_t hash_combine(uint64_t h, uint64_t v) {
  return mix64(h ^ (v + 0x632be59bd9b4e019ULL + (h << 6) + (h >> 2)));
}
// Uniform in (0,1): 53 random bits, never exactly 0.
static inline double u01(uint64_t key, uint64_t ctr) {
  uint64_t r = mix64(key ^ mix64(ctr));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Host loops of plan construction split over up to 16 threads: f(begin, end)
// runs on disjoint index ranges covering [0, n). Loops shorter than
// serial_below run inline.
template <class F>
void parallel_for(int64_t n, F f, int64_t serial_below = 65536) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t nt = std::min<int64_t>(std::max<int64_t>(1, hw), 16);
  if (n < serial_below || nt == 1) {
    f((int64_t)0, n);
    return;
  }
  const int64_t chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int64_t b = 0; b < n; b += chunk) th.emplace_back(f, b, std::min(n, b + chunk));
  for (auto &t : th) t.join();
}

// Wall-clock phases of plan construction, printed to stderr when the
// environment variable CQMM_PLAN_TIMING is set (host phases include the waits
// for the device work they launch).
struct PhaseTimer {
  bool on;
  std::chrono::steady_clock::time_point t;
  PhaseTimer() : on(enabled()), t(std::chrono::steady_clock::now()) {}
  static bool enabled() {
    static const bool v = std::getenv("CQMM_PLAN_TIMING") != nullptr;
    return v;
  }
  void mark(const char *what) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[cqmm plan] %-28s %9.3f ms\n", what,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
};

}  // namespace cq
