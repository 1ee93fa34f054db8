// Internal helpers shared by the host and device sources of libcqmm.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <new>
#include <stdexcept>
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
  // An exception in a worker (host allocation failure) is passed to the
  // caller after every started worker has joined.
  std::vector<std::exception_ptr> err((size_t)nt);
  std::vector<std::thread> th;
  try {
    int64_t k = 0;
    for (int64_t b = 0; b < n; b += chunk, ++k)
      th.emplace_back([f, &err, k, b, e = std::min(n, b + chunk)]() mutable {
        try {
          f(b, e);
        } catch (...) {
          err[(size_t)k] = std::current_exception();
        }
      });
  } catch (...) {
    for (auto &t : th) t.join();
    throw;
  }
  for (auto &t : th) t.join();
  for (auto &x : err)
    if (x) std::rethrow_exception(x);
}

// Body of a C entry point: a C++ exception (host allocation failure, thread
// creation failure) becomes a negative status and a message instead of
// unwinding into the C or Fortran caller.
template <class F>
int guarded(const char *fn, F &&body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return set_errorf(-3, "%s: host allocation failed", fn);
  } catch (const std::length_error &) {
    return set_errorf(-3, "%s: host allocation failed (size)", fn);
  } catch (const std::exception &e) {
    return set_errorf(-5, "%s: %s", fn, e.what());
  } catch (...) {
    return set_errorf(-5, "%s: unexpected exception", fn);
  }
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
