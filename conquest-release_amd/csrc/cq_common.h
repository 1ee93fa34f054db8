// Internal helpers shared by the host and device sources of libcqmm.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

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

}  // namespace cq
