// Temporary matrices of the density-matrix loops, freed in reverse order of
// allocation (the context enforces free_temp_matrix's LIFO rule,
// src/mult_module.f90:3360-3362). The destructor frees what an early error
// return left, so a failing loop leaves the caller's stack as it found it.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_matrix.h"

namespace cq {

struct Temps {
  cqmm_ctx *ctx;
  std::vector<int32_t> h;
  explicit Temps(cqmm_ctx *c) : ctx(c) {}
  Temps(const Temps &) = delete;
  Temps &operator=(const Temps &) = delete;
  int alloc(int32_t range, int32_t *out) {
    const int st = cqmm_matrix_allocate(ctx, range, out);
    if (!st) h.push_back(*out);
    return st;
  }
  int release() {
    while (!h.empty()) {
      const int st = cqmm_matrix_free(ctx, h.back());
      h.pop_back();
      if (st) return st;
    }
    return CQMM_OK;
  }
  ~Temps() { release(); }
};

}  // namespace cq
