// Internal layout of a matrix context (cqmm_ctx) and the helpers shared by
// the sources that run density-matrix loops on it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_matrix.h"
#include "cq_device.h"
#include "cq_ops.h"

namespace cq {

constexpr int kMaxRanges = 64;

// Host copy of one range's type(matrix) index plus its device maps.
struct RangeRec {
  bool set = false;
  int64_t length = 0;
  int32_t n_prim = 0;
  std::vector<int32_t> n_nab, ndimi;           // per primary atom
  std::vector<int64_t> onsite;                 // 1-based onsite block offset, 0 if none
  std::vector<int64_t> row_nab, row_off;       // first neighbour / 0-based data offset of each row
  std::vector<int32_t> i_part, i_seq, ndimj;   // per neighbour
  DiagMap diag;
  PairMap pairs;
  bool has_pairs = false;
};

struct MultRec {
  cqmm_plan *plan = nullptr;
  int32_t ra = 0, rb = 0, rc = 0;
};

struct MatRec {
  int32_t range = -1;
  bool live = false;
  DevBuf data;
};

}  // namespace cq

struct cqmm_ctx {
  cq::RangeRec ranges[cq::kMaxRanges];
  std::map<int32_t, cq::MultRec> mults;
  std::vector<cq::MatRec> mats;                         // handle h -> mats[h-1]
  std::map<std::pair<int32_t, int32_t>, std::unique_ptr<cq::RunMap>> summaps;
  std::multimap<size_t, cq::DevBuf> pool;               // freed matrix buffers by size
  cq::DevBuf red_part, red_out;
  std::map<std::string, cq::DevBuf> scratch;            // FP32 operand copies
  int32_t product_precision = 0;                        // 0 FP64, else the FP32 mode of cqmm_matrix_product
  ~cqmm_ctx() {
    for (auto &m : mults) cqmm_plan_destroy(m.second.plan);
  }
};

namespace cq {

int ctx_check(const cqmm_ctx *ctx, int32_t h, const char *who);
inline double *dptr(cqmm_ctx *ctx, int32_t h) { return ctx->mats[h - 1].data.as<double>(); }
inline int32_t hrange(const cqmm_ctx *ctx, int32_t h) { return ctx->mats[h - 1].range; }
inline int64_t hlen(const cqmm_ctx *ctx, int32_t h) { return ctx->ranges[hrange(ctx, h)].length; }
// Device buffer of at least `bytes` kept in the context under `key`.
int get_scratch(cqmm_ctx *ctx, const std::string &key, size_t bytes, void **p);
// Element runs of the blocks common to ranges ra and rb (built on first use).
int summap(cqmm_ctx *ctx, int32_t ra, int32_t rb, const RunMap **out);
// Partial-sum and result buffers of the context's deterministic reductions.
int ensure_red(cqmm_ctx *ctx);

}  // namespace cq
