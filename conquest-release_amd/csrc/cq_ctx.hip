// Device-resident matrix API (mult_module mirror) and the density-matrix loops
// that drive it: McWeeny purification (src/McWeeny.f90), its initial guess
// InitMcW, and the Hotelling iteration for S^-1 (src/S_matrix_module.f90).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_matrix.h"
#include "cq_common.h"
#include "cq_device.h"
#include "cq_ops.h"
#include "cq_temps.h"
#include "cq_ctx.h"

namespace cq {

int ctx_check(const cqmm_ctx *ctx, int32_t h, const char *who) {
  if (!ctx) return set_errorf(CQMM_ERR_ARG, "%s: null context", who);
  if (h < 1 || h > (int32_t)ctx->mats.size() || !ctx->mats[h - 1].live)
    return set_errorf(CQMM_ERR_ARG, "%s: invalid matrix handle %d", who, h);
  return CQMM_OK;
}

int get_scratch(cqmm_ctx *ctx, const std::string &key, size_t bytes, void **p) {
  DevBuf &b = ctx->scratch[key];
  if (b.bytes < bytes) {
    int st = b.alloc(bytes);
    if (st) return st;
  }
  *p = b.p;
  return CQMM_OK;
}

// addscan's merge (src/multiply_module.f90:849-943) of two neighbour lists sorted
// by (i_part, i_seq): element runs of the blocks common to both patterns.
static int build_summap(const RangeRec &A, const RangeRec &B, cq::RunMap &m) {
  if (A.n_prim != B.n_prim) return cq::set_error(CQMM_ERR_ARG, "matrix_sum: ranges with different primary sets");
  std::vector<int64_t> dst, src;
  std::vector<int32_t> len;
  for (int32_t p = 0; p < A.n_prim; ++p) {
    int64_t ia = A.row_nab[p], ea = ia + A.n_nab[p];
    int64_t ib = B.row_nab[p], eb = ib + B.n_nab[p];
    int64_t oa = A.row_off[p], ob = B.row_off[p];
    const int32_t nd1 = A.ndimi[p];
    while (ia < ea && ib < eb) {
      const int64_t ka = ((int64_t)A.i_part[ia] << 32) | (uint32_t)A.i_seq[ia];
      const int64_t kb = ((int64_t)B.i_part[ib] << 32) | (uint32_t)B.i_seq[ib];
      if (ka < kb) {
        oa += (int64_t)nd1 * A.ndimj[ia++];
      } else if (kb < ka) {
        ob += (int64_t)nd1 * B.ndimj[ib++];
      } else {
        const int32_t n = nd1 * A.ndimj[ia];
        if (!len.empty() && dst.back() + len.back() == oa && src.back() + len.back() == ob && len.back() + n < (1 << 20)) {
          len.back() += n;
        } else {
          dst.push_back(oa);
          src.push_back(ob);
          len.push_back(n);
        }
        oa += n;
        ob += (int64_t)nd1 * B.ndimj[ib];
        ++ia;
        ++ib;
      }
    }
  }
  hipStream_t s = cq::stream();
  m.nruns = (int64_t)dst.size();
  int64_t ne = 0;
  for (int32_t l : len) ne += l;
  m.nelem = ne;
  int st;
  if ((st = m.dst.alloc(std::max<size_t>(1, dst.size()) * 8))) return st;
  if ((st = m.src.alloc(std::max<size_t>(1, src.size()) * 8))) return st;
  if ((st = m.len.alloc(std::max<size_t>(1, len.size()) * 4))) return st;
  if (m.nruns) {
    if ((st = cq::hip_check(hipMemcpyAsync(m.dst.p, dst.data(), dst.size() * 8, hipMemcpyHostToDevice, s), "H2D"))) return st;
    if ((st = cq::hip_check(hipMemcpyAsync(m.src.p, src.data(), src.size() * 8, hipMemcpyHostToDevice, s), "H2D"))) return st;
    if ((st = cq::hip_check(hipMemcpyAsync(m.len.p, len.data(), len.size() * 4, hipMemcpyHostToDevice, s), "H2D"))) return st;
  }
  return cq::hip_check(hipStreamSynchronize(s), "summap upload");
}

int summap(cqmm_ctx *ctx, int32_t ra, int32_t rb, const cq::RunMap **out) {
  auto key = std::make_pair(ra, rb);
  auto it = ctx->summaps.find(key);
  if (it == ctx->summaps.end()) {
    std::unique_ptr<cq::RunMap> m(new cq::RunMap());
    int st = build_summap(ctx->ranges[ra], ctx->ranges[rb], *m);
    if (st) return st;
    it = ctx->summaps.emplace(key, std::move(m)).first;
  }
  *out = it->second.get();
  return CQMM_OK;
}

static int reduce_result(cqmm_ctx *ctx, double *result) {
  hipStream_t s = cq::stream();
  int st = cq::hip_check(hipMemcpyAsync(result, ctx->red_out.p, 8, hipMemcpyDeviceToHost, s), "D2H reduction");
  if (st) return st;
  return cq::hip_check(hipStreamSynchronize(s), "reduction");
}

int ensure_red(cqmm_ctx *ctx) {
  int st;
  if (!ctx->red_part.p && (st = ctx->red_part.alloc((size_t)cq::reduction_partials() * 8))) return st;
  if (!ctx->red_out.p && (st = ctx->red_out.alloc(64))) return st;
  return CQMM_OK;
}


}  // namespace cq

using namespace cq;

extern "C" {

int cqmm_ctx_create(cqmm_ctx **ctx) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx) return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_create: null argument");
    int st = cq::ensure_init();
    if (st) return st;
    *ctx = new cqmm_ctx();
    return CQMM_OK;
  });
}

int cqmm_ctx_destroy(cqmm_ctx *ctx) {
  return cq::guarded(__func__, [&]() -> int {
    if (ctx) {
      (void)hipStreamSynchronize(cq::stream());
      delete ctx;
    }
    return CQMM_OK;
  });
}

int cqmm_ctx_add_range(cqmm_ctx *ctx, int32_t range, const cqmm_matrix_index *idx) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !idx || range < 0 || range >= kMaxRanges) return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_add_range: bad argument");
    if (idx->n_parts <= 0 || idx->n_prim <= 0 || !idx->n_atoms || !idx->nab_offset || !idx->nd_offset || !idx->n_nab ||
        !idx->i_acc || !idx->i_nd_acc || !idx->ndimi || !idx->i_part || !idx->i_seq || !idx->ndimj || idx->length <= 0)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_add_range: incomplete matrix index");
    if (ctx->ranges[range].set) {
      // A live matrix or a registered product sized by the old index would
      // read or write past its buffers under the new one.
      for (const MatRec &m : ctx->mats)
        if (m.live && m.range == range)
          return cq::set_errorf(CQMM_ERR_STATE, "cqmm_ctx_add_range: range %d has live matrices", range);
      for (const auto &kv : ctx->mults)
        if (kv.second.ra == range || kv.second.rb == range || kv.second.rc == range)
          return cq::set_errorf(CQMM_ERR_STATE, "cqmm_ctx_add_range: range %d is used by mult %d", range, kv.first);
    }
    RangeRec r;
    r.n_prim = idx->n_prim;
    r.length = idx->length;
    r.n_nab.assign(idx->n_nab, idx->n_nab + idx->n_prim);
    r.ndimi.assign(idx->ndimi, idx->ndimi + idx->n_prim);
    r.onsite.assign(idx->n_prim, 0);
    if (idx->onsite) r.onsite.assign(idx->onsite, idx->onsite + idx->n_prim);
    r.i_part.assign(idx->i_part, idx->i_part + idx->n_nabs);
    r.i_seq.assign(idx->i_seq, idx->i_seq + idx->n_nabs);
    r.ndimj.assign(idx->ndimj, idx->ndimj + idx->n_nabs);
    r.row_nab.resize(idx->n_prim);
    r.row_off.resize(idx->n_prim);
    int32_t p = 0;
    for (int32_t nn = 0; nn < idx->n_parts; ++nn)
      for (int32_t a = 0; a < idx->n_atoms[nn]; ++a, ++p) {
        if (p >= idx->n_prim) return cq::set_error(CQMM_ERR_RANGE, "cqmm_ctx_add_range: n_atoms exceed n_prim");
        r.row_nab[p] = (int64_t)idx->nab_offset[nn] + idx->i_acc[p] - 1;
        r.row_off[p] = idx->nd_offset[nn] + idx->i_nd_acc[p] - 1;
        if (r.row_nab[p] < 0 || r.row_nab[p] + r.n_nab[p] > idx->n_nabs)
          return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_ctx_add_range: row %d outside the neighbour arrays", p + 1);
        int64_t sz = 0;
        for (int32_t e = 0; e < r.n_nab[p]; ++e) sz += (int64_t)r.ndimi[p] * r.ndimj[r.row_nab[p] + e];
        if (r.row_off[p] < 0 || r.row_off[p] + sz > r.length)
          return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_ctx_add_range: row %d data outside length", p + 1);
      }
    if (p != idx->n_prim) return cq::set_error(CQMM_ERR_RANGE, "cqmm_ctx_add_range: n_atoms do not sum to n_prim");
    // Onsite blocks for matrix_trace.
    std::vector<int64_t> doff;
    std::vector<int32_t> dn;
    for (int32_t q = 0; q < r.n_prim; ++q)
      if (r.onsite[q] > 0) {
        if (r.onsite[q] - 1 + (int64_t)r.ndimi[q] * r.ndimi[q] > r.length)
          return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_ctx_add_range: onsite(%d) outside length", q + 1);
        doff.push_back(r.onsite[q] - 1);
        dn.push_back(r.ndimi[q]);
      }
    int st;
    hipStream_t s = cq::stream();
    r.diag.nblk = (int64_t)doff.size();
    if ((st = r.diag.off.alloc(std::max<size_t>(1, doff.size()) * 8))) return st;
    if ((st = r.diag.nsf.alloc(std::max<size_t>(1, dn.size()) * 4))) return st;
    if (!doff.empty()) {
      if ((st = cq::hip_check(hipMemcpyAsync(r.diag.off.p, doff.data(), doff.size() * 8, hipMemcpyHostToDevice, s), "H2D"))) return st;
      if ((st = cq::hip_check(hipMemcpyAsync(r.diag.nsf.p, dn.data(), dn.size() * 4, hipMemcpyHostToDevice, s), "H2D"))) return st;
    }
    if ((st = cq::hip_check(hipStreamSynchronize(s), "range upload"))) return st;
    r.set = true;
    // Replacing a range invalidates its cached sum maps.
    for (auto it = ctx->summaps.begin(); it != ctx->summaps.end();)
      it = (it->first.first == range || it->first.second == range) ? ctx->summaps.erase(it) : std::next(it);
    ctx->ranges[range] = std::move(r);
    return CQMM_OK;
  });
}

int cqmm_ctx_set_transpose(cqmm_ctx *ctx, int32_t range, int64_t npairs, const int64_t *p, const int64_t *q,
                           const int32_t *n1, const int32_t *n2) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || range < 0 || range >= kMaxRanges || !ctx->ranges[range].set || npairs < 0 || (npairs && (!p || !q || !n1 || !n2)))
      return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_set_transpose: bad argument");
    RangeRec &r = ctx->ranges[range];
    std::vector<int64_t> pp(npairs), qq(npairs);
    std::vector<uint8_t> a(npairs), b(npairs);
    for (int64_t t = 0; t < npairs; ++t) {
      if (p[t] < 1 || q[t] < 1 || n1[t] < 1 || n2[t] < 1 || n1[t] > 64 || n2[t] > 64 ||
          p[t] - 1 + (int64_t)n1[t] * n2[t] > r.length || q[t] - 1 + (int64_t)n1[t] * n2[t] > r.length)
        return cq::set_errorf(CQMM_ERR_RANGE, "cqmm_ctx_set_transpose: pair %lld outside the matrix", (long long)t + 1);
      if (p[t] == q[t] && n1[t] != n2[t]) return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_set_transpose: non-square self pair");
      pp[t] = p[t] - 1;
      qq[t] = q[t] - 1;
      a[t] = (uint8_t)n1[t];
      b[t] = (uint8_t)n2[t];
    }
    int st;
    hipStream_t s = cq::stream();
    r.pairs.npairs = npairs;
    if ((st = r.pairs.poff.alloc(std::max<int64_t>(1, npairs) * 8))) return st;
    if ((st = r.pairs.qoff.alloc(std::max<int64_t>(1, npairs) * 8))) return st;
    if ((st = r.pairs.n1.alloc(std::max<int64_t>(1, npairs)))) return st;
    if ((st = r.pairs.n2.alloc(std::max<int64_t>(1, npairs)))) return st;
    if (npairs) {
      if ((st = cq::hip_check(hipMemcpyAsync(r.pairs.poff.p, pp.data(), npairs * 8, hipMemcpyHostToDevice, s), "H2D"))) return st;
      if ((st = cq::hip_check(hipMemcpyAsync(r.pairs.qoff.p, qq.data(), npairs * 8, hipMemcpyHostToDevice, s), "H2D"))) return st;
      if ((st = cq::hip_check(hipMemcpyAsync(r.pairs.n1.p, a.data(), npairs, hipMemcpyHostToDevice, s), "H2D"))) return st;
      if ((st = cq::hip_check(hipMemcpyAsync(r.pairs.n2.p, b.data(), npairs, hipMemcpyHostToDevice, s), "H2D"))) return st;
    }
    r.has_pairs = true;
    return cq::hip_check(hipStreamSynchronize(s), "transpose upload");
  });
}

int cqmm_ctx_add_mult(cqmm_ctx *ctx, int32_t mult_id, int32_t ra, int32_t rb, int32_t rc, cqmm_plan *plan) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !plan) return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_add_mult: null argument");
    for (int32_t r : {ra, rb, rc})
      if (r < 0 || r >= kMaxRanges || !ctx->ranges[r].set)
        return cq::set_errorf(CQMM_ERR_ARG, "cqmm_ctx_add_mult: range %d not registered", r);
    if (plan->lena != ctx->ranges[ra].length || plan->lenb != ctx->ranges[rb].length || plan->lenc != ctx->ranges[rc].length)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_add_mult: plan lengths differ from the ranges' lengths");
    auto it = ctx->mults.find(mult_id);
    if (it != ctx->mults.end()) cqmm_plan_destroy(it->second.plan);
    MultRec m;
    m.plan = plan;
    m.ra = ra;
    m.rb = rb;
    m.rc = rc;
    ctx->mults[mult_id] = m;
    return CQMM_OK;
  });
}

int cqmm_matrix_allocate(cqmm_ctx *ctx, int32_t range, int32_t *handle) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !handle || range < 0 || range >= kMaxRanges || !ctx->ranges[range].set)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_matrix_allocate: range %d not registered", range);
    const size_t bytes = (size_t)ctx->ranges[range].length * 8;
    MatRec m;
    m.range = range;
    auto it = ctx->pool.find(bytes);
    if (it != ctx->pool.end()) {
      m.data = std::move(it->second);
      ctx->pool.erase(it);
    } else {
      int st = m.data.alloc(bytes);
      if (st) return st;
    }
    int st = cq::hip_check(hipMemsetAsync(m.data.p, 0, bytes, cq::stream()), "zero matrix");
    if (st) return st;
    m.live = true;
    ctx->mats.push_back(std::move(m));
    *handle = (int32_t)ctx->mats.size();
    return CQMM_OK;
  });
}

int cqmm_matrix_free(cqmm_ctx *ctx, int32_t handle) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, handle, "cqmm_matrix_free");
    if (st) return st;
    if (handle != (int32_t)ctx->mats.size())
      return cq::set_errorf(CQMM_ERR_STATE, "cqmm_matrix_free: handle %d freed out of order (top is %d)", handle,
                            (int)ctx->mats.size());
    MatRec &m = ctx->mats.back();
    const size_t bytes = m.data.bytes;
    ctx->pool.emplace(bytes, std::move(m.data));
    ctx->mats.pop_back();
    return CQMM_OK;
  });
}

int cqmm_matrix_length(const cqmm_ctx *ctx, int32_t h, int64_t *len) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, h, "cqmm_matrix_length");
    if (st) return st;
    *len = hlen(ctx, h);
    return CQMM_OK;
  });
}

int cqmm_matrix_range(const cqmm_ctx *ctx, int32_t h, int32_t *range) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, h, "cqmm_matrix_range");
    if (st) return st;
    *range = hrange(ctx, h);
    return CQMM_OK;
  });
}

int cqmm_matrix_device_ptr(const cqmm_ctx *ctx, int32_t h, double **ptr) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, h, "cqmm_matrix_device_ptr");
    if (st) return st;
    *ptr = ctx->mats[h - 1].data.as<double>();
    return CQMM_OK;
  });
}

int cqmm_matrix_set(cqmm_ctx *ctx, int32_t h, const double *host) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, h, "cqmm_matrix_set");
    if (st) return st;
    hipStream_t s = cq::stream();
    if ((st = cq::hip_check(hipMemcpyAsync(dptr(ctx, h), host, hlen(ctx, h) * 8, hipMemcpyHostToDevice, s), "H2D matrix"))) return st;
    return cq::hip_check(hipStreamSynchronize(s), "cqmm_matrix_set");
  });
}

int cqmm_matrix_get(cqmm_ctx *ctx, int32_t h, double *host) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, h, "cqmm_matrix_get");
    if (st) return st;
    hipStream_t s = cq::stream();
    if ((st = cq::hip_check(hipMemcpyAsync(host, dptr(ctx, h), hlen(ctx, h) * 8, hipMemcpyDeviceToHost, s), "D2H matrix"))) return st;
    return cq::hip_check(hipStreamSynchronize(s), "cqmm_matrix_get");
  });
}

int cqmm_matrix_product(cqmm_ctx *ctx, int32_t A, int32_t B, int32_t C, int32_t mult_id) {
  return cq::guarded(__func__, [&]() -> int {
    int st;
    if ((st = ctx_check(ctx, A, "cqmm_matrix_product")) || (st = ctx_check(ctx, B, "cqmm_matrix_product")) ||
        (st = ctx_check(ctx, C, "cqmm_matrix_product")))
      return st;
    auto it = ctx->mults.find(mult_id);
    if (it == ctx->mults.end()) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_matrix_product: mult %d not registered", mult_id);
    const MultRec &m = it->second;
    const cq::TripletPlan &tp = m.plan->tp;
    // matrix_product: type 1 -> mat_mult(A,B,C); type 2 -> mat_mult(C,B,A).
    const int32_t a_role = tp.mult_type == 1 ? A : C, c_role = tp.mult_type == 1 ? C : A;
    if (hrange(ctx, a_role) != m.ra || hrange(ctx, B) != m.rb || hrange(ctx, c_role) != m.rc)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_matrix_product: operand ranges (%d,%d,%d) do not match mult %d",
                            hrange(ctx, A), hrange(ctx, B), hrange(ctx, C), mult_id);
    hipStream_t s = cq::stream();
    if (ctx->product_precision == 0) return cq::launch_block_products(tp, dptr(ctx, A), dptr(ctx, B), dptr(ctx, C), false, s);
    // FP32 products (cqmm_ctx_set_product_precision): FP32 copies of both
    // operands, the FP32 kernels of that mode, the result widened to FP64.
    // The kernels take the left operand from A and write C for both types
    // (type 2: A = C.B^T in mat_mult's naming is P = A, O = C here), as the
    // FP64 launch above.
    if (hlen(ctx, A) < tp.len_p || hlen(ctx, B) < tp.len_q || hlen(ctx, C) < tp.len_out)
      return cq::set_errorf(CQMM_ERR_STATE, "cqmm_matrix_product: operand lengths below mult %d's", mult_id);
    void *fp, *fq, *fo;
    if ((st = get_scratch(ctx, "prodP", tp.len_p * 4, &fp)) || (st = get_scratch(ctx, "prodQ", tp.len_q * 4, &fq)) ||
        (st = get_scratch(ctx, "prodO", tp.len_out * 4, &fo)))
      return st;
    // Split-BF16 x3 where the plan's kernel has it, else FP32 MFMA with FP64 sums.
    const int mode = (ctx->product_precision == 4 && !(cq::uses_tiles(tp) && cq::tiles_take_bf16x3(tp.tiles)))
                         ? 2 : ctx->product_precision;
    if (mode == 4) {
      if ((st = cq::op_to_bf16x2(tp.len_p, dptr(ctx, A), (uint32_t *)fp, s))) return st;
      if ((st = cq::op_to_bf16x2(tp.len_q, dptr(ctx, B), (uint32_t *)fq, s))) return st;
    } else {
      if ((st = cq::op_to_f32(tp.len_p, dptr(ctx, A), (float *)fp, s))) return st;
      if ((st = cq::op_to_f32(tp.len_q, dptr(ctx, B), (float *)fq, s))) return st;
    }
    if ((st = cq::launch_block_products_f32(tp, (const float *)fp, (const float *)fq, (float *)fo, mode, s)))
      return st;
    return cq::op_to_f64(tp.len_out, (const float *)fo, dptr(ctx, C), s);
  });
}

int cqmm_ctx_set_product_precision(cqmm_ctx *ctx, int32_t precision) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx) return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_set_product_precision: null context");
    if (precision < 0 || precision > 4)
      return cq::set_errorf(CQMM_ERR_ARG, "cqmm_ctx_set_product_precision: precision %d (0..4)", precision);
    ctx->product_precision = precision;
    return CQMM_OK;
  });
}

int cqmm_ctx_get_product_precision(const cqmm_ctx *ctx, int32_t *precision) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !precision) return cq::set_error(CQMM_ERR_ARG, "cqmm_ctx_get_product_precision: null argument");
    *precision = ctx->product_precision;
    return CQMM_OK;
  });
}

int cqmm_matrix_sum(cqmm_ctx *ctx, double alpha, int32_t A, double beta, int32_t B) {
  return cq::guarded(__func__, [&]() -> int {
    int st;
    if ((st = ctx_check(ctx, A, "cqmm_matrix_sum")) || (st = ctx_check(ctx, B, "cqmm_matrix_sum"))) return st;
    hipStream_t s = cq::stream();
    const int32_t ra = hrange(ctx, A), rb = hrange(ctx, B);
    if (ra == rb) return cq::op_axpby(hlen(ctx, A), alpha, dptr(ctx, A), beta, dptr(ctx, B), s);
    const cq::RunMap *m = nullptr;
    if ((st = summap(ctx, ra, rb, &m))) return st;
    return cq::op_runs_axpby(*m, alpha, dptr(ctx, A), beta, dptr(ctx, B), s);
  });
}

int cqmm_matrix_scale(cqmm_ctx *ctx, double alpha, int32_t A) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, A, "cqmm_matrix_scale");
    if (st) return st;
    return cq::op_scale(hlen(ctx, A), alpha, dptr(ctx, A), cq::stream());
  });
}

int cqmm_matrix_product_trace(cqmm_ctx *ctx, int32_t A, int32_t B, double *result) {
  return cq::guarded(__func__, [&]() -> int {
    int st;
    if ((st = ctx_check(ctx, A, "cqmm_matrix_product_trace")) || (st = ctx_check(ctx, B, "cqmm_matrix_product_trace"))) return st;
    if (!result) return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_product_trace: null result");
    if (hrange(ctx, A) != hrange(ctx, B))
      return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_product_trace: matrices of different ranges");
    if ((st = ensure_red(ctx))) return st;
    if ((st = cq::op_dot(hlen(ctx, A), dptr(ctx, A), dptr(ctx, B), ctx->red_part.as<double>(), ctx->red_out.as<double>(),
                         cq::stream())))
      return st;
    return reduce_result(ctx, result);
  });
}

int cqmm_matrix_trace(cqmm_ctx *ctx, int32_t A, double *result) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, A, "cqmm_matrix_trace");
    if (st) return st;
    if (!result) return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_trace: null result");
    if ((st = ensure_red(ctx))) return st;
    const RangeRec &r = ctx->ranges[hrange(ctx, A)];
    if ((st = cq::op_trace(r.diag, dptr(ctx, A), ctx->red_part.as<double>(), ctx->red_out.as<double>(), cq::stream())))
      return st;
    return reduce_result(ctx, result);
  });
}

int cqmm_matrix_symmetrise(cqmm_ctx *ctx, int32_t A) {
  return cq::guarded(__func__, [&]() -> int {
    int st = ctx_check(ctx, A, "cqmm_matrix_symmetrise");
    if (st) return st;
    const RangeRec &r = ctx->ranges[hrange(ctx, A)];
    if (!r.has_pairs) return cq::set_errorf(CQMM_ERR_STATE, "cqmm_matrix_symmetrise: range %d has no transpose pairs", hrange(ctx, A));
    return cq::op_symmetrise(r.pairs, dptr(ctx, A), cq::stream());
  });
}

int cqmm_matrix_transpose(cqmm_ctx *ctx, int32_t A, int32_t B) {
  return cq::guarded(__func__, [&]() -> int {
    int st;
    if ((st = ctx_check(ctx, A, "cqmm_matrix_transpose")) || (st = ctx_check(ctx, B, "cqmm_matrix_transpose"))) return st;
    if (A == B) return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_transpose: A and B must differ");
    if (hrange(ctx, A) != hrange(ctx, B))
      return cq::set_error(CQMM_ERR_ARG, "cqmm_matrix_transpose: A and B must share a range");
    const RangeRec &r = ctx->ranges[hrange(ctx, A)];
    if (!r.has_pairs) return cq::set_errorf(CQMM_ERR_STATE, "cqmm_matrix_transpose: range %d has no transpose pairs", hrange(ctx, A));
    return cq::op_transpose(r.pairs, dptr(ctx, A), dptr(ctx, B), cq::stream());
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Density-matrix loops.
// ---------------------------------------------------------------------------
namespace {

#define CQ_TRY(x)            \
  do {                       \
    int st_ = (x);           \
    if (st_) return st_;     \
  } while (0)

// Checks of InitMcW's handles against its setup.
int initmcw_check(cqmm_ctx *ctx, const cqmm_initmcw_setup *su, int32_t matH, int32_t matS, int32_t matT,
                  int32_t matL, const char *who) {
  CQ_TRY(ctx_check(ctx, matH, who));
  CQ_TRY(ctx_check(ctx, matS, who));
  CQ_TRY(ctx_check(ctx, matT, who));
  CQ_TRY(ctx_check(ctx, matL, who));
  for (int32_t r : {su->range_T, su->range_S, su->range_H, su->range_L, su->range_TS, su->range_TH})
    if (r < 0 || r >= kMaxRanges || !ctx->ranges[r].set) return cq::set_errorf(CQMM_ERR_ARG, "%s: range %d not registered", who, r);
  if (hrange(ctx, matH) != su->range_H || hrange(ctx, matS) != su->range_S || hrange(ctx, matT) != su->range_T ||
      hrange(ctx, matL) != su->range_L)
    return cq::set_errorf(CQMM_ERR_ARG, "%s: handle ranges differ from the setup", who);
  return CQMM_OK;
}

// n_o: support functions of the cell (the row dimensions of the H range).
double initmcw_orbitals(const cqmm_ctx *ctx, const cqmm_initmcw_setup *su) {
  const RangeRec &rh = ctx->ranges[su->range_H];
  double n_o = 0.0;
  for (int32_t p = 0; p < rh.n_prim; ++p) n_o += rh.ndimi[p];
  return n_o;
}

// What InitMcW needs of one spin (src/McWeeny.f90:413-452): SX = Tr[T S];
// McWXHX's XHX = T H T at L range (a temporary of `tmp`, left allocated) and
// SXHX = Tr[S (THT)_S]; GetHLimits' hmin, hmax (:1020-1109) from a host copy of H.
struct InitMcwParts {
  double SX = 0.0, SXHX = 0.0, hmin = 1e30, hmax = -1e30;
  int32_t XHX = 0;
};

int initmcw_parts(cqmm_ctx *ctx, const cqmm_initmcw_setup *su, int32_t matH, int32_t matS, int32_t matT,
                  cq::Temps &tmp, InitMcwParts *q) {
  int32_t matTS, matTH, matTmp;
  CQ_TRY(tmp.alloc(su->range_TS, &matTS));
  CQ_TRY(cqmm_matrix_product(ctx, matT, matS, matTS, su->mult_T_S_TS));
  CQ_TRY(cqmm_matrix_trace(ctx, matTS, &q->SX));
  CQ_TRY(tmp.alloc(su->range_TH, &matTH));
  CQ_TRY(tmp.alloc(su->range_L, &q->XHX));
  CQ_TRY(tmp.alloc(su->range_S, &matTmp));
  CQ_TRY(cqmm_matrix_product(ctx, matT, matH, matTH, su->mult_T_H_TH));
  CQ_TRY(cqmm_matrix_product(ctx, matTH, matT, q->XHX, su->mult_TH_T_L));
  CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matTmp, 1.0, q->XHX));
  CQ_TRY(cqmm_matrix_product_trace(ctx, matS, matTmp, &q->SXHX));
  const RangeRec &rh = ctx->ranges[su->range_H];
  std::vector<double> H(rh.length);
  CQ_TRY(cqmm_matrix_get(ctx, matH, H.data()));
  for (int32_t p = 0; p < rh.n_prim; ++p) {
    const int32_t nsfi = rh.ndimi[p];
    std::vector<double> lo(64, 0.0), hi(64, 0.0);
    int64_t off = rh.row_off[p];
    for (int32_t nb = 0; nb < rh.n_nab[p]; ++nb) {
      const int32_t nsfj = rh.ndimj[rh.row_nab[p] + nb];
      const bool onsite = rh.onsite[p] > 0 && off == rh.onsite[p] - 1;
      for (int32_t i = 0; i < nsfj; ++i)
        for (int32_t j = 0; j < nsfi; ++j) {
          const double v = std::fabs(H[off + j + (int64_t)nsfi * i]);
          lo[i] += onsite ? v : -v;
          hi[i] += v;
        }
      off += (int64_t)nsfi * nsfj;
    }
    for (int32_t i = 0; i < nsfi; ++i) {
      q->hmin = std::min(q->hmin, lo[i]);
      q->hmax = std::max(q->hmax, hi[i]);
    }
  }
  return CQMM_OK;
}

// InitMcW's choice of mubar and lambda (src/McWeeny.f90:476-507 per spin,
// :522-545 on the direct sum of both spins, whose second branch tests `<`
// where the per-spin one tests `>`); n_o == n_e gives mubar = lambda = 0.
int initmcw_mubar(double n_o, double ne, double SXHX, double hmin, double hmax, double A, bool direct_sum,
                  double *mubar, double *lambda) {
  *mubar = 0.0;
  *lambda = 0.0;
  if (n_o == ne) return CQMM_OK;
  const double mu1 = SXHX / n_o + hmax * A;
  const double mu2 = (hmin * A * ne / (n_o - ne) - SXHX / n_o) / (n_o * A / (n_o - ne) - 1.0);
  const double lhs2 = ne / (hmax - mu2), rhs2 = (n_o - ne) / (mu2 - hmin);
  if (ne / (hmax - mu1) < (n_o - ne) / (mu1 - hmin)) {
    *mubar = mu1;
    *lambda = ne / (hmax - mu1);
  } else if (direct_sum ? lhs2 < rhs2 : lhs2 > rhs2) {
    *mubar = mu2;
    *lambda = (n_o - ne) / (mu2 - hmin);
  } else {
    return cq::set_error(CQMM_ERR_NOTCONV, "InitMcW: cannot find mubar");
  }
  return CQMM_OK;
}

}  // namespace

extern "C" {

int cqmm_hotelling(cqmm_ctx *ctx, const cqmm_hotelling_setup *su, int32_t matS, int32_t matT, int32_t *iters,
                   double *omega_out) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !su) return cq::set_error(CQMM_ERR_ARG, "cqmm_hotelling: null argument");
    int st;
    if ((st = ctx_check(ctx, matS, "cqmm_hotelling")) || (st = ctx_check(ctx, matT, "cqmm_hotelling"))) return st;
    if (hrange(ctx, matS) != su->range_S || hrange(ctx, matT) != su->range_T)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_hotelling: handle ranges differ from the setup");
    hipStream_t s = cq::stream();
    cq::Temps tmp(ctx);
    int32_t matI, matT1, matTold, matA, matTS, matGrad;
    CQ_TRY(tmp.alloc(su->range_TS, &matI));
    CQ_TRY(tmp.alloc(su->range_T, &matT1));
    CQ_TRY(tmp.alloc(su->range_T, &matTold));
    CQ_TRY(tmp.alloc(su->range_TS, &matA));
    CQ_TRY(tmp.alloc(su->range_TS, &matTS));
    CQ_TRY(tmp.alloc(su->range_T, &matGrad));
    CQ_TRY(cq::op_set_diag(ctx->ranges[su->range_TS].diag, 1.0, dptr(ctx, matI), s));
    // T0 = eps S with eps = 1 / sum S_jk^2 (Iter_Hott_InvS, src/S_matrix_module.f90:846-870).
    double tot = 0.0;
    CQ_TRY(cqmm_matrix_product_trace(ctx, matS, matS, &tot));
    CQ_TRY(cqmm_matrix_scale(ctx, 0.0, matT));
    CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matT, 1.0 / tot, matS));
    double deltaomega = 0.0, oldomega = 0.0, omega = 0.0;
    int n = 0;
    const int max_iter = su->max_iter > 0 ? su->max_iter : 50;
    for (n = 1; n <= max_iter; ++n) {
      deltaomega *= 0.5;
      if (!(n < 3 || std::fabs(deltaomega) > su->tol)) break;
      // HotInvS_mm (src/S_matrix_module.f90:1005-1034): TS = T.S, A = I - TS,
      // Grad = TS.T, T1 = 2T - Grad, omega = Tr[A A].
      CQ_TRY(cqmm_matrix_product(ctx, matT, matS, matTS, su->mult_T_S_TS));
      CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matA, 1.0, matI));
      CQ_TRY(cqmm_matrix_sum(ctx, 1.0, matA, -1.0, matTS));
      CQ_TRY(cqmm_matrix_product(ctx, matTS, matT, matGrad, su->mult_TS_T_T));
      CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matT1, 2.0, matT));
      CQ_TRY(cqmm_matrix_sum(ctx, 1.0, matT1, -1.0, matGrad));
      CQ_TRY(cqmm_matrix_product_trace(ctx, matA, matA, &omega));
      deltaomega = omega - oldomega;
      if (omega > oldomega && oldomega != 0.0) {
        CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matT, 1.0, matTold));
        omega = oldomega;
        break;
      }
      oldomega = omega;
      CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matTold, 1.0, matT));
      CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matT, 1.0, matT1));
    }
    if (iters) *iters = std::min(n, max_iter);
    if (omega_out) *omega_out = omega;
    CQ_TRY(tmp.release());
    return cq::hip_check(hipStreamSynchronize(s), "cqmm_hotelling");
  });
}

int cqmm_init_mcweeny(cqmm_ctx *ctx, const cqmm_initmcw_setup *su, int32_t matH, int32_t matS, int32_t matT,
                      int32_t matL, double out_mu_lambda[2]) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !su) return cq::set_error(CQMM_ERR_ARG, "cqmm_init_mcweeny: null argument");
    CQ_TRY(initmcw_check(ctx, su, matH, matS, matT, matL, "cqmm_init_mcweeny"));
    const double n_o = initmcw_orbitals(ctx, su);
    const double ne = su->n_electrons;
    cq::Temps tmp(ctx);
    InitMcwParts q;
    CQ_TRY(initmcw_parts(ctx, su, matH, matS, matT, tmp, &q));
    double mubar = 0.0, lambda = 0.0;
    CQ_TRY(initmcw_mubar(n_o, ne, q.SXHX, q.hmin, q.hmax, 1.0 - q.SX / n_o, false, &mubar, &lambda));
    // McWRho0 (src/McWeeny.f90:961-1000): L = ((l m + Ne)/No) T - (l/No) XHX.
    CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matL, (lambda * mubar + ne) / n_o, matT));
    CQ_TRY(cqmm_matrix_sum(ctx, 1.0, matL, -lambda / n_o, q.XHX));
    if (out_mu_lambda) {
      out_mu_lambda[0] = mubar;
      out_mu_lambda[1] = lambda;
    }
    CQ_TRY(tmp.release());
    return cq::hip_check(hipStreamSynchronize(cq::stream()), "cqmm_init_mcweeny");
  });
}

int cqmm_init_mcweeny_spin(cqmm_ctx *ctx, const cqmm_initmcw_setup *su, int32_t fixed_populations,
                           const int32_t matH[2], const int32_t matS[2], const int32_t matT[2], const int32_t matL[2],
                           const double n_electrons[2], double out_mu_lambda[4]) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !su || !matH || !matS || !matT || !matL || !n_electrons)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_init_mcweeny_spin: null argument");
    if (matL[0] == matL[1]) return cq::set_error(CQMM_ERR_ARG, "cqmm_init_mcweeny_spin: one L handle for both spins");
    for (int sp = 0; sp < 2; ++sp)
      CQ_TRY(initmcw_check(ctx, su, matH[sp], matS[sp], matT[sp], matL[sp], "cqmm_init_mcweeny_spin"));
    if (fixed_populations) {
      // Each spin as the spin-unpolarised InitMcW with its own population.
      for (int sp = 0; sp < 2; ++sp) {
        cqmm_initmcw_setup one = *su;
        one.n_electrons = n_electrons[sp];
        CQ_TRY(cqmm_init_mcweeny(ctx, &one, matH[sp], matS[sp], matT[sp], matL[sp],
                                 out_mu_lambda ? out_mu_lambda + 2 * sp : nullptr));
      }
      return CQMM_OK;
    }
    // Variable populations: InitMcW on the direct sum of the two spin channels
    // (src/McWeeny.f90:380-395, 460-474, 516-554): twice the orbitals, all the
    // electrons, SXHX summed over spins, the extremes of both spins' limits,
    // one mubar and lambda (its second branch test as written there, with <).
    const double n_o = initmcw_orbitals(ctx, su);
    const double n_o_ds = n_o + n_o, n_e_ds = n_electrons[0] + n_electrons[1];
    cq::Temps tmp(ctx);
    InitMcwParts q[2];
    for (int sp = 0; sp < 2; ++sp) CQ_TRY(initmcw_parts(ctx, su, matH[sp], matS[sp], matT[sp], tmp, &q[sp]));
    const double SXHX_ds = q[0].SXHX + q[1].SXHX;
    const double hmin_ds = std::min(q[0].hmin, q[1].hmin), hmax_ds = std::max(q[0].hmax, q[1].hmax);
    const double A_ds = 1.0 - (q[0].SX + q[1].SX) / n_o_ds;
    double mubar = 0.0, lambda = 0.0;
    CQ_TRY(initmcw_mubar(n_o_ds, n_e_ds, SXHX_ds, hmin_ds, hmax_ds, A_ds, true, &mubar, &lambda));
    for (int sp = 0; sp < 2; ++sp) {
      CQ_TRY(cqmm_matrix_sum(ctx, 0.0, matL[sp], (lambda * mubar + n_e_ds) / n_o_ds, matT[sp]));
      CQ_TRY(cqmm_matrix_sum(ctx, 1.0, matL[sp], -lambda / n_o_ds, q[sp].XHX));
      if (out_mu_lambda) {
        out_mu_lambda[2 * sp] = mubar;
        out_mu_lambda[2 * sp + 1] = lambda;
      }
    }
    CQ_TRY(tmp.release());
    return cq::hip_check(hipStreamSynchronize(cq::stream()), "cqmm_init_mcweeny_spin");
  });
}

}  // extern "C"
