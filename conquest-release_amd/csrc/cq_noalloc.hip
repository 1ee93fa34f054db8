// FP32 products of cqmm_mat_mult without device allocation.
//
// cqmm_mat_mult with CQMM_FP32 converts both operands to FP32, runs the FP32
// product kernels and widens the result, with the three FP32 arrays held in
// buffers the plan allocates on first use. cqmm_mat_mult_fp32_noalloc runs the
// same kernels with those arrays carved from one file-scope device array
// (`scratch`, kScratchBytes per device, held until the process ends), so that
// a caller that cannot allocate (code under stream capture, a serving loop
// that forbids hipMalloc) can still run it. The device address of the array is
// looked up once per device at cqmm_init (hipGetSymbolAddress); host code uses
// that address only, never the array's name.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"
#include "cq_ops.h"

namespace {

constexpr int64_t kScratchBytes = 64ll << 20;  // 64 MiB per device
constexpr int64_t kAlignFloats = 64;           // 256-byte alignment of each carved array
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kScratchBytes / 4];

// Device address of `scratch` on each device (null until cqmm_init ran there).
float *g_scratch[kMaxDevices] = {};

int64_t round_up(int64_t n) { return (n + kAlignFloats - 1) / kAlignFloats * kAlignFloats; }

}  // namespace

namespace cq {

// Called by cqmm_init with `device` current.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_errorf(CQMM_ERR_ARG, "cqmm_init: device %d beyond the no-allocation table (%d)", device, kMaxDevices);
  if (g_scratch[device]) return CQMM_OK;
  void *p = nullptr;
  const int st = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (no-allocation scratch)");
  if (st) return st;
  g_scratch[device] = static_cast<float *>(p);
  return CQMM_OK;
}

}  // namespace cq

extern "C" {

int64_t cqmm_noalloc_scratch_bytes(void) { return kScratchBytes; }

int cqmm_mat_mult_fp32_noalloc(cqmm_plan *plan, double *a, int64_t lena, const double *b, int64_t lenb, double *c,
                               int64_t lenc, uint32_t flags) {
  return cq::guarded(__func__, [&]() -> int {
    if (!plan || !a || !b || !c) return cq::set_error(CQMM_ERR_ARG, "cqmm_mat_mult_fp32_noalloc: null argument");
    if (!(flags & CQMM_DEVICE_PTRS))
      return cq::set_error(CQMM_ERR_ARG, "cqmm_mat_mult_fp32_noalloc: needs device pointers (CQMM_DEVICE_PTRS)");
    if (lena < plan->lena || lenb < plan->lenb || lenc < plan->lenc)
      return cq::set_errorf(CQMM_ERR_ARG,
                            "cqmm_mat_mult_fp32_noalloc: lengths (%lld,%lld,%lld) below the plan's (%lld,%lld,%lld)",
                            (long long)lena, (long long)lenb, (long long)lenc, (long long)plan->lena,
                            (long long)plan->lenb, (long long)plan->lenc);
    const cq::Runtime &rt = cq::runtime();
    if (!rt.initialised || rt.device < 0 || rt.device >= kMaxDevices || !g_scratch[rt.device])
      return cq::set_error(CQMM_ERR_STATE, "cqmm_mat_mult_fp32_noalloc: call cqmm_init first");
    const cq::TripletPlan &tp = plan->tp;
    const int64_t np = round_up(tp.len_p), nq = round_up(tp.len_q), no = round_up(tp.len_out);
    if (np + nq + no > kScratchBytes / 4)
      return cq::set_errorf(CQMM_ERR_ARG,
                            "cqmm_mat_mult_fp32_noalloc: operands need %lld FP32 elements, above the %lld of the "
                            "no-allocation scratch",
                            (long long)(np + nq + no), (long long)(kScratchBytes / 4));
    float *fp = g_scratch[rt.device], *fq = fp + np, *fo = fq + nq;
    hipStream_t s = cq::stream();
    const double *dP = tp.mult_type == 1 ? a : c;
    double *dO = tp.mult_type == 1 ? c : a;
    int st;
    // The FP32 copies and the FP32 result are fully overwritten by every call
    // (each output block is written once, zeros where it has no products).
    const int mode = (flags & CQMM_BF16X3) ? 4 : (flags & CQMM_FP32_RUN64) ? 3 : (flags & CQMM_FP32_ACC64) ? 2 : 1;
    if (mode == 4) {
      if ((st = cq::op_to_bf16x2(tp.len_p, dP, (uint32_t *)fp, s))) return st;
      if ((st = cq::op_to_bf16x2(tp.len_q, b, (uint32_t *)fq, s))) return st;
    } else {
      if ((st = cq::op_to_f32(tp.len_p, dP, fp, s))) return st;
      if ((st = cq::op_to_f32(tp.len_q, b, fq, s))) return st;
    }
    if ((st = cq::launch_block_products_f32(tp, fp, fq, fo, mode, s))) return st;
    if ((st = cq::op_to_f64(tp.len_out, fo, dO, s))) return st;
    if (!(flags & CQMM_NO_SYNC)) return cq::hip_check(hipStreamSynchronize(s), "cqmm_mat_mult_fp32_noalloc");
    return CQMM_OK;
  });
}

}  // extern "C"
