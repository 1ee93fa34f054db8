// Block-product kernels over a triplet plan (the arithmetic of m_kern_max /
// m_kern_min, src/multiply_kernel_default.f90:204-214 and :392-402).
//
// Type 1:  C_ij(n1,n2) += sum_n3 A_ik(n1,n3) * B_kj(n3,n2)
// Type 2:  X_ik(n1,n3) += sum_n2 Y_ij(n1,n2) * B_kj(n3,n2)
// Both are  O(m,n) = sum_t sum_q P_t(m,q) * Q_t(q,n)  with P column-major
// (m fastest) and Q either column-major (q fastest, type 1) or used transposed
// (n fastest, type 2). All blocks are at most 16x16.
//
// One wavefront owns one output block (a segment of the plan) and accumulates
// its products with v_mfma_f64_16x16x4 in registers, then writes the block
// once: no atomics, a fixed summation order, the output fully overwritten as
// mat_mult overwrites it (src/multiply_module.f90:204-208).
//
// MFMA operand layout (16x16x4, f64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; the accumulator holds D[(l/16)+4*r][l%16] in element r
// (checked on gfx950 against the CPU oracle).
#include <hip/hip_runtime.h>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;  // wavefronts (output blocks) per workgroup

template <typename OffT, bool QT, int NSF>
__global__ void __launch_bounds__(64 * kWaves) k_block_products_f64(
    const int64_t *__restrict__ seg_beg, const int64_t *__restrict__ seg_out, const uint8_t *__restrict__ seg_mn,
    const OffT *__restrict__ tp, const OffT *__restrict__ tq, const uint8_t *__restrict__ tk,
    const double *__restrict__ P, const double *__restrict__ Q, double *__restrict__ O, int64_t n_seg) {
  const int lane = threadIdx.x & 63;
  const int64_t seg = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (seg >= n_seg) return;
  int m, n;
  if (NSF > 0) {
    m = NSF;
    n = NSF;
  } else {
    m = seg_mn[2 * seg];
    n = seg_mn[2 * seg + 1];
  }
  const int r = lane & 15, g = lane >> 4;
  const int64_t t0 = seg_beg[seg], t1 = seg_beg[seg + 1];
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t p = (int64_t)tp[t], q = (int64_t)tq[t];
    const int kk = NSF > 0 ? NSF : (int)tk[t];
    const int nsteps = (kk + 3) >> 2;
    for (int s4 = 0; s4 < nsteps; ++s4) {
      const int qi = 4 * s4 + g;
      double a = 0.0, b = 0.0;
      if (r < m && qi < kk) a = P[p + r + (int64_t)m * qi];
      if (r < n && qi < kk) b = QT ? Q[q + r + (int64_t)n * qi] : Q[q + qi + (int64_t)kk * r];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
  }
  const int64_t o = seg_out[seg];
  if (r < n) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = g + 4 * e;
      if (row < m) O[o + row + (int64_t)m * r] = acc[e];
    }
  }
}

template <typename OffT, bool QT>
int launch_t(const cq::TripletPlan &tp, const double *P, const double *Q, double *O, hipStream_t s) {
  const unsigned grid = (unsigned)((tp.n_seg + kWaves - 1) / kWaves);
  const int64_t *sb = tp.seg_beg.as<int64_t>();
  const int64_t *so = tp.seg_out.as<int64_t>();
  const uint8_t *mn = tp.seg_mn.as<uint8_t>();
  const OffT *pp = tp.tp.as<OffT>();
  const OffT *qq = tp.tq.as<OffT>();
  const uint8_t *kk = tp.tk.as<uint8_t>();
  dim3 blk(64 * kWaves);
  switch (tp.uniform) {
    case 4: hipLaunchKernelGGL((k_block_products_f64<OffT, QT, 4>), dim3(grid), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg); break;
    case 9: hipLaunchKernelGGL((k_block_products_f64<OffT, QT, 9>), dim3(grid), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg); break;
    case 13: hipLaunchKernelGGL((k_block_products_f64<OffT, QT, 13>), dim3(grid), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg); break;
    case 16: hipLaunchKernelGGL((k_block_products_f64<OffT, QT, 16>), dim3(grid), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg); break;
    default:
      if (tp.uniform == 0) {
        hipLaunchKernelGGL((k_block_products_f64<OffT, QT, 0>), dim3(grid), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg);
      } else {
        return cq::set_errorf(CQMM_ERR_ARG, "block products: uniform block size %d has no kernel instance", tp.uniform);
      }
  }
  return cq::hip_check(hipGetLastError(), "block-product kernel launch");
}

}  // namespace

namespace cq {

int launch_block_products(const TripletPlan &tp, const double *P, const double *Q, double *O, bool fp32,
                          hipStream_t s) {
  if (fp32) return set_error(CQMM_ERR_ARG, "block products: FP32 path not built in this library");
  if (tp.n_seg == 0) return CQMM_OK;
  if (tp.uniform != 0 && tp.uniform != 4 && tp.uniform != 9 && tp.uniform != 13 && tp.uniform != 16) {
    // Other uniform sizes run through the general instance with per-block sizes.
    return set_errorf(CQMM_ERR_ARG, "block products: uniform block size %d unsupported", tp.uniform);
  }
  if (tp.off64) return tp.q_trans ? launch_t<int64_t, true>(tp, P, Q, O, s) : launch_t<int64_t, false>(tp, P, Q, O, s);
  return tp.q_trans ? launch_t<int32_t, true>(tp, P, Q, O, s) : launch_t<int32_t, false>(tp, P, Q, O, s);
}

}  // namespace cq
