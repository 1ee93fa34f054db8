// Block-product kernels over a triplet plan (the arithmetic of m_kern_max /
// m_kern_min, src/multiply_kernel_default.f90:204-214 and :392-402).
//
// Type 1:  C_ij(n1,n2) += sum_n3 A_ik(n1,n3) * B_kj(n3,n2)
// Type 2:  X_ik(n1,n3) += sum_n2 Y_ij(n1,n2) * B_kj(n3,n2)
// Both are  O(m,n) = sum_t sum_q P_t(m,q) * Q_t(q,n)  with P column-major
// (m fastest) and Q either column-major (q fastest, type 1) or used transposed
// (n fastest, type 2). Blocks are at most 64x64: up to 16x16 one wavefront
// owns an output block; larger blocks are split into 16x16 sub-blocks, one
// wavefront each (k_block_products_big).
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

#include <cstdlib>
#include <string>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

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

// Uniform block size NSF (<= 16): the same products as k_block_products_f64,
// with
//  * an XCD-aware workgroup order: hardware dispatches workgroup b to XCD b % 8,
//    so workgroup b takes segment group (b % 8) * ceil(G/8) + b / 8 (the grid is
//    padded to 8*ceil(G/8) workgroups so every group runs once). Consecutive
//    segments share a B row (type 2: same halo atom k) or an A row (type 1:
//    same primary atom i) and now run on one XCD, whose L2 keeps that row;
//  * wave-uniform segment and triplet indices (scalar loads);
//  * a two-stage register pipeline: the operands of product t+1 are loaded
//    while the MFMAs of product t run.
template <typename OffT, bool QT, int NSF>
__global__ void __launch_bounds__(64 * kWaves) k_block_products_f64_u(
    const int64_t *__restrict__ seg_beg, const int64_t *__restrict__ seg_out, const OffT *__restrict__ tp,
    const OffT *__restrict__ tq, const double *__restrict__ P, const double *__restrict__ Q,
    double *__restrict__ O, int64_t n_seg, int32_t n_groups) {
  constexpr int NS = (NSF + 3) / 4;  // MFMA k-steps per product
  const int lane = threadIdx.x & 63;
  const int xcd = blockIdx.x & 7;
  const int32_t per = (n_groups + 7) >> 3;
  const int32_t grp = xcd * per + (int32_t)(blockIdx.x >> 3);
  if (grp >= n_groups) return;
  const int64_t seg = (int64_t)grp * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (seg >= n_seg) return;
  const int r = lane & 15, g = lane >> 4;
  const int64_t t0 = seg_beg[seg], t1 = seg_beg[seg + 1];
  // Per-lane element offsets inside a block for each k-step (constant per lane).
  int oa[NS], ob[NS];
  bool va[NS], vb[NS];
#pragma unroll
  for (int s4 = 0; s4 < NS; ++s4) {
    const int qi = 4 * s4 + g;
    va[s4] = (r < NSF) && (qi < NSF);
    vb[s4] = va[s4];
    oa[s4] = r + NSF * qi;
    ob[s4] = QT ? (r + NSF * qi) : (qi + NSF * r);
  }
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  double a0[NS], b0[NS];
  if (t0 < t1) {
    const int64_t p = (int64_t)tp[t0], q = (int64_t)tq[t0];
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4) {
      a0[s4] = va[s4] ? P[p + oa[s4]] : 0.0;
      b0[s4] = vb[s4] ? Q[q + ob[s4]] : 0.0;
    }
  }
  for (int64_t t = t0; t < t1; ++t) {
    double a1[NS], b1[NS];
    if (t + 1 < t1) {
      const int64_t p = (int64_t)tp[t + 1], q = (int64_t)tq[t + 1];
#pragma unroll
      for (int s4 = 0; s4 < NS; ++s4) {
        a1[s4] = va[s4] ? P[p + oa[s4]] : 0.0;
        b1[s4] = vb[s4] ? Q[q + ob[s4]] : 0.0;
      }
    }
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s4], b0[s4], acc, 0, 0, 0);
#pragma unroll
    for (int s4 = 0; s4 < NS; ++s4) {
      a0[s4] = a1[s4];
      b0[s4] = b1[s4];
    }
  }
  const int64_t o = seg_out[seg];
  if (r < NSF) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = g + 4 * e;
      if (row < NSF) O[o + row + (int64_t)NSF * r] = acc[e];
    }
  }
}

// Blocks larger than 16x16 (up to 64x64): wavefront w of the grid takes
// sub-block (w % nsub) of output block (w / nsub), rows 16*(st / nsc) .. +15 and
// columns 16*(st % nsc) .. +15 (nsc = sub-blocks per row of the largest block,
// nsub = nsc^2; sub-blocks past a smaller block leave at once). The
// contraction runs over the product's full inner dimension, 4 per MFMA step,
// in the same order for every sub-block. FP64, or FP32 operands with FP32
// accumulation (ACC64 unset) or each product's FP32 result added in FP64.
template <typename OffT, bool QT, typename T, bool ACC64>
__global__ void __launch_bounds__(64 * kWaves) k_block_products_big(
    const int64_t *__restrict__ seg_beg, const int64_t *__restrict__ seg_out, const uint8_t *__restrict__ seg_mn,
    const OffT *__restrict__ tp, const OffT *__restrict__ tq, const uint8_t *__restrict__ tk,
    const T *__restrict__ P, const T *__restrict__ Q, T *__restrict__ O, int64_t n_seg, int32_t nsc) {
  const int lane = threadIdx.x & 63;
  const int nsub = nsc * nsc;
  const int64_t wv = (int64_t)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t seg = wv / nsub;
  if (seg >= n_seg) return;
  const int st = (int)(wv - seg * nsub);
  const int r0 = 16 * (st / nsc), c0 = 16 * (st % nsc);
  const int m = seg_mn[2 * seg], n = seg_mn[2 * seg + 1];
  if (r0 >= m || c0 >= n) return;
  const int r = lane & 15, g = lane >> 4;
  const bool ra = r0 + r < m, rb = c0 + r < n;
  const int64_t t0 = seg_beg[seg], t1 = seg_beg[seg + 1];
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  flt4 acc32 = {0.f, 0.f, 0.f, 0.f};
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t p = (int64_t)tp[t], q = (int64_t)tq[t];
    const int kk = (int)tk[t];
    const int nsteps = (kk + 3) >> 2;
    for (int s4 = 0; s4 < nsteps; ++s4) {
      const int qi = 4 * s4 + g;
      T a = (T)0, b = (T)0;
      if (ra && qi < kk) a = P[p + (r0 + r) + (int64_t)m * qi];
      if (rb && qi < kk) b = QT ? Q[q + (c0 + r) + (int64_t)n * qi] : Q[q + qi + (int64_t)kk * (c0 + r)];
      if constexpr (sizeof(T) == 8) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      else acc32 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc32, 0, 0, 0);
    }
    if constexpr (sizeof(T) == 4 && ACC64) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        acc[e] += (double)acc32[e];
        acc32[e] = 0.f;
      }
    }
  }
  if (!rb) return;
  const int64_t o = seg_out[seg];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    // FP64 accumulator: D[(l/16)+4e][l%16]; FP32: D[4(l/16)+e][l%16].
    const int row = r0 + (sizeof(T) == 8 ? g + 4 * e : 4 * g + e);
    if (row < m) {
      T v;
      if constexpr (sizeof(T) == 8) v = acc[e];
      else v = ACC64 ? (T)acc[e] : acc32[e];
      O[o + row + (int64_t)m * (c0 + r)] = v;
    }
  }
}

template <typename OffT, bool QT, typename T, bool ACC64>
int launch_big(const cq::TripletPlan &tp, const T *P, const T *Q, T *O, hipStream_t s) {
  const int32_t nsc = (tp.maxn + 15) / 16;
  const int64_t waves = tp.n_seg * (int64_t)nsc * nsc;
  const int64_t grid = (waves + kWaves - 1) / kWaves;
  if (grid > INT32_MAX) return cq::set_error(CQMM_ERR_ARG, "block products: too many output sub-blocks for one launch");
  hipLaunchKernelGGL((k_block_products_big<OffT, QT, T, ACC64>), dim3((unsigned)grid), dim3(64 * kWaves), 0, s,
                     tp.seg_beg.as<int64_t>(), tp.seg_out.as<int64_t>(), tp.seg_mn.as<uint8_t>(), tp.tp.as<OffT>(),
                     tp.tq.as<OffT>(), tp.tk.as<uint8_t>(), P, Q, O, tp.n_seg, nsc);
  return cq::hip_check(hipGetLastError(), "block-product kernel launch (large blocks)");
}

// CQMM_KERNEL=simple selects the one-stage triplet kernel for every block size
// (used to measure what the pipelined, XCD-ordered and tiled kernels gain).
bool force_general() {
  static const int v = [] {
    const char *e = getenv("CQMM_KERNEL");
    return (e && std::string(e) == "simple") ? 1 : 0;
  }();
  return v != 0;
}

template <typename OffT, bool QT>
int launch_t(const cq::TripletPlan &tp, const double *P, const double *Q, double *O, hipStream_t s) {
  if (tp.maxn > 16) return launch_big<OffT, QT, double, false>(tp, P, Q, O, s);
  const unsigned grid = (unsigned)((tp.n_seg + kWaves - 1) / kWaves);
  const int64_t *sb = tp.seg_beg.as<int64_t>();
  const int64_t *so = tp.seg_out.as<int64_t>();
  const uint8_t *mn = tp.seg_mn.as<uint8_t>();
  const OffT *pp = tp.tp.as<OffT>();
  const OffT *qq = tp.tq.as<OffT>();
  const uint8_t *kk = tp.tk.as<uint8_t>();
  dim3 blk(64 * kWaves);
  // The XCD-aware order needs every (xcd, slot) pair: pad the grid to 8*ceil(G/8).
  const int32_t ngrp = (int32_t)grid;
  const unsigned gridx = 8u * (unsigned)((ngrp + 7) / 8);
  switch (force_general() ? -1 : tp.uniform) {
    case 4: hipLaunchKernelGGL((k_block_products_f64_u<OffT, QT, 4>), dim3(gridx), blk, 0, s, sb, so, pp, qq, P, Q, O, tp.n_seg, ngrp); break;
    case 9: hipLaunchKernelGGL((k_block_products_f64_u<OffT, QT, 9>), dim3(gridx), blk, 0, s, sb, so, pp, qq, P, Q, O, tp.n_seg, ngrp); break;
    case 13: hipLaunchKernelGGL((k_block_products_f64_u<OffT, QT, 13>), dim3(gridx), blk, 0, s, sb, so, pp, qq, P, Q, O, tp.n_seg, ngrp); break;
    case 16: hipLaunchKernelGGL((k_block_products_f64_u<OffT, QT, 16>), dim3(gridx), blk, 0, s, sb, so, pp, qq, P, Q, O, tp.n_seg, ngrp); break;
    default: break;
  }
  if (tp.uniform != 0 && !force_general()) return cq::hip_check(hipGetLastError(), "block-product kernel launch");
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

// FP32 variant (v_mfma_f32_16x16x4_f32, FP32 accumulate) for the bulk of the
// mixed-precision purification. Same segment/XCD/pipeline structure as the
// FP64 kernel; general block sizes use per-segment/per-product dimensions.

// ACC64: each product's FP32 MFMA result is added into FP64 registers (the
// MFMA inner sums stay FP32; the sum over products of a block is FP64).
template <typename OffT, bool QT, int NSF, bool ACC64>
__global__ void __launch_bounds__(64 * kWaves) k_block_products_f32(
    const int64_t *__restrict__ seg_beg, const int64_t *__restrict__ seg_out, const uint8_t *__restrict__ seg_mn,
    const OffT *__restrict__ tp, const OffT *__restrict__ tq, const uint8_t *__restrict__ tk,
    const float *__restrict__ P, const float *__restrict__ Q, float *__restrict__ O, int64_t n_seg, int32_t n_groups) {
  const int lane = threadIdx.x & 63;
  const int xcd = blockIdx.x & 7;
  const int32_t per = (n_groups + 7) >> 3;
  const int32_t grp = xcd * per + (int32_t)(blockIdx.x >> 3);
  if (grp >= n_groups) return;
  const int64_t seg = (int64_t)grp * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
  flt4 acc = {0.f, 0.f, 0.f, 0.f};
  double a64[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t p = (int64_t)tp[t], q = (int64_t)tq[t];
    const int kk = NSF > 0 ? NSF : (int)tk[t];
    const int nsteps = (kk + 3) >> 2;
    for (int s4 = 0; s4 < nsteps; ++s4) {
      const int qi = 4 * s4 + g;
      float a = 0.f, b = 0.f;
      if (r < m && qi < kk) a = P[p + r + (int64_t)m * qi];
      if (r < n && qi < kk) b = QT ? Q[q + r + (int64_t)n * qi] : Q[q + qi + (int64_t)kk * r];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
    if (ACC64) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a64[e] += (double)acc[e];
        acc[e] = 0.f;
      }
    }
  }
  const int64_t o = seg_out[seg];
  if (r < n) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 4 * g + e;  // FP32 accumulator: D[4*(l/16)+e][l%16]
      if (row < m) O[o + row + (int64_t)m * r] = ACC64 ? (float)a64[e] : acc[e];
    }
  }
}

template <typename OffT, bool QT, bool ACC64>
int launch_f32_t(const cq::TripletPlan &tp, const float *P, const float *Q, float *O, hipStream_t s) {
  if (tp.maxn > 16) return launch_big<OffT, QT, float, ACC64>(tp, P, Q, O, s);
  const int32_t ngrp = (int32_t)((tp.n_seg + kWaves - 1) / kWaves);
  const unsigned gridx = 8u * (unsigned)((ngrp + 7) / 8);
  const int64_t *sb = tp.seg_beg.as<int64_t>();
  const int64_t *so = tp.seg_out.as<int64_t>();
  const uint8_t *mn = tp.seg_mn.as<uint8_t>();
  const OffT *pp = tp.tp.as<OffT>();
  const OffT *qq = tp.tq.as<OffT>();
  const uint8_t *kk = tp.tk.as<uint8_t>();
  dim3 blk(64 * kWaves);
  switch (tp.uniform) {
    case 4: hipLaunchKernelGGL((k_block_products_f32<OffT, QT, 4, ACC64>), dim3(gridx), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg, ngrp); break;
    case 9: hipLaunchKernelGGL((k_block_products_f32<OffT, QT, 9, ACC64>), dim3(gridx), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg, ngrp); break;
    case 13: hipLaunchKernelGGL((k_block_products_f32<OffT, QT, 13, ACC64>), dim3(gridx), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg, ngrp); break;
    case 16: hipLaunchKernelGGL((k_block_products_f32<OffT, QT, 16, ACC64>), dim3(gridx), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg, ngrp); break;
    default: hipLaunchKernelGGL((k_block_products_f32<OffT, QT, 0, ACC64>), dim3(gridx), blk, 0, s, sb, so, mn, pp, qq, kk, P, Q, O, tp.n_seg, ngrp); break;
  }
  return cq::hip_check(hipGetLastError(), "FP32 block-product kernel launch");
}

}  // namespace

namespace cq {

bool uses_tiles(const TripletPlan &tp) {
  return tp.tiles.n_tiles > 0 && tp.kernel != CQMM_KERNEL_TRIPLET && !force_general();
}

// The launches of a plan's products (tile schedule or per-output-block kernels)
// without the symmetric-result mirror.
int products_f32(const TripletPlan &tp, const float *P, const float *Q, float *O, int mode, hipStream_t s);
int products_f64(const TripletPlan &tp, const double *P, const double *Q, double *O, bool fp32, hipStream_t s);

int launch_block_products_f32(const TripletPlan &tp, const float *P, const float *Q, float *O, int mode,
                              hipStream_t s) {
  const int st = products_f32(tp, P, Q, O, mode, s);
  return st ? st : mirror_result_f32(tp, O, s);
}

int launch_block_products(const TripletPlan &tp, const double *P, const double *Q, double *O, bool fp32,
                          hipStream_t s) {
  const int st = products_f64(tp, P, Q, O, fp32, s);
  return st ? st : mirror_result(tp, O, s);
}

int products_f32(const TripletPlan &tp, const float *P, const float *Q, float *O, int mode, hipStream_t s) {
  if (tp.n_seg == 0) return CQMM_OK;
  if (mode < 1 || mode > 4) return set_errorf(CQMM_ERR_ARG, "block products: FP32 mode %d", mode);
  if (uses_tiles(tp))
    return launch_tile_products_f32(tp.tiles, tp.q_trans, P, Q, O, mode, s);
  if (mode == 4) return set_error(CQMM_ERR_STATE, "split-BF16 products need a tile schedule of blocks 9 to 16");
  const bool acc64 = mode != 1;
  if (acc64) {
    if (tp.off64) return tp.q_trans ? launch_f32_t<int64_t, true, true>(tp, P, Q, O, s) : launch_f32_t<int64_t, false, true>(tp, P, Q, O, s);
    return tp.q_trans ? launch_f32_t<int32_t, true, true>(tp, P, Q, O, s) : launch_f32_t<int32_t, false, true>(tp, P, Q, O, s);
  }
  if (tp.off64) return tp.q_trans ? launch_f32_t<int64_t, true, false>(tp, P, Q, O, s) : launch_f32_t<int64_t, false, false>(tp, P, Q, O, s);
  return tp.q_trans ? launch_f32_t<int32_t, true, false>(tp, P, Q, O, s) : launch_f32_t<int32_t, false, false>(tp, P, Q, O, s);
}

int products_f64(const TripletPlan &tp, const double *P, const double *Q, double *O, bool fp32, hipStream_t s) {
  if (fp32) return set_error(CQMM_ERR_ARG, "block products: FP32 path not built in this library");
  if (tp.n_seg == 0) return CQMM_OK;
  if (tp.uniform != 0 && tp.uniform != 4 && tp.uniform != 9 && tp.uniform != 13 && tp.uniform != 16) {
    // Other uniform sizes run through the general instance with per-block sizes.
    return set_errorf(CQMM_ERR_ARG, "block products: uniform block size %d unsupported", tp.uniform);
  }
  if (uses_tiles(tp))
    return launch_tile_products(tp.tiles, tp.q_trans, P, Q, O, s);
  if (tp.off64) return tp.q_trans ? launch_t<int64_t, true>(tp, P, Q, O, s) : launch_t<int64_t, false>(tp, P, Q, O, s);
  return tp.q_trans ? launch_t<int32_t, true>(tp, P, Q, O, s) : launch_t<int32_t, false>(tp, P, Q, O, s);
}

}  // namespace cq
