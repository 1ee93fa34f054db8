// Output-tile schedule and FP64 tile kernel for uniform block sizes.
//
// The same block products as the triplet kernels (m_kern_max / m_kern_min,
// src/multiply_kernel_default.f90:204-214 and :392-402), grouped so that
// operand blocks are read from HBM once per tile instead of once per product:
//
//  * a tile is 8 primary atoms (rows, consecutive in primary order) x 8 column
//    atoms (consecutive in the Morton order of their halo partitions) of the
//    output matrix: type 1 C(i,j), columns in the C halo, contraction over the
//    A-halo atom k; type 2 X(i,k), columns in the A halo, contraction over the
//    C-halo atom j;
//  * an entry of a tile is one contraction atom x with its (up to 8) left
//    operand blocks P(row, x) and (up to 8) right operand blocks Q(column, x);
//    entries of a tile are in ascending x (type 1: the reference's k order);
//  * one workgroup (4 wavefronts, three workgroups per CU) runs one tile.
//    Chunks of entries have their operand blocks copied HBM->LDS
//    (global_load_lds, no registers) into one of two stages while the
//    previous chunk is computed from the other (one barrier per chunk); each
//    wavefront accumulates 16 output blocks of the tile with
//    v_mfma_f64_16x16x4 from LDS and finally writes each of them once (zeros
//    where a block has no products), so the output is fully overwritten as
//    mat_mult overwrites it (src/multiply_module.f90:204-208). Ownership is a
//    Latin square (two blocks per tile row and column per wavefront), chosen
//    when the plan is built to balance the wavefronts' product counts;
//  * the plan holds each wavefront's products of each chunk as a list of
//    packed LDS offsets (left | right << 16, in elements), sorted by the
//    wavefront's output block and, within a block, by entry, with the count per
//    block in one 64-bit word (16 x 4 bits). The kernel walks that list: no
//    product discovery at run time.
// Summation order per output block is fixed (ascending x), so results are
// bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

// Four split-BF16 words -> their hi halves (bits 0-15) or lo halves (bits
// 16-31) as the four BF16 elements of an MFMA operand.
__device__ inline s16x4 pack_bf16_lo(const uint32_t w[4]) {
  const uint32_t x = __builtin_amdgcn_perm(w[1], w[0], 0x05040100u);
  const uint32_t y = __builtin_amdgcn_perm(w[3], w[2], 0x05040100u);
  return __builtin_bit_cast(s16x4, (uint64_t)x | ((uint64_t)y << 32));
}
__device__ inline s16x4 pack_bf16_hi(const uint32_t w[4]) {
  const uint32_t x = __builtin_amdgcn_perm(w[1], w[0], 0x07060302u);
  const uint32_t y = __builtin_amdgcn_perm(w[3], w[2], 0x07060302u);
  return __builtin_bit_cast(s16x4, (uint64_t)x | ((uint64_t)y << 32));
}

constexpr int kWaves = 4;                    // wavefronts per workgroup (one tile)
constexpr int kTileThreads = 64 * kWaves;
constexpr int kSlots = 64 / kWaves;          // output blocks per wavefront
static_assert(8 % kWaves == 0, "a tile row of 8 blocks is split evenly over the wavefronts");
constexpr int kStageBytes = 24 * 1024;       // one of the two operand stages (three workgroups per CU)
// LDS reads of a block product never mask rows or columns beyond NSF: those
// lanes only feed accumulator rows/columns >= NSF, which are never written.
// Only the contraction index 4*s+g >= NSF at the last k-step is zeroed (both
// operands), so the padding of the 16x16x4 steps adds nothing. kLdsPad doubles
// after each stage keep the over-reads of its last block inside the array.
constexpr int kLdsPad = 256;
constexpr int kChunkEntries = 15;            // entries per chunk at most (4-bit product counts)
constexpr int kWaveProducts = 64;            // products per wavefront per chunk at most (one per lane)
constexpr int kDenseNsf = 4;                 // block size of the dense-quadrant kernel
constexpr int kDenseEntries = 64;            // entries per chunk at most there (one per lane)

// Operand blocks per chunk (FP64 sizing): what fits kStageBytes, but at
// least the 16 blocks one entry can bring (8 left + 8 right), so every entry
// fits a stage on its own (16x16 blocks: 32 KB stages, two workgroups per CU).
constexpr int stage_cap(int nsf) {
  return kStageBytes / (nsf * nsf * 8) > 16 ? kStageBytes / (nsf * nsf * 8) : 16;
}

template <int NSF>
struct TileCfg {
  static constexpr int NS2 = NSF * NSF;
  static constexpr int CAP = stage_cap(NSF);
  static constexpr int KS = (NSF + 3) / 4;              // MFMA k-steps per product
};

int tile_cap(int nsf) { return stage_cap(nsf); }

// Prefetch loads are unconditional (indices clamped or arrays padded) so that
// no value is consumed before the next barrier: a use right after the load
// would make the wavefront wait for every outstanding load, including the
// stage copies that are meant to overlap the compute.
//
// Operand list of a chunk with nb > 0 blocks: lane k holds the operand of
// block min(w + kWaves*k, nb - 1) (lanes past wavefront w's blocks are unused).
__device__ inline int32_t load_chunk_list(const int32_t *__restrict__ list, int64_t lbeg, int32_t nb, int w,
                                          int lane) {
  return list[lbeg + min(w + kWaves * lane, nb - 1)];
}

// One 256-byte piece of a block copy: lane l moves dword OFF/4 + l. The
// immediate offset applies to both the global and the LDS address (one M0 per
// block).
template <int OFF>
__device__ inline void dma_piece(const uint32_t *src, uint32_t *dst) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                   (__attribute__((address_space(3))) void *)dst, 4, OFF, 0);
}

template <int... I>
__device__ inline void dma_full_pieces(const uint32_t *src, uint32_t *dst, std::integer_sequence<int, I...>) {
  (dma_piece<I * 256>(src, dst), ...);
}

// Issue the HBM->LDS copies of a chunk's operand blocks without waiting:
// wavefront w copies blocks w, w+kWaves, ... (operands in lv) with
// global_load_lds_dword; lane l moves dword k*64+l of a block to
// stage + block*DW + k*64 + l.
template <int NSF, class T>
__device__ inline void issue_stage(int32_t lv, int32_t nb, const T *__restrict__ P, const T *__restrict__ Q,
                                   T *stage, int w, int lane) {
  constexpr int NS2 = NSF * NSF;
  constexpr int DW = NS2 * (int)sizeof(T) / 4;
  constexpr int NF = DW / 64;          // full 64-dword pieces
  constexpr int TAIL = DW - 64 * NF;   // dwords of the last piece
  const int32_t nbw = nb > w ? (nb - w + kWaves - 1) / kWaves : 0;
  for (int32_t k = 0; k < nbw; ++k) {
    const int32_t lb = __builtin_amdgcn_readlane(lv, k);
    const uint32_t *src = (const uint32_t *)((lb < 0 ? Q : P) + (int64_t)(lb & 0x7fffffff) * NS2) + lane;
    uint32_t *dst = (uint32_t *)(stage + (w + kWaves * k) * NS2);
    dma_full_pieces(src, dst, std::make_integer_sequence<int, NF>{});
    if constexpr (TAIL > 0) {
      if (lane < TAIL) dma_piece<NF * 256>(src, dst);
    }
  }
}

using cq::TileArgs;
using cq::TileGroup;
using cq::kMaxGroup;

// Tile of this workgroup, XCD-aware: workgroup b runs on XCD b % 8 and takes
// tile (b % 8) * ceil(T/8) + b / 8, so neighbouring tiles (same rows) share an
// L2. Returns -1 past the last tile.
__device__ inline int32_t xcd_tile(int32_t n) {
  const int32_t per = (n + 7) >> 3;
  const int32_t t = (int32_t)(blockIdx.x & 7) * per + (int32_t)(blockIdx.x >> 3);
  return t < n ? t : -1;
}

__device__ inline int32_t tile_select(const TileArgs &arg, TileArgs &a) {
  a = arg;
  return xcd_tile(a.n_tiles);
}

// A grouped launch numbers the tiles of its products consecutively (product j
// from g.a[j].tile_base); the arguments (kernel-argument memory, read with
// scalar loads) are those of the product the tile belongs to.
__device__ inline int32_t tile_select(const TileGroup &g, TileArgs &a) {
  const int32_t t = xcd_tile(g.total);
  if (t < 0) return -1;
  int32_t j = 0;
#pragma unroll
  for (int k = 1; k < kMaxGroup; ++k) j += (k < g.ngrp && g.a[k].tile_base <= t) ? 1 : 0;
  a = g.a[j];
  return t - a.tile_base;
}

__device__ inline int64_t readlane64(int64_t v, int k) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), k);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// MODE 0: FP64 operands, v_mfma_f64_16x16x4 (accumulator D[(l/16)+4e][l%16]).
// MODE 1: FP32 operands, v_mfma_f32_16x16x4, FP32 accumulation.
// MODE 2: FP32 operands and MFMA, each product's result added into FP64
//         registers (accumulator D[4(l/16)+e][l%16]).
// MODE 3: FP32 operands and MFMA; the products of one output block within a
//         chunk (consecutive in the wavefront's list, at most 15) chain in one
//         FP32 accumulator, which is then added into the block's FP64 registers.
// Pipeline per chunk c (one barrier): wait for chunk c's copies and the
// prefetched registers, issue chunk c+1's copies into the other stage and the
// register loads of chunk c+1's product list and chunk c+2's operand list,
// then compute chunk c from LDS.
template <int MODE, bool QT, int NSF, class ARG>
__global__ void __launch_bounds__(kTileThreads) k_tile_products(
    const ARG arg) {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  using C = TileCfg<NSF>;
  constexpr int STAGE = C::CAP * C::NS2 + kLdsPad * (int)(8 / sizeof(T));
  __shared__ T stage0[STAGE];
  __shared__ T stage1[STAGE];
  TileArgs a;
  const int32_t tile = tile_select(arg, a);
  if (tile < 0) return;  // the whole workgroup leaves together
  const int32_t *__restrict__ tile_chunk = a.tile_chunk;
  const int64_t *__restrict__ chunk_lbeg = a.chunk_lbeg;
  const int32_t *__restrict__ list = a.list;
  const int64_t *__restrict__ pstart = a.pstart;
  const uint64_t *__restrict__ pcnt = a.pcnt;
  const uint32_t *__restrict__ prod = a.prod;
  const int32_t *__restrict__ wave_out = a.wave_out;
  const T *__restrict__ P = static_cast<const T *>(a.P);
  const T *__restrict__ Q = static_cast<const T *>(a.Q);
  T *__restrict__ O = static_cast<T *>(a.O);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, r = lane & 15, g = lane >> 4;
  constexpr int SA = 4 * NSF;             // A offset step per k-step
  constexpr int SB = QT ? 4 * NSF : 4;    // B offset step per k-step
  constexpr int REM = NSF - 4 * (C::KS - 1);  // valid g at the last k-step
  const int oA = r + NSF * g;
  const int oB = QT ? (r + NSF * g) : (g + NSF * r);
  const bool last_ok = g < REM;
  // The wavefront's output blocks (-1: none), set when the plan was built.
  const int32_t *wo = wave_out + ((int64_t)tile * kWaves + w) * kSlots;
  int32_t oblk[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) oblk[u] = wo[u];
  dbl4 a64[kSlots];
  flt4 a32[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    a64[u] = dbl4{0.0, 0.0, 0.0, 0.0};
    a32[u] = flt4{0.f, 0.f, 0.f, 0.f};
  }
  const int32_t c0 = tile_chunk[tile], c1 = tile_chunk[tile + 1];
  if (c0 < c1) {
    // Per-chunk data of the tile held in registers, 64 chunks at a time: lane
    // l has chunk_lbeg and this wavefront's product-list start and counts of
    // chunk wb + l (index clamped to c1). Vector loads, refilled after the
    // compute of a chunk, so the loop has no scalar loads (a scalar load
    // outstanding during the compute would make every LDS wait also wait for
    // it).
    int32_t wb = c0;
    int32_t cc = min(wb + lane, c1);
    int64_t wcl = chunk_lbeg[cc];
    int64_t wps = pstart[(int64_t)cc * kWaves + w];
    int64_t wpc = (int64_t)pcnt[(int64_t)cc * kWaves + w];
    auto cl_at = [&](int32_t x) { return readlane64(wcl, min(x, c1) - wb); };
    auto ps_at = [&](int32_t x) { return readlane64(wps, min(x, c1) - wb); };
    auto pc_at = [&](int32_t x) { return (uint64_t)readlane64(wpc, min(x, c1) - wb); };
    uint32_t pl;
    int32_t lvn = 0;
    {
      const int64_t cl0 = cl_at(c0), cl1 = cl_at(c0 + 1);
      const int32_t lv0 = load_chunk_list(list, cl0, (int32_t)(cl1 - cl0), w, lane);
      issue_stage<NSF, T>(lv0, (int32_t)(cl1 - cl0), P, Q, stage0, w, lane);
      pl = prod[ps_at(c0) + lane];
      if (c0 + 1 < c1) lvn = load_chunk_list(list, cl1, (int32_t)(cl_at(c0 + 2) - cl1), w, lane);
    }
    for (int32_t c = c0; c < c1; ++c) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      const bool odd = ((c - c0) & 1) != 0;
      const T *cur = odd ? stage1 : stage0;
      const uint32_t plc = pl;
      const uint64_t cnt = pc_at(c);
      if (c + 1 < c1) {
        const int64_t cl1 = cl_at(c + 1), cl2 = cl_at(c + 2);
        issue_stage<NSF, T>(lvn, (int32_t)(cl2 - cl1), P, Q, odd ? stage0 : stage1, w, lane);
        pl = prod[ps_at(c + 1) + lane];
        if (c + 2 < c1) lvn = load_chunk_list(list, cl2, (int32_t)(cl_at(c + 3) - cl2), w, lane);
      }
      int p = 0;
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int n = (int)((cnt >> (4 * u)) & 15u);
        flt4 run = {0.f, 0.f, 0.f, 0.f};  // MODE 3: this block's products of the chunk
        for (int i = 0; i < n; ++i, ++p) {
          const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)plc, p);
          if constexpr (MODE == 4) {
            // Split-BF16 x3: the whole contraction (NSF <= 16) in one K = 16
            // step; lane (r, g) holds contraction indices 4g..4g+3 as four
            // LDS words (bits 0-15 hi, 16-31 lo; indices >= NSF zeroed),
            // repacked into the hi and lo operands of v_mfma_f32_16x16x16_bf16.
            // hi.hi + hi.lo + lo.hi chain in FP32; each product's sum is
            // added into the FP64 accumulator (D[4g + e][r]).
            const T *la = cur + (int)(v & 0xffffu) + r + NSF * 4 * g;
            const T *lq = cur + (int)(v >> 16) + (QT ? r + NSF * 4 * g : 4 * g + NSF * r);
            uint32_t wa[4], wb[4];
#pragma unroll
            for (int i4 = 0; i4 < 4; ++i4) {
              wa[i4] = __float_as_uint(la[NSF * i4]);
              wb[i4] = __float_as_uint(lq[QT ? NSF * i4 : i4]);
              if (NSF < 16 && 4 * g + i4 >= NSF) {
                wa[i4] = 0u;
                wb[i4] = 0u;
              }
            }
            const s16x4 ahi = pack_bf16_lo(wa), alo = pack_bf16_hi(wa);
            const s16x4 bhi = pack_bf16_lo(wb), blo = pack_bf16_hi(wb);
            flt4 t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, flt4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, t, 0, 0, 0);
            t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, t, 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) a64[u][e] += (double)t[e];
            continue;
          }
          if constexpr (REM == 1) {
            // NSF = 4m + 1: the last contraction index k = NSF - 1 is a
            // rank-1 update on the VALU, D[row][r] += A[row][k] * B[k][r]
            // (one fused multiply-add per accumulator element, FP64 or FP32),
            // instead of an MFMA step with three zeroed k (a quarter fewer
            // MFMAs; accumulators in VGPRs, built with the VGPR MFMA form).
            // Rows g+4e >= NSF read past the block (inside the stage and its
            // pad) and only feed accumulator rows that are never written.
            const T *la = cur + (int)(v & 0xffffu) + oA;
            const T *lq = cur + (int)(v >> 16) + oB;
            T a[C::KS - 1], bb[C::KS - 1];
#pragma unroll
            for (int s4 = 0; s4 < C::KS - 1; ++s4) {
              a[s4] = la[SA * s4];
              bb[s4] = lq[SB * s4];
            }
            // Accumulator rows: FP64 g+4e, FP32 4g+e.
            const T *ca = cur + (int)(v & 0xffffu) + NSF * (NSF - 1) + (MODE == 0 ? g : 4 * g);
            const T bl = QT ? cur[(int)(v >> 16) + NSF * (NSF - 1) + r] : cur[(int)(v >> 16) + (NSF - 1) + NSF * r];
            T ce[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) ce[e] = ca[MODE == 0 ? 4 * e : e];
            if constexpr (MODE == 0) {
#pragma unroll
              for (int s4 = 0; s4 < C::KS - 1; ++s4)
                a64[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s4], bb[s4], a64[u], 0, 0, 0);
#pragma unroll
              for (int e = 0; e < 4; ++e) a64[u][e] = __builtin_fma(ce[e], bl, a64[u][e]);
            } else if constexpr (MODE == 1) {
#pragma unroll
              for (int s4 = 0; s4 < C::KS - 1; ++s4)
                a32[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4], bb[s4], a32[u], 0, 0, 0);
#pragma unroll
              for (int e = 0; e < 4; ++e) a32[u][e] = __builtin_fmaf(ce[e], bl, a32[u][e]);
            } else if constexpr (MODE == 3) {
#pragma unroll
              for (int s4 = 0; s4 < C::KS - 1; ++s4) run = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4], bb[s4], run, 0, 0, 0);
#pragma unroll
              for (int e = 0; e < 4; ++e) run[e] = __builtin_fmaf(ce[e], bl, run[e]);
            } else {
              flt4 t = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
              for (int s4 = 0; s4 < C::KS - 1; ++s4) t = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4], bb[s4], t, 0, 0, 0);
#pragma unroll
              for (int e = 0; e < 4; ++e) a64[u][e] += (double)__builtin_fmaf(ce[e], bl, t[e]);
            }
            continue;
          }
          const T *la = cur + (int)(v & 0xffffu) + oA;
          const T *lq = cur + (int)(v >> 16) + oB;
          T a[C::KS], bb[C::KS];
#pragma unroll
          for (int s4 = 0; s4 < C::KS; ++s4) {
            a[s4] = la[SA * s4];
            bb[s4] = lq[SB * s4];
          }
          if (REM < 4 && !last_ok) {
            a[C::KS - 1] = (T)0;
            bb[C::KS - 1] = (T)0;
          }
          if constexpr (MODE == 0) {
#pragma unroll
            for (int s4 = 0; s4 < C::KS; ++s4)
              a64[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s4], bb[s4], a64[u], 0, 0, 0);
          } else if constexpr (MODE == 1) {
#pragma unroll
            for (int s4 = 0; s4 < C::KS; ++s4)
              a32[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4], bb[s4], a32[u], 0, 0, 0);
          } else if constexpr (MODE == 3) {
#pragma unroll
            for (int s4 = 0; s4 < C::KS; ++s4) run = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4], bb[s4], run, 0, 0, 0);
          } else {
            flt4 t = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s4 = 0; s4 < C::KS; ++s4) t = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s4], bb[s4], t, 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) a64[u][e] += (double)t[e];
          }
        }
        if constexpr (MODE == 3) {
          if (n > 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) a64[u][e] += (double)run[e];
          }
        }
      }
      if (c + 4 > wb + 63 && c + 1 < c1) {  // next iteration reads up to chunk c + 3
        wb = c + 1;
        cc = min(wb + lane, c1);
        wcl = chunk_lbeg[cc];
        wps = pstart[(int64_t)cc * kWaves + w];
        wpc = (int64_t)pcnt[(int64_t)cc * kWaves + w];
      }
    }
  }
  if (r < NSF) {
#pragma unroll
    for (int u = 0; u < kSlots; ++u) {
      if (oblk[u] < 0) continue;
      T *o = O + (int64_t)oblk[u] * C::NS2;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = MODE == 0 ? g + 4 * e : 4 * g + e;
        if (row < NSF) o[row + NSF * r] = MODE == 0 ? (T)a64[u][e] : (MODE == 1 ? (T)a32[u][e] : (T)a64[u][e]);
      }
    }
  }
}

// 4x4 blocks: a tile is 32x32 elements, and wavefront w owns its 16x16
// quadrant (block rows 4*(w/2)..+3, block columns 4*(w%2)..+3). One
// 16x16x4 MFMA per entry covers the entry's whole contraction (4) for all 16
// blocks of the quadrant at once: lane (r, g) supplies P(row 4*(w/2)+r/4,
// x)[r%4][g] and Q(column 4*(w%2)+r/4, x)[g][r%4], zero where the entry has no
// such block. A block pair whose output block does not exist only feeds
// accumulator elements that are never written; every output block sums its
// products in entry (x) order, one fused step per product, as the triplet
// kernel does. Same chunks, stages and copies as k_tile_products; the chunk's
// entries (at most 64, one per lane) replace the product lists.
template <int MODE, bool QT, class ARG>
__global__ void __launch_bounds__(kTileThreads) k_tile_dense4(
    const ARG arg) {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  constexpr int NS2 = kDenseNsf * kDenseNsf;
  constexpr int STAGE = stage_cap(kDenseNsf) * NS2 + kLdsPad * (int)(8 / sizeof(T));
  __shared__ T stage0[STAGE];
  __shared__ T stage1[STAGE];
  TileArgs a;
  const int32_t tile = tile_select(arg, a);
  if (tile < 0) return;  // the whole workgroup leaves together
  const int32_t *__restrict__ tile_chunk = a.tile_chunk;
  const int64_t *__restrict__ chunk_lbeg = a.chunk_lbeg;
  const int32_t *__restrict__ list = a.list;
  const int32_t *__restrict__ chunk_entry = a.chunk_entry;
  const uint32_t *__restrict__ emask = a.emask;
  const int32_t *__restrict__ ecoff = a.ecoff;
  const int32_t *__restrict__ wave_out = a.wave_out;
  const T *__restrict__ P = static_cast<const T *>(a.P);
  const T *__restrict__ Q = static_cast<const T *>(a.Q);
  T *__restrict__ O = static_cast<T *>(a.O);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, r = lane & 15, g = lane >> 4;
  const int qr = w >> 1, qc = w & 1;
  const int ua = 4 * qr + (r >> 2);  // tile row of this lane's left operand
  const int cb = 4 * qc + (r >> 2);  // tile column of this lane's right operand
  const int oa = (r & 3) + 4 * g;    // P(row, x)[r % 4][g]
  const int ob = QT ? (r & 3) + 4 * g : g + 4 * (r & 3);  // Q[g][r % 4]
  const uint32_t below_a = (1u << ua) - 1u, below_b = (1u << cb) - 1u;
  const uint32_t quad_rows = 0xfu << (4 * qr), quad_cols = 0xfu << (8 + 4 * qc);
  dbl4 a64 = {0.0, 0.0, 0.0, 0.0};
  flt4 a32 = {0.f, 0.f, 0.f, 0.f};
  const int32_t c0 = tile_chunk[tile], c1 = tile_chunk[tile + 1];
  if (c0 < c1) {
    // Per-chunk operand-list and entry bounds of the tile in registers, 64
    // chunks at a time (lane l: chunk wb + l, clamped to c1).
    int32_t wb = c0;
    int32_t cc = min(wb + lane, c1);
    int64_t wcl = chunk_lbeg[cc];
    int32_t wce = chunk_entry[cc];
    auto cl_at = [&](int32_t x) { return readlane64(wcl, min(x, c1) - wb); };
    auto ce_at = [&](int32_t x) { return __builtin_amdgcn_readlane(wce, min(x, c1) - wb); };
    // The entries of a chunk, lane i = entry i (index clamped to the chunk).
    auto entry_regs = [&](int32_t e0, int32_t e1, uint32_t &m, int32_t &o) {
      const int32_t e = min(e0 + lane, e1 - 1);
      m = emask[e];
      o = ecoff[e];
    };
    uint32_t em, emn = 0;
    int32_t eo, eon = 0;
    int32_t lvn = 0;
    {
      const int64_t cl0 = cl_at(c0), cl1 = cl_at(c0 + 1);
      const int32_t lv0 = load_chunk_list(list, cl0, (int32_t)(cl1 - cl0), w, lane);
      issue_stage<kDenseNsf, T>(lv0, (int32_t)(cl1 - cl0), P, Q, stage0, w, lane);
      entry_regs(ce_at(c0), ce_at(c0 + 1), emn, eon);
      if (c0 + 1 < c1) lvn = load_chunk_list(list, cl1, (int32_t)(cl_at(c0 + 2) - cl1), w, lane);
    }
    for (int32_t c = c0; c < c1; ++c) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      const bool odd = ((c - c0) & 1) != 0;
      const T *cur = odd ? stage1 : stage0;
      em = emn;
      eo = eon;
      const int32_t ne = ce_at(c + 1) - ce_at(c);
      if (c + 1 < c1) {
        const int64_t cl1 = cl_at(c + 1), cl2 = cl_at(c + 2);
        issue_stage<kDenseNsf, T>(lvn, (int32_t)(cl2 - cl1), P, Q, odd ? stage0 : stage1, w, lane);
        entry_regs(ce_at(c + 1), ce_at(c + 2), emn, eon);
        if (c + 2 < c1) lvn = load_chunk_list(list, cl2, (int32_t)(cl_at(c + 3) - cl2), w, lane);
      }
      for (int i = 0; i < ne; ++i) {
        const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)em, i);
        if ((m & quad_rows) == 0 || (m & quad_cols) == 0) continue;  // no block pair in this quadrant
        const int32_t base = __builtin_amdgcn_readlane(eo, i);
        const uint32_t ym = m & 0xffu, bm = (m >> 8) & 0xffu;
        const int ia = base + __popc(ym & below_a);
        const int ib = base + __popc(ym) + __popc(bm & below_b);
        T a = cur[ia * NS2 + oa], b = cur[ib * NS2 + ob];
        if (((ym >> ua) & 1u) == 0) a = (T)0;
        if (((bm >> cb) & 1u) == 0) b = (T)0;
        if constexpr (MODE == 0) {
          a64 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, a64, 0, 0, 0);
        } else if constexpr (MODE == 1) {
          a32 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, a32, 0, 0, 0);
        } else {
          const flt4 t = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, flt4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
          for (int e = 0; e < 4; ++e) a64[e] += (double)t[e];
        }
      }
      if (c + 4 > wb + 63 && c + 1 < c1) {  // next iteration reads up to chunk c + 3
        wb = c + 1;
        cc = min(wb + lane, c1);
        wcl = chunk_lbeg[cc];
        wce = chunk_entry[cc];
      }
    }
  }
  // FP64 accumulator D[g+4e][r], FP32 D[4g+e][r]: quadrant row R, column r is
  // element (R % 4, r % 4) of block (R / 4, r / 4).
  const int32_t *wo = wave_out + ((int64_t)tile * kWaves + w) * 16;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int rq = MODE == 0 ? e : g, ri = MODE == 0 ? g : e;
    const int32_t blk = wo[rq * 4 + (r >> 2)];
    if (blk >= 0)
      O[(int64_t)blk * NS2 + ri + 4 * (r & 3)] = MODE == 1 ? (T)a32[e] : (T)a64[e];
  }
}

// ---------------------------------------------------------------- builder

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__global__ void k_fill_i32(int32_t *__restrict__ v, int64_t n, int32_t val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = val;
}

__global__ void k_chunk_lbeg(const int32_t *__restrict__ chunk_entry, const int64_t *__restrict__ entry_lbeg,
                             int64_t nc1, int64_t *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc1) out[c] = entry_lbeg[chunk_entry[c]];
}

// Operand-list start of each entry relative to its chunk's first operand.
__global__ void k_entry_coff(const int32_t *__restrict__ chunk_entry, const int64_t *__restrict__ entry_lbeg,
                             int64_t nc, int32_t *__restrict__ coff) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int64_t base = entry_lbeg[chunk_entry[c]];
  for (int32_t e = chunk_entry[c]; e < chunk_entry[c + 1]; ++e) coff[e] = (int32_t)(entry_lbeg[e] - base);
}

__global__ void k_iota32(int32_t *__restrict__ v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = (int32_t)t;
}

// Products per output slot of each tile.
__global__ void k_slot_counts(const uint64_t *__restrict__ key, const uint8_t *__restrict__ slot, int64_t n,
                              int64_t nx, int32_t *__restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) atomicAdd(&cnt[(int64_t)(key[t] / (uint64_t)nx) * 64 + slot[t]], 1);
}

// flag[pos] = 1 where a new (tile, x) key starts in the sorted list.
__global__ void k_entry_flags(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) flag[t] = (t == 0 || key[t] != key[t - 1]) ? 1 : 0;
}

// Scatter each product's operand blocks into its entry's row / column slots.
// Products of one entry that share a row (column) write the same block index.
__global__ void k_entry_fill(const uint64_t *__restrict__ key, const int32_t *__restrict__ perm,
                             const int32_t *__restrict__ eid, const uint8_t *__restrict__ slot,
                             const int64_t *__restrict__ p, const int64_t *__restrict__ q, int64_t n, int64_t nx,
                             int64_t ns2, int32_t *__restrict__ ep, int32_t *__restrict__ eq,
                             int32_t *__restrict__ etile) {
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  const int64_t e = (int64_t)eid[pos] - 1;
  const int32_t t = perm[pos];
  const int sl = slot[t];
  ep[e * 8 + (sl >> 3)] = (int32_t)(p[t] / ns2);
  eq[e * 8 + (sl & 7)] = (int32_t)(q[t] / ns2);
  if (pos == 0 || eid[pos - 1] != eid[pos]) etile[e] = (int32_t)(key[pos] / (uint64_t)nx);
}

__global__ void k_entry_mask(const int32_t *__restrict__ ep, const int32_t *__restrict__ eq, int64_t ne,
                             uint32_t *__restrict__ mask, int64_t *__restrict__ cnt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  uint32_t m = 0;
  for (int u = 0; u < 8; ++u) {
    if (ep[e * 8 + u] >= 0) m |= 1u << u;
    if (eq[e * 8 + u] >= 0) m |= 1u << (8 + u);
  }
  mask[e] = m;
  cnt[e] = __popc(m);
}

// Operand list of each entry: present left blocks by row, then right blocks by column.
__global__ void k_list_fill(const int32_t *__restrict__ ep, const int32_t *__restrict__ eq,
                            const int64_t *__restrict__ lbeg, int64_t ne, int32_t *__restrict__ list) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int64_t o = lbeg[e];
  for (int u = 0; u < 8; ++u)
    if (ep[e * 8 + u] >= 0) list[o++] = ep[e * 8 + u];
  for (int u = 0; u < 8; ++u)
    if (eq[e * 8 + u] >= 0) list[o++] = (int32_t)((uint32_t)eq[e * 8 + u] | 0x80000000u);
}

// tebeg[t] = first entry of tile t (entries are sorted by tile), t = 0..n_tiles.
__global__ void k_tile_ebeg(const int32_t *__restrict__ etile, int64_t ne, int32_t n_tiles,
                            int32_t *__restrict__ tebeg) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  int64_t lo = 0, hi = ne;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (etile[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  tebeg[t] = (int32_t)lo;
}

// Wavefront products of entry e (tile t): rows u of the entry's left blocks x
// columns of its right blocks whose output block exists (owner != 0xff);
// owner byte = wavefront * kSlots + slot.
__device__ inline void entry_wave_counts(uint32_t m, const uint8_t *__restrict__ own, int pw[kWaves]) {
#pragma unroll
  for (int w = 0; w < kWaves; ++w) pw[w] = 0;
  for (uint32_t ym = m & 0xffu; ym; ym &= ym - 1u) {
    const int u = __builtin_ctz(ym);
    for (uint32_t bm = (m >> 8) & 0xffu; bm; bm &= bm - 1u) {
      const uint8_t o = own[u * 8 + __builtin_ctz(bm)];
      if (o != 0xff) ++pw[o / kSlots];
    }
  }
}

// Chunks of a tile: at most max_entries consecutive entries whose operand
// blocks fit `cap` and (with an owner table) that give no wavefront more than
// kWaveProducts products. With FILL unset, counts them; with FILL set, writes each chunk's
// first entry.
template <bool FILL>
__global__ void k_chunks(const int32_t *__restrict__ tebeg, const uint32_t *__restrict__ mask,
                         const uint8_t *__restrict__ owner, int32_t n_tiles, int32_t cap, int32_t max_entries,
                         const int32_t *__restrict__ tchunk, int32_t *__restrict__ out) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  int32_t nch = 0, sum = 0, ne = 0;
  int wc[kWaves] = {};
  for (int32_t e = tebeg[t]; e < tebeg[t + 1]; ++e) {
    const uint32_t m = mask[e];
    const int32_t c = __popc(m);
    int pw[kWaves] = {};
    if (owner) entry_wave_counts(m, owner + (int64_t)t * 64, pw);  // no owner: no per-wavefront limit
    bool over = false;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) over |= wc[w] + pw[w] > kWaveProducts;
    if (nch == 0 || sum + c > cap || ne == max_entries || over) {
      if (FILL) out[tchunk[t] + nch] = e;
      ++nch;
      sum = 0;
      ne = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) wc[w] = 0;
    }
    sum += c;
    ++ne;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wc[w] += pw[w];
  }
  if (!FILL) out[t] = nch;
}

// Tile of each chunk.
__global__ void k_chunk_tile(const int32_t *__restrict__ tchunk, int32_t n_tiles, int32_t *__restrict__ ctile) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  for (int32_t c = tchunk[t]; c < tchunk[t + 1]; ++c) ctile[c] = t;
}

// Product lists of each chunk, sorted by (wavefront, slot, entry). Count pass
// (FILL unset): products per chunk into np. Fill pass: with pbase = the
// chunk's first product, writes pstart/pcnt of its wavefronts and the packed
// LDS offsets (left | right << 16, elements) of every product.
template <bool FILL>
__global__ void k_chunk_prods(int32_t nch, const int32_t *__restrict__ chunk_entry, const int32_t *__restrict__ ctile,
                              const uint32_t *__restrict__ mask, const int32_t *__restrict__ coff,
                              const uint8_t *__restrict__ owner, int32_t ns2, int64_t *__restrict__ np,
                              const int64_t *__restrict__ pbase, int64_t *__restrict__ pstart,
                              uint64_t *__restrict__ pcnt, uint32_t *__restrict__ prod) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const uint8_t *own = owner + (int64_t)ctile[c] * 64;
  const int32_t e0 = chunk_entry[c], e1 = chunk_entry[c + 1];
  uint8_t cnt[64];
  for (int i = 0; i < 64; ++i) cnt[i] = 0;
  int32_t tot = 0;
  for (int32_t e = e0; e < e1; ++e) {
    const uint32_t m = mask[e];
    for (uint32_t ym = m & 0xffu; ym; ym &= ym - 1u) {
      const int u = __builtin_ctz(ym);
      for (uint32_t bm = (m >> 8) & 0xffu; bm; bm &= bm - 1u) {
        const uint8_t o = own[u * 8 + __builtin_ctz(bm)];
        if (o != 0xff) {
          ++cnt[o];
          ++tot;
        }
      }
    }
  }
  if (!FILL) {
    np[c] = tot;
    return;
  }
  uint16_t off[64];
  int32_t run = 0;
  for (int w = 0; w < kWaves; ++w) {
    pstart[(int64_t)c * kWaves + w] = pbase[c] + run;
    uint64_t word = 0;
    for (int l = 0; l < kSlots; ++l) {
      off[w * kSlots + l] = (uint16_t)run;
      run += cnt[w * kSlots + l];
      word |= (uint64_t)cnt[w * kSlots + l] << (4 * l);
    }
    pcnt[(int64_t)c * kWaves + w] = word;
  }
  for (int32_t e = e0; e < e1; ++e) {
    const uint32_t m = mask[e];
    const uint32_t ym = m & 0xffu, bmask = (m >> 8) & 0xffu;
    const int32_t base = coff[e];
    for (uint32_t yb = ym; yb; yb &= yb - 1u) {
      const int u = __builtin_ctz(yb);
      const uint32_t aoff = (uint32_t)((base + __popc(ym & ((1u << u) - 1u))) * ns2);
      for (uint32_t bm = bmask; bm; bm &= bm - 1u) {
        const int col = __builtin_ctz(bm);
        const uint8_t o = own[u * 8 + col];
        if (o == 0xff) continue;
        const uint32_t boff = (uint32_t)((base + __popc(ym) + __popc(bmask & ((1u << col) - 1u))) * ns2);
        prod[pbase[c] + off[o]++] = aoff | (boff << 16);
      }
    }
  }
}

template <class T>
int inclusive_scan(const T *in, T *out, size_t n, hipStream_t s) {
  if (n == 0) return CQMM_OK;
  size_t bytes = 0;
  hipError_t e = rocprim::inclusive_scan(nullptr, bytes, in, out, n, rocprim::plus<T>(), s);
  if (e != hipSuccess) return cq::hip_check(e, "inclusive_scan (size)");
  cq::DevBuf tmp;
  int st = tmp.alloc(bytes);
  if (st) return st;
  e = rocprim::inclusive_scan(tmp.p, bytes, in, out, n, rocprim::plus<T>(), s);
  if (e != hipSuccess) return cq::hip_check(e, "inclusive_scan");
  // tmp is freed on return: wait for the scan that uses it.
  return cq::hip_check(hipStreamSynchronize(s), "inclusive_scan");
}

// Kernel family of a tile schedule: schedules of one family (with the same
// precision mode and operand transposition) can share a grouped launch.
enum TileKind {
  kKindDense4 = 1,
  kKindDenseMx = 2,
  kKindMixedSmall = 3,
  kKindMixedLarge = 4,
  kKindQuadSmall = 5,
  kKindQuadLarge = 6,
  kKindUniform = 16
};
int tile_kind(const cq::TilePlan &tp);

// ------------------------------------------------ mixed block sizes
// Same schedule with per-atom block dimensions (up to 16): operands are
// addressed by element offset and packed in LDS by size; each product word
// carries its contraction dimension; the k-steps of a product (1-4) and the
// lane offsets follow the dimensions of its blocks.

// Operand elements a chunk stage holds: STG_SMALL for blocks up to 13x13
// (three workgroups per CU), STG_LARGE up to 16x16 (an entry brings at most
// 16 blocks, so it always fits).
constexpr int kStgSmall = 3072;
constexpr int kStgLarge = 4096;
int stage_elems(int maxn) { return maxn <= 13 ? kStgSmall : kStgLarge; }

__global__ void k_entry_fill_mx(const uint64_t *__restrict__ key, const int32_t *__restrict__ perm,
                                const int32_t *__restrict__ eid, const uint8_t *__restrict__ slot,
                                const int64_t *__restrict__ p, const int64_t *__restrict__ q, int64_t n, int64_t nx,
                                int64_t *__restrict__ ep, int64_t *__restrict__ eq, int32_t *__restrict__ etile,
                                int32_t *__restrict__ ex) {
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  const int64_t e = (int64_t)eid[pos] - 1;
  const int32_t t = perm[pos];
  const int sl = slot[t];
  ep[e * 8 + (sl >> 3)] = p[t];
  eq[e * 8 + (sl & 7)] = q[t];
  if (pos == 0 || eid[pos - 1] != eid[pos]) {
    etile[e] = (int32_t)(key[pos] / (uint64_t)nx);
    ex[e] = (int32_t)(key[pos] % (uint64_t)nx);
  }
}

// Mask (left rows | right columns << 8 | (contraction dimension - 1) << 16),
// operand elements and operand count of each entry.
__global__ void k_entry_mask_mx(const int64_t *__restrict__ ep, const int64_t *__restrict__ eq,
                                const int32_t *__restrict__ etile, const int32_t *__restrict__ ex,
                                const uint8_t *__restrict__ rdim, const uint8_t *__restrict__ cdim,
                                const uint8_t *__restrict__ xdim, int64_t ne, uint32_t *__restrict__ mask,
                                int32_t *__restrict__ esz, int64_t *__restrict__ cnt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int64_t t = etile[e];
  const int32_t nxd = xdim[ex[e]];
  uint32_t m = 0;
  int32_t sz = 0;
  for (int u = 0; u < 8; ++u) {
    if (ep[e * 8 + u] >= 0) {
      m |= 1u << u;
      sz += rdim[t * 8 + u] * nxd;
    }
    if (eq[e * 8 + u] >= 0) {
      m |= 1u << (8 + u);
      sz += nxd * cdim[t * 8 + u];
    }
  }
  mask[e] = m | (uint32_t)(nxd - 1) << 16;  // contraction dimension in bits 16-19
  esz[e] = sz;
  cnt[e] = __popc(m);
}

// Operand list of each entry (left blocks by row, then right blocks by
// column): element offsets (bit 63: right operand) and element counts.
__global__ void k_list_fill_mx(const int64_t *__restrict__ ep, const int64_t *__restrict__ eq,
                               const int32_t *__restrict__ etile, const int32_t *__restrict__ ex,
                               const uint8_t *__restrict__ rdim, const uint8_t *__restrict__ cdim,
                               const uint8_t *__restrict__ xdim, const int64_t *__restrict__ lbeg, int64_t ne,
                               uint64_t *__restrict__ list, uint16_t *__restrict__ lsz) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int64_t t = etile[e];
  const int32_t nxd = xdim[ex[e]];
  int64_t o = lbeg[e];
  for (int u = 0; u < 8; ++u)
    if (ep[e * 8 + u] >= 0) {
      list[o] = (uint64_t)ep[e * 8 + u];
      lsz[o++] = (uint16_t)(rdim[t * 8 + u] * nxd);
    }
  for (int u = 0; u < 8; ++u)
    if (eq[e * 8 + u] >= 0) {
      list[o] = (uint64_t)eq[e * 8 + u] | (1ull << 63);
      lsz[o++] = (uint16_t)(nxd * cdim[t * 8 + u]);
    }
}

// Chunks of a tile by operand elements (at most cap), entries, operand
// blocks (at most kWaves * 64) and wavefront products (as k_chunks).
template <bool FILL>
__global__ void k_chunks_mx(const int32_t *__restrict__ tebeg, const uint32_t *__restrict__ mask,
                            const int32_t *__restrict__ esz, const uint8_t *__restrict__ owner, int32_t n_tiles,
                            int32_t cap, int32_t max_entries, const int32_t *__restrict__ tchunk,
                            int32_t *__restrict__ out) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  int32_t nch = 0, sum = 0, ne = 0, nb = 0;
  int wc[kWaves] = {};
  for (int32_t e = tebeg[t]; e < tebeg[t + 1]; ++e) {
    const uint32_t m = mask[e];
    const int32_t c = esz[e];
    const int32_t b = __popc(m & 0xffffu);
    int pw[kWaves] = {};
    if (owner) entry_wave_counts(m, owner + (int64_t)t * 64, pw);  // no owner: no per-wavefront limit
    bool over = false;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) over |= wc[w] + pw[w] > kWaveProducts;
    // A chunk's operand blocks are spread over the lanes of the wavefronts'
    // operand lists: at most kWaves * 64 of them.
    if (nch == 0 || sum + c > cap || ne == max_entries || nb + b > kWaves * 64 || over) {
      if (FILL) out[tchunk[t] + nch] = e;
      ++nch;
      sum = 0;
      ne = 0;
      nb = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) wc[w] = 0;
    }
    sum += c;
    nb += b;
    ++ne;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wc[w] += pw[w];
  }
  if (!FILL) out[t] = nch;
}

// First LDS element of each entry in its chunk's stage.
__global__ void k_entry_coff_mx(const int32_t *__restrict__ chunk_entry, const int32_t *__restrict__ esz, int64_t nc,
                                int32_t *__restrict__ coff) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int32_t run = 0;
  for (int32_t e = chunk_entry[c]; e < chunk_entry[c + 1]; ++e) {
    coff[e] = run;
    run += esz[e];
  }
}

// LDS element offset | elements << 16 of each operand of each chunk.
__global__ void k_list_lds_mx(const int64_t *__restrict__ chunk_lbeg, const uint16_t *__restrict__ lsz, int64_t nc,
                              uint32_t *__restrict__ listls) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  uint32_t run = 0;
  for (int64_t o = chunk_lbeg[c]; o < chunk_lbeg[c + 1]; ++o) {
    listls[o] = run | ((uint32_t)lsz[o] << 16);
    run += lsz[o];
  }
}

// Product lists (as k_chunk_prods): word = left LDS offset | right LDS offset
// << 12 | (contraction dimension nx - 1) << 24 | (k-steps - 1) << 28 |
// (contraction indices of the last k-step - 1) << 30.
template <bool FILL>
__global__ void k_chunk_prods_mx(int32_t nch, const int32_t *__restrict__ chunk_entry,
                                 const int32_t *__restrict__ ctile, const uint32_t *__restrict__ mask,
                                 const int32_t *__restrict__ coff, const int32_t *__restrict__ ex,
                                 const uint8_t *__restrict__ owner, const uint8_t *__restrict__ rdim,
                                 const uint8_t *__restrict__ cdim, const uint8_t *__restrict__ xdim,
                                 int64_t *__restrict__ np, const int64_t *__restrict__ pbase,
                                 int64_t *__restrict__ pstart, uint64_t *__restrict__ pcnt,
                                 uint32_t *__restrict__ prod) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const int64_t t = ctile[c];
  const uint8_t *own = owner + t * 64;
  const int32_t e0 = chunk_entry[c], e1 = chunk_entry[c + 1];
  uint8_t cnt[64];
  for (int i = 0; i < 64; ++i) cnt[i] = 0;
  int32_t tot = 0;
  for (int32_t e = e0; e < e1; ++e) {
    const uint32_t m = mask[e];
    for (uint32_t ym = m & 0xffu; ym; ym &= ym - 1u) {
      const int u = __builtin_ctz(ym);
      for (uint32_t bm = (m >> 8) & 0xffu; bm; bm &= bm - 1u) {
        const uint8_t o = own[u * 8 + __builtin_ctz(bm)];
        if (o != 0xff) {
          ++cnt[o];
          ++tot;
        }
      }
    }
  }
  if (!FILL) {
    np[c] = tot;
    return;
  }
  uint16_t off[64];
  int32_t run = 0;
  for (int w = 0; w < kWaves; ++w) {
    pstart[(int64_t)c * kWaves + w] = pbase[c] + run;
    uint64_t word = 0;
    for (int l = 0; l < kSlots; ++l) {
      off[w * kSlots + l] = (uint16_t)run;
      run += cnt[w * kSlots + l];
      word |= (uint64_t)cnt[w * kSlots + l] << (4 * l);
    }
    pcnt[(int64_t)c * kWaves + w] = word;
  }
  for (int32_t e = e0; e < e1; ++e) {
    const uint32_t m = mask[e];
    const uint32_t ym = m & 0xffu, bmask = (m >> 8) & 0xffu;
    const uint32_t nxd = xdim[ex[e]];
    uint32_t rows = 0;  // left elements of the entry
    for (uint32_t yb = ym; yb; yb &= yb - 1u) rows += rdim[t * 8 + __builtin_ctz(yb)] * nxd;
    uint32_t aoff = (uint32_t)coff[e];
    for (uint32_t yb = ym; yb; yb &= yb - 1u) {
      const int u = __builtin_ctz(yb);
      uint32_t boff = (uint32_t)coff[e] + rows;
      for (uint32_t bm = bmask; bm; bm &= bm - 1u) {
        const int col = __builtin_ctz(bm);
        const uint8_t o = own[u * 8 + col];
        if (o != 0xff) {
          const uint32_t ks = (nxd + 3u) >> 2, rem = nxd - 4u * (ks - 1u);
          prod[pbase[c] + off[o]++] = aoff | (boff << 12) | ((nxd - 1u) << 24) | ((ks - 1u) << 28) | ((rem - 1u) << 30);
        }
        boff += nxd * cdim[t * 8 + col];
      }
      aoff += rdim[t * 8 + u] * nxd;
    }
  }
}

// Copies of one operand block (up to 16x16) with one M0: pieces I < NP of
// 256 bytes, the last masked to the block's dwords.
template <int... I>
__device__ inline void dma_block_pieces(const uint32_t *src, uint32_t *dst, int dwords, int lane,
                                        std::integer_sequence<int, I...>) {
  ((I * 64 < dwords ? (I * 64 + lane < dwords ? (dma_piece<I * 256>(src, dst), 0) : 0) : 0), ...);
}

template <class T>
__device__ inline void issue_stage_mx(uint64_t lv, uint32_t lls, int32_t nb, const T *__restrict__ P,
                                      const T *__restrict__ Q, T *stage, int w, int lane) {
  constexpr int MAXP = 256 * (int)sizeof(T) / 4 / 64;  // pieces of a 16x16 block
  const int32_t nbw = nb > w ? (nb - w + kWaves - 1) / kWaves : 0;
  for (int32_t k = 0; k < nbw; ++k) {
    const uint64_t g = (uint64_t)readlane64((int64_t)lv, k);
    const uint32_t ls = (uint32_t)__builtin_amdgcn_readlane((int)lls, k);
    const T *base = (g >> 63) ? Q : P;
    const uint32_t *src = (const uint32_t *)(base + (int64_t)(g & 0x7fffffffffffffffull)) + lane;
    uint32_t *dst = (uint32_t *)(stage + (ls & 0xffffu));
    const int dwords = (int)(ls >> 16) * (int)sizeof(T) / 4;
    dma_block_pieces(src, dst, dwords, lane, std::make_integer_sequence<int, MAXP>{});
  }
}

// Mixed-size tile kernel: as k_tile_products with per-slot output dimensions
// (wave_dim) and per-product contraction dimension (product word bits 24-27).
template <int MODE, bool QT, int STG, class ARG>
__global__ void __launch_bounds__(kTileThreads) k_tile_mixed(
    const ARG arg) {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  constexpr int STAGE = STG + kLdsPad * (int)(8 / sizeof(T));
  __shared__ T stage0[STAGE];
  __shared__ T stage1[STAGE];
  TileArgs a;
  const int32_t tile = tile_select(arg, a);
  if (tile < 0) return;  // the whole workgroup leaves together
  const int32_t *__restrict__ tile_chunk = a.tile_chunk;
  const int64_t *__restrict__ chunk_lbeg = a.chunk_lbeg;
  const uint64_t *__restrict__ list64 = a.list64;
  const uint32_t *__restrict__ listls = a.listls;
  const int64_t *__restrict__ pstart = a.pstart;
  const uint64_t *__restrict__ pcnt = a.pcnt;
  const uint32_t *__restrict__ prod = a.prod;
  const int64_t *__restrict__ wave_out = a.wave_out64;
  const uint8_t *__restrict__ wave_dim = a.wave_dim;
  const T *__restrict__ P = static_cast<const T *>(a.P);
  const T *__restrict__ Q = static_cast<const T *>(a.Q);
  T *__restrict__ O = static_cast<T *>(a.O);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, r = lane & 15, g = lane >> 4;
  const int64_t wbase = ((int64_t)tile * kWaves + w) * kSlots;
  int64_t oblk[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) oblk[u] = wave_out[wbase + u];
  const uint32_t *wd32 = (const uint32_t *)(wave_dim + wbase);
  uint32_t wdim[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) wdim[k] = wd32[k];
  dbl4 a64[kSlots];
  flt4 a32[kSlots];
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    a64[u] = dbl4{0.0, 0.0, 0.0, 0.0};
    a32[u] = flt4{0.f, 0.f, 0.f, 0.f};
  }
  const int32_t c0 = tile_chunk[tile], c1 = tile_chunk[tile + 1];
  if (c0 < c1) {
    int32_t wb = c0;
    int32_t cc = min(wb + lane, c1);
    int64_t wcl = chunk_lbeg[cc];
    int64_t wps = pstart[(int64_t)cc * kWaves + w];
    int64_t wpc = (int64_t)pcnt[(int64_t)cc * kWaves + w];
    auto cl_at = [&](int32_t x) { return readlane64(wcl, min(x, c1) - wb); };
    auto ps_at = [&](int32_t x) { return readlane64(wps, min(x, c1) - wb); };
    auto pc_at = [&](int32_t x) { return (uint64_t)readlane64(wpc, min(x, c1) - wb); };
    auto list_idx = [&](int64_t lbeg, int32_t nb) { return lbeg + min(w + kWaves * lane, nb - 1); };
    uint32_t pl;
    uint64_t lvn = 0;
    uint32_t lsn = 0;
    {
      const int64_t cl0 = cl_at(c0), cl1 = cl_at(c0 + 1);
      const int64_t i0 = list_idx(cl0, (int32_t)(cl1 - cl0));
      const uint64_t lv0 = list64[i0];
      const uint32_t ls0 = listls[i0];
      issue_stage_mx<T>(lv0, ls0, (int32_t)(cl1 - cl0), P, Q, stage0, w, lane);
      pl = prod[ps_at(c0) + lane];
      if (c0 + 1 < c1) {
        const int64_t i1 = list_idx(cl1, (int32_t)(cl_at(c0 + 2) - cl1));
        lvn = list64[i1];
        lsn = listls[i1];
      }
    }
    for (int32_t c = c0; c < c1; ++c) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      const bool odd = ((c - c0) & 1) != 0;
      const T *cur = odd ? stage1 : stage0;
      const uint32_t plc = pl;
      const uint64_t cnt = pc_at(c);
      if (c + 1 < c1) {
        const int64_t cl1 = cl_at(c + 1), cl2 = cl_at(c + 2);
        issue_stage_mx<T>(lvn, lsn, (int32_t)(cl2 - cl1), P, Q, odd ? stage0 : stage1, w, lane);
        pl = prod[ps_at(c + 1) + lane];
        if (c + 2 < c1) {
          const int64_t i2 = list_idx(cl2, (int32_t)(cl_at(c + 3) - cl2));
          lvn = list64[i2];
          lsn = listls[i2];
        }
      }
      // Lane coordinates made opaque per chunk, so the per-slot lane offsets
      // below are computed where used instead of being kept for all 16 slots
      // across the chunk loop (which costs 32 registers).
      int lr = r, lg = g;
      asm volatile("" : "+v"(lr), "+v"(lg));
      int p = 0;
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int n = (int)((cnt >> (4 * u)) & 15u);
        if (n == 0) continue;
        const uint32_t dm = (wdim[u >> 2] >> (8 * (u & 3))) & 0xffu;
        const int nrow = (int)(dm & 15u) + 1, ncol = (int)(dm >> 4) + 1;
        // Lanes past the block's rows (left) or columns (right) read the last
        // valid row/column (a broadcast, no bank conflicts); they feed only
        // accumulator rows/columns that are never written.
        const int rA = min(lr, nrow - 1), rB = min(lr, ncol - 1);
        const int oA = rA + nrow * lg, SA = 4 * nrow;
        for (int i = 0; i < n; ++i, ++p) {
          const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)plc, p);
          const int nx = (int)((v >> 24) & 15u) + 1;
          const int ks1 = (int)((v >> 28) & 3u), rem = (int)(v >> 30) + 1;
          const T *la = cur + (int)(v & 0xfffu) + oA;
          const T *lq = cur + (int)((v >> 12) & 0xfffu) + (QT ? rB + ncol * lg : lg + nx * rB);
          const int SB = QT ? 4 * ncol : 4;
          flt4 t = {0.f, 0.f, 0.f, 0.f};
          // Full k-steps, then the last one (contraction indices >= nx zeroed).
#pragma unroll
          for (int s4 = 0; s4 < 3; ++s4) {
            if (s4 < ks1) {
              const T a = la[SA * s4], bb = lq[SB * s4];
              if constexpr (MODE == 0) a64[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, a64[u], 0, 0, 0);
              else if constexpr (MODE == 1) a32[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, a32[u], 0, 0, 0);
              else t = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, t, 0, 0, 0);
            }
          }
          {
            T a = la[SA * ks1], bb = lq[SB * ks1];
            if (lg >= rem) {
              a = (T)0;
              bb = (T)0;
            }
            if constexpr (MODE == 0) a64[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, a64[u], 0, 0, 0);
            else if constexpr (MODE == 1) a32[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, a32[u], 0, 0, 0);
            else t = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, t, 0, 0, 0);
          }
          if constexpr (MODE == 2) {
#pragma unroll
            for (int e = 0; e < 4; ++e) a64[u][e] += (double)t[e];
          }
        }
      }
      if (c + 4 > wb + 63 && c + 1 < c1) {
        wb = c + 1;
        cc = min(wb + lane, c1);
        wcl = chunk_lbeg[cc];
        wps = pstart[(int64_t)cc * kWaves + w];
        wpc = (int64_t)pcnt[(int64_t)cc * kWaves + w];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kSlots; ++u) {
    if (oblk[u] < 0) continue;
    const uint32_t dm = (wdim[u >> 2] >> (8 * (u & 3))) & 0xffu;
    const int nrow = (int)(dm & 15u) + 1, ncol = (int)(dm >> 4) + 1;
    if (r >= ncol) continue;
    T *o = O + oblk[u];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = MODE == 0 ? g + 4 * e : 4 * g + e;
      if (row < nrow) o[row + nrow * r] = MODE == 0 ? (T)a64[u][e] : (MODE == 1 ? (T)a32[u][e] : (T)a64[u][e]);
    }
  }
}

// Mixed block sizes of at most 4x4 (e.g. C and H with single-zeta bases):
// the dense-quadrant scheme of k_tile_dense4 over the mixed-size schedule.
// Wavefront w owns quadrant (w / 2, w % 2) of the tile (4x4 blocks, each
// inside a 4x4 corner of the quadrant's 16x16 MFMA tile) and issues one
// 16x16x4 MFMA per entry: lane (r, g) supplies element (r % 4, g) of
// P(row 4*(w/2) + r/4, x) and element (g, r % 4) of Q(column 4*(w%2) + r/4, x),
// zero outside the blocks' shapes or where the entry lacks the block. A
// block's LDS offset within its entry is the entry offset plus nx times the
// rows (columns) of the entry's blocks before it, summed from the tile's
// block-shape bit masks with three popcounts.
template <int MODE, bool QT, class ARG>
__global__ void __launch_bounds__(kTileThreads) k_tile_dense_mx(
    const ARG arg) {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  constexpr int STAGE = kStgSmall + kLdsPad * (int)(8 / sizeof(T));
  __shared__ T stage0[STAGE];
  __shared__ T stage1[STAGE];
  TileArgs a;
  const int32_t tile = tile_select(arg, a);
  if (tile < 0) return;  // the whole workgroup leaves together
  const int32_t *__restrict__ tile_chunk = a.tile_chunk;
  const int64_t *__restrict__ chunk_lbeg = a.chunk_lbeg;
  const uint64_t *__restrict__ list64 = a.list64;
  const uint32_t *__restrict__ listls = a.listls;
  const int32_t *__restrict__ chunk_entry = a.chunk_entry;
  const uint32_t *__restrict__ emask = a.emask;
  const int32_t *__restrict__ ecoff = a.ecoff;
  const uint8_t *__restrict__ rdim = a.rdim;
  const uint8_t *__restrict__ cdim = a.cdim;
  const int64_t *__restrict__ wave_out = a.wave_out64;
  const T *__restrict__ P = static_cast<const T *>(a.P);
  const T *__restrict__ Q = static_cast<const T *>(a.Q);
  T *__restrict__ O = static_cast<T *>(a.O);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, r = lane & 15, g = lane >> 4, rr = r & 3;
  const int qr = w >> 1, qc = w & 1;
  const int ua = 4 * qr + (r >> 2), cb = 4 * qc + (r >> 2);
  // Block shapes of the tile: this lane's row and column, and bit masks of the
  // rows (columns) whose size has bit 0, 1 or 2 set.
  const int rd = rdim[(int64_t)tile * 8 + ua], cd = cdim[(int64_t)tile * 8 + cb];
  uint32_t rb0 = 0, rb1 = 0, rb2 = 0, cb0 = 0, cb1 = 0, cb2 = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t a = rdim[(int64_t)tile * 8 + k], c = cdim[(int64_t)tile * 8 + k];
    rb0 |= (a & 1u) << k;
    rb1 |= ((a >> 1) & 1u) << k;
    rb2 |= ((a >> 2) & 1u) << k;
    cb0 |= (c & 1u) << k;
    cb1 |= ((c >> 1) & 1u) << k;
    cb2 |= ((c >> 2) & 1u) << k;
  }
  auto rows_of = [&](uint32_t m) { return __popc(m & rb0) + 2 * __popc(m & rb1) + 4 * __popc(m & rb2); };
  auto cols_of = [&](uint32_t m) { return __popc(m & cb0) + 2 * __popc(m & cb1) + 4 * __popc(m & cb2); };
  const uint32_t below_a = (1u << ua) - 1u, below_b = (1u << cb) - 1u;
  const uint32_t quad_rows = 0xfu << (4 * qr), quad_cols = 0xfu << (8 + 4 * qc);
  dbl4 a64 = {0.0, 0.0, 0.0, 0.0};
  flt4 a32 = {0.f, 0.f, 0.f, 0.f};
  const int32_t c0 = tile_chunk[tile], c1 = tile_chunk[tile + 1];
  if (c0 < c1) {
    int32_t wb = c0;
    int32_t cc = min(wb + lane, c1);
    int64_t wcl = chunk_lbeg[cc];
    int32_t wce = chunk_entry[cc];
    auto cl_at = [&](int32_t x) { return readlane64(wcl, min(x, c1) - wb); };
    auto ce_at = [&](int32_t x) { return __builtin_amdgcn_readlane(wce, min(x, c1) - wb); };
    auto list_idx = [&](int64_t lbeg, int32_t nb) { return lbeg + min(w + kWaves * lane, nb - 1); };
    auto entry_regs = [&](int32_t e0, int32_t e1, uint32_t &m, int32_t &o) {
      const int32_t e = min(e0 + lane, e1 - 1);
      m = emask[e];
      o = ecoff[e];
    };
    uint32_t em, emn = 0;
    int32_t eo, eon = 0;
    uint64_t lvn = 0;
    uint32_t lsn = 0;
    {
      const int64_t cl0 = cl_at(c0), cl1 = cl_at(c0 + 1);
      const int64_t i0 = list_idx(cl0, (int32_t)(cl1 - cl0));
      issue_stage_mx<T>(list64[i0], listls[i0], (int32_t)(cl1 - cl0), P, Q, stage0, w, lane);
      entry_regs(ce_at(c0), ce_at(c0 + 1), emn, eon);
      if (c0 + 1 < c1) {
        const int64_t i1 = list_idx(cl1, (int32_t)(cl_at(c0 + 2) - cl1));
        lvn = list64[i1];
        lsn = listls[i1];
      }
    }
    for (int32_t c = c0; c < c1; ++c) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      const bool odd = ((c - c0) & 1) != 0;
      const T *cur = odd ? stage1 : stage0;
      em = emn;
      eo = eon;
      const int32_t ne = ce_at(c + 1) - ce_at(c);
      if (c + 1 < c1) {
        const int64_t cl1 = cl_at(c + 1), cl2 = cl_at(c + 2);
        issue_stage_mx<T>(lvn, lsn, (int32_t)(cl2 - cl1), P, Q, odd ? stage0 : stage1, w, lane);
        entry_regs(ce_at(c + 1), ce_at(c + 2), emn, eon);
        if (c + 2 < c1) {
          const int64_t i2 = list_idx(cl2, (int32_t)(cl_at(c + 3) - cl2));
          lvn = list64[i2];
          lsn = listls[i2];
        }
      }
      for (int i = 0; i < ne; ++i) {
        const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)em, i);
        if ((m & quad_rows) == 0 || (m & quad_cols) == 0) continue;  // no block pair in this quadrant
        const int32_t base = __builtin_amdgcn_readlane(eo, i);
        const uint32_t ym = m & 0xffu, bm = (m >> 8) & 0xffu;
        const int nx = (int)((m >> 16) & 15u) + 1;
        const bool va = ((ym >> ua) & 1u) != 0 && rr < rd && g < nx;
        const bool vb = ((bm >> cb) & 1u) != 0 && rr < cd && g < nx;
        const int pa = base + nx * rows_of(ym & below_a);
        const int pb = base + nx * (rows_of(ym) + cols_of(bm & below_b));
        const int ia = va ? pa + rr + rd * g : 0;
        const int ib = vb ? (QT ? pb + rr + cd * g : pb + g + nx * rr) : 0;
        T a = cur[ia], b = cur[ib];
        if (!va) a = (T)0;
        if (!vb) b = (T)0;
        if constexpr (MODE == 0) {
          a64 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, a64, 0, 0, 0);
        } else if constexpr (MODE == 1) {
          a32 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, a32, 0, 0, 0);
        } else {
          const flt4 t = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, flt4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
          for (int e = 0; e < 4; ++e) a64[e] += (double)t[e];
        }
      }
      if (c + 4 > wb + 63 && c + 1 < c1) {
        wb = c + 1;
        cc = min(wb + lane, c1);
        wcl = chunk_lbeg[cc];
        wce = chunk_entry[cc];
      }
    }
  }
  // Quadrant row R, column r: element (R % 4, r % 4) of block (R / 4, r / 4);
  // written where it lies inside that block's shape.
  const int64_t *wo = wave_out + ((int64_t)tile * kWaves + w) * 16;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int rq = MODE == 0 ? e : g, ri = MODE == 0 ? g : e;
    const int64_t off = wo[rq * 4 + (r >> 2)];
    const int nr = rdim[(int64_t)tile * 8 + 4 * qr + rq];
    if (off >= 0 && ri < nr && rr < cd) O[off + ri + nr * rr] = MODE == 1 ? (T)a32[e] : (T)a64[e];
  }
}

// Class tiles (species mixes, blocks up to 16x16): the tile's rows all have
// one block size NI and its columns one size NJ (the plan groups tile rows
// and columns by species), so the tile's blocks pack into 16x16-element MFMA
// tiles of RP row blocks x CP column blocks (RP = 4, 2, 1 for NI <= 4, <= 8,
// <= 16; CP likewise): sixteen 4x4 block pairs share one MFMA, four 4x13
// pairs, one 13x13 pair. Wavefront w owns quadrant (w / 2, w % 2) of the tile
// (4 block rows x 4 block columns, 16 / (RP CP) MFMA tiles). Per chunk entry
// (contraction atom x with nx contraction indices; at most 64 entries per
// chunk, one per lane) and k-step of 4 indices, each lane reads one element
// of each of the wavefront's row units and column units from LDS (its block's
// offset from the entry's offset and the popcount of the entry's blocks
// before it) and every MFMA tile with a block pair present accumulates;
// absent blocks, rows or columns past NI / NJ and indices past nx read as
// zero. Each output block sums its products in entry (x) order. The entry
// list replaces per-product lists: no product decode at run time.
template <int MODE, bool QT>
__device__ __forceinline__ void quad_tile(const TileArgs &a, const int32_t tile,
                                          typename std::conditional<MODE == 0, double, float>::type *stage0,
                                          typename std::conditional<MODE == 0, double, float>::type *stage1) {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  const int32_t *__restrict__ tile_chunk = a.tile_chunk;
  const int64_t *__restrict__ chunk_lbeg = a.chunk_lbeg;
  const uint64_t *__restrict__ list64 = a.list64;
  const uint32_t *__restrict__ listls = a.listls;
  const int32_t *__restrict__ chunk_entry = a.chunk_entry;
  const uint32_t *__restrict__ emask = a.emask;
  const int32_t *__restrict__ ecoff = a.ecoff;
  const uint8_t *__restrict__ rdim = a.rdim;
  const uint8_t *__restrict__ cdim = a.cdim;
  const int64_t *__restrict__ wave_out = a.wave_out64;
  const T *__restrict__ P = static_cast<const T *>(a.P);
  const T *__restrict__ Q = static_cast<const T *>(a.Q);
  T *__restrict__ O = static_cast<T *>(a.O);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, r = lane & 15, g = lane >> 4;
  const int qr = w >> 1, qc = w & 1;
  // Block sizes of the tile (rows of absent tile rows read 0).
  int NI = 0, NJ = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    NI = max(NI, (int)rdim[(int64_t)tile * 8 + k]);
    NJ = max(NJ, (int)cdim[(int64_t)tile * 8 + k]);
  }
  NI = __builtin_amdgcn_readfirstlane(NI);
  NJ = __builtin_amdgcn_readfirstlane(NJ);
  // Packing: RP row blocks of RH element rows per MFMA tile (CP, CH for columns).
  const int lrh = NI <= 4 ? 2 : NI <= 8 ? 3 : 4, lch = NJ <= 4 ? 2 : NJ <= 8 ? 3 : 4;
  const int RP = 16 >> lrh, CP = 16 >> lch;
  const int NTA = 4 / RP, NTB = 4 / CP;
  const int ir = r & ((1 << lrh) - 1), jc = r & ((1 << lch) - 1);
  const int ru = r >> lrh, cu = r >> lch;
  const bool rowok = ir < NI, colok = jc < NJ;
  dbl4 a64[16];
  flt4 a32[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    a64[u] = dbl4{0.0, 0.0, 0.0, 0.0};
    a32[u] = flt4{0.f, 0.f, 0.f, 0.f};
  }
  const int32_t c0 = tile_chunk[tile], c1 = tile_chunk[tile + 1];
  if (c0 < c1) {
    int32_t wb = c0;
    int32_t cc = min(wb + lane, c1);
    int64_t wcl = chunk_lbeg[cc];
    int32_t wce = chunk_entry[cc];
    auto cl_at = [&](int32_t x) { return readlane64(wcl, min(x, c1) - wb); };
    auto ce_at = [&](int32_t x) { return __builtin_amdgcn_readlane(wce, min(x, c1) - wb); };
    auto list_idx = [&](int64_t lbeg, int32_t nb) { return lbeg + min(w + kWaves * lane, nb - 1); };
    auto entry_regs = [&](int32_t e0, int32_t e1, uint32_t &m, int32_t &o) {
      const int32_t e = min(e0 + lane, e1 - 1);
      m = emask[e];
      o = ecoff[e];
    };
    uint32_t em, emn = 0;
    int32_t eo, eon = 0;
    uint64_t lvn = 0;
    uint32_t lsn = 0;
    {
      const int64_t cl0 = cl_at(c0), cl1 = cl_at(c0 + 1);
      const int64_t i0 = list_idx(cl0, (int32_t)(cl1 - cl0));
      issue_stage_mx<T>(list64[i0], listls[i0], (int32_t)(cl1 - cl0), P, Q, stage0, w, lane);
      entry_regs(ce_at(c0), ce_at(c0 + 1), emn, eon);
      if (c0 + 1 < c1) {
        const int64_t i1 = list_idx(cl1, (int32_t)(cl_at(c0 + 2) - cl1));
        lvn = list64[i1];
        lsn = listls[i1];
      }
    }
    const uint32_t quad_rows = 0xfu << (4 * qr), quad_cols = 0xfu << (4 * qc);
    for (int32_t c = c0; c < c1; ++c) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      const bool odd = ((c - c0) & 1) != 0;
      const T *cur = odd ? stage1 : stage0;
      em = emn;
      eo = eon;
      const int32_t ne = ce_at(c + 1) - ce_at(c);
      if (c + 1 < c1) {
        const int64_t cl1 = cl_at(c + 1), cl2 = cl_at(c + 2);
        issue_stage_mx<T>(lvn, lsn, (int32_t)(cl2 - cl1), P, Q, odd ? stage0 : stage1, w, lane);
        entry_regs(ce_at(c + 1), ce_at(c + 2), emn, eon);
        if (c + 2 < c1) {
          const int64_t i2 = list_idx(cl2, (int32_t)(cl_at(c + 3) - cl2));
          lvn = list64[i2];
          lsn = listls[i2];
        }
      }
      for (int i = 0; i < ne; ++i) {
        const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)em, i);
        const uint32_t ym = m & 0xffu, bm = (m >> 8) & 0xffu;
        if ((ym & quad_rows) == 0 || (bm & quad_cols) == 0) continue;  // no block pair in this quadrant
        const int32_t base = __builtin_amdgcn_readlane(eo, i);
        const int nx = (int)((m >> 16) & 15u) + 1;
        const int ablk = NI * nx, bblk = nx * NJ;
        const int bbase = base + ablk * __popc(ym);
        if constexpr (MODE == 4) {
          // Split-BF16 x3: one K = 16 step (nx <= 16); lane (r, g) holds
          // contraction indices 4g..4g+3 of its row (column) as four LDS words
          // (bits 0-15 hi, 16-31 lo), repacked into the hi and lo operands of
          // v_mfma_f32_16x16x16_bf16; hi.hi + hi.lo + lo.hi chain in FP32 and
          // each entry's sum is added into the FP64 accumulator.
          s16x4 ahi[4], alo[4], bhi[4], blo[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int u = 4 * qr + t * RP + ru, v = 4 * qc + t * CP + cu;
            const bool oka = t < NTA && rowok && ((ym >> u) & 1u) != 0;
            const bool okb = t < NTB && colok && ((bm >> v) & 1u) != 0;
            const int a0 = base + ablk * __popc(ym & ((1u << u) - 1u)) + ir + NI * 4 * g;
            const int b0 = bbase + bblk * __popc(bm & ((1u << v) - 1u)) + (QT ? jc + NJ * 4 * g : 4 * g + nx * jc);
            uint32_t wa[4], wb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const bool kin = 4 * g + i < nx;
              const bool ra = oka && kin, rb = okb && kin;
              wa[i] = __float_as_uint(cur[ra ? a0 + NI * i : 0]);
              wb[i] = __float_as_uint(cur[rb ? b0 + (QT ? NJ * i : i) : 0]);
              if (!ra) wa[i] = 0u;
              if (!rb) wb[i] = 0u;
            }
            ahi[t] = pack_bf16_lo(wa);
            alo[t] = pack_bf16_hi(wa);
            bhi[t] = pack_bf16_lo(wb);
            blo[t] = pack_bf16_hi(wb);
          }
#pragma unroll
          for (int ta = 0; ta < 4; ++ta) {
            if (ta >= NTA || (ym & (((1u << RP) - 1u) << (4 * qr + ta * RP))) == 0) continue;
#pragma unroll
            for (int tb = 0; tb < 4; ++tb) {
              if (tb >= NTB || (bm & (((1u << CP) - 1u) << (4 * qc + tb * CP))) == 0) continue;
              const int u = ta * 4 + tb;
              flt4 t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[ta], bhi[tb], flt4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
              t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi[ta], blo[tb], t, 0, 0, 0);
              t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo[ta], bhi[tb], t, 0, 0, 0);
#pragma unroll
              for (int e = 0; e < 4; ++e) a64[u][e] += (double)t[e];
            }
          }
          continue;
        }
        // This lane's element of each row unit (ta) and column unit (tb) at k-step 0.
        int pa[4], pb[4];
        bool va[4], vb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int u = 4 * qr + t * RP + ru, v = 4 * qc + t * CP + cu;
          va[t] = t < NTA && rowok && ((ym >> u) & 1u) != 0;
          vb[t] = t < NTB && colok && ((bm >> v) & 1u) != 0;
          pa[t] = base + ablk * __popc(ym & ((1u << u) - 1u)) + ir + NI * g;
          pb[t] = bbase + bblk * __popc(bm & ((1u << v) - 1u)) + (QT ? jc + NJ * g : g + nx * jc);
        }
        const int KS = (nx + 3) >> 2;
        const int sa = 4 * NI, sb = QT ? 4 * NJ : 4;
        for (int s4 = 0; s4 < KS; ++s4) {
          const bool kv = s4 + 1 < KS || 4 * s4 + g < nx;
          T av[4], bv[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const bool oka = va[t] && kv, okb = vb[t] && kv;
            av[t] = cur[oka ? pa[t] + sa * s4 : 0];
            bv[t] = cur[okb ? pb[t] + sb * s4 : 0];
            if (!oka) av[t] = (T)0;
            if (!okb) bv[t] = (T)0;
          }
#pragma unroll
          for (int ta = 0; ta < 4; ++ta) {
            if (ta >= NTA || (ym & (((1u << RP) - 1u) << (4 * qr + ta * RP))) == 0) continue;
#pragma unroll
            for (int tb = 0; tb < 4; ++tb) {
              if (tb >= NTB || (bm & (((1u << CP) - 1u) << (4 * qc + tb * CP))) == 0) continue;
              const int u = ta * 4 + tb;
              if constexpr (MODE == 0) {
                a64[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ta], bv[tb], a64[u], 0, 0, 0);
              } else if constexpr (MODE == 1) {
                a32[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ta], bv[tb], a32[u], 0, 0, 0);
              } else {
                const flt4 t = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ta], bv[tb], flt4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) a64[u][e] += (double)t[e];
              }
            }
          }
        }
      }
      if (c + 4 > wb + 63 && c + 1 < c1) {
        wb = c + 1;
        cc = min(wb + lane, c1);
        wcl = chunk_lbeg[cc];
        wce = chunk_entry[cc];
      }
    }
  }
  // Element (R, r) of MFMA tile (ta, tb): row R % RH of quadrant block row
  // ta RP + R / RH, column r % CH of quadrant block column tb CP + r / CH;
  // written where it lies inside that block's shape.
  const int64_t *wo = wave_out + ((int64_t)tile * kWaves + w) * 16;
#pragma unroll
  for (int ta = 0; ta < 4; ++ta) {
    if (ta >= NTA) continue;
#pragma unroll
    for (int tb = 0; tb < 4; ++tb) {
      if (tb >= NTB) continue;
      const int u = ta * 4 + tb;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // FP64 MFMA accumulators hold D[g + 4e][r]; FP32 ones (modes 1 and 2,
        // whose FP64 sums add FP32 results element for element) D[4g + e][r].
        const int R = MODE == 0 ? g + 4 * e : 4 * g + e;
        const int bu = ta * RP + (R >> lrh), rowi = R & ((1 << lrh) - 1);
        const int bv = tb * CP + cu;
        const int64_t off = wo[bu * 4 + bv];
        if (off >= 0 && rowi < NI && colok) O[off + rowi + (int64_t)NI * jc] = MODE == 1 ? (T)a32[u][e] : (T)a64[u][e];
      }
    }
  }
}

// Global tile number of a launch -> the product's arguments and its tile.
__device__ inline int32_t tile_at(const TileArgs &arg, int32_t t, TileArgs &a) {
  a = arg;
  return t;
}
__device__ inline int32_t tile_at(const TileGroup &g, int32_t t, TileArgs &a) {
  int32_t j = 0;
#pragma unroll
  for (int k = 1; k < kMaxGroup; ++k) j += (k < g.ngrp && g.a[k].tile_base <= t) ? 1 : 0;
  a = g.a[j];
  return t - a.tile_base;
}
__device__ inline int32_t total_tiles(const TileArgs &a) { return a.n_tiles; }
__device__ inline int32_t total_tiles(const TileGroup &g) { return g.total; }

// Persistent launch of class tiles: the grid (a multiple of 8 workgroups, at
// most one per tile) splits the tiles into 8 contiguous ranges, one per XCD
// (workgroup b runs on XCD b % 8), and each workgroup takes every (grid/8)-th
// tile of its XCD's range, so neighbouring tiles (same rows) share an L2.
// With one workgroup per tile this is the one-tile-per-workgroup order of the
// other tile kernels.
template <int MODE, bool QT, int STG, class ARG>
__global__ void __launch_bounds__(kTileThreads) k_tile_quad(const ARG arg) {
  using T = typename std::conditional<MODE == 0, double, float>::type;
  constexpr int STAGE = STG + kLdsPad * (int)(8 / sizeof(T));
  __shared__ T stage0[STAGE];
  __shared__ T stage1[STAGE];
  const int32_t total = total_tiles(arg);
  const int32_t per = (total + 7) >> 3;
  const int32_t x = (int32_t)(blockIdx.x & 7), step = (int32_t)(gridDim.x >> 3);
  const int32_t tend = min(total, (x + 1) * per);
  for (int32_t tg = x * per + (int32_t)(blockIdx.x >> 3); tg < tend; tg += step) {
    __syncthreads();  // the previous tile's last chunk is read before this tile's first copies
    TileArgs a;
    const int32_t tile = tile_at(arg, tg, a);
    quad_tile<MODE, QT>(a, tile, stage0, stage1);
  }
}

template <int MODE, bool QT, class ARG>
int launch_kind(int kind, const ARG &arg, int32_t n, int32_t wgs, hipStream_t s) {
  // One workgroup per tile; the class-tile kernel also runs persistent grids
  // of wgs workgroups (rounded up to a multiple of 8).
  unsigned gridx = 8u * (unsigned)((n + 7) / 8);
  if ((kind == kKindQuadSmall || kind == kKindQuadLarge) && wgs > 0)
    gridx = std::min(gridx, 8u * (unsigned)((wgs + 7) / 8));
  const dim3 blk(kTileThreads);
  switch (kind) {
    case kKindDense4: hipLaunchKernelGGL((k_tile_dense4<MODE, QT, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindDenseMx: hipLaunchKernelGGL((k_tile_dense_mx<MODE, QT, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindMixedSmall: hipLaunchKernelGGL((k_tile_mixed<MODE, QT, kStgSmall, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindMixedLarge: hipLaunchKernelGGL((k_tile_mixed<MODE, QT, kStgLarge, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindQuadSmall: hipLaunchKernelGGL((k_tile_quad<MODE, QT, kStgSmall, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindQuadLarge: hipLaunchKernelGGL((k_tile_quad<MODE, QT, kStgLarge, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindUniform + 4: hipLaunchKernelGGL((k_tile_products<MODE, QT, 4, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindUniform + 9: hipLaunchKernelGGL((k_tile_products<MODE, QT, 9, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindUniform + 13: hipLaunchKernelGGL((k_tile_products<MODE, QT, 13, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindUniform + 16: hipLaunchKernelGGL((k_tile_products<MODE, QT, 16, ARG>), dim3(gridx), blk, 0, s, arg); break;
    default: return cq::set_errorf(CQMM_ERR_ARG, "tile products: kernel kind %d has no instance", kind);
  }
  return cq::hip_check(hipGetLastError(), "tile-product kernel launch");
}

template <int MODE, bool QT, class ARG>
int launch_uniform(int kind, const ARG &arg, int32_t n, hipStream_t s) {
  const unsigned gridx = 8u * (unsigned)((n + 7) / 8);
  const dim3 blk(kTileThreads);
  switch (kind) {
    case kKindUniform + 9: hipLaunchKernelGGL((k_tile_products<MODE, QT, 9, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindUniform + 13: hipLaunchKernelGGL((k_tile_products<MODE, QT, 13, ARG>), dim3(gridx), blk, 0, s, arg); break;
    case kKindUniform + 16: hipLaunchKernelGGL((k_tile_products<MODE, QT, 16, ARG>), dim3(gridx), blk, 0, s, arg); break;
    default: return cq::set_errorf(CQMM_ERR_ARG, "tile products: kernel kind %d has no instance", kind);
  }
  return cq::hip_check(hipGetLastError(), "tile-product kernel launch");
}

template <int MODE, bool QT, class ARG>
int launch_quad(int kind, const ARG &arg, int32_t n, int32_t wgs, hipStream_t s) {
  unsigned gridx = 8u * (unsigned)((n + 7) / 8);
  if (wgs > 0) gridx = std::min(gridx, 8u * (unsigned)((wgs + 7) / 8));
  if (kind == kKindQuadSmall) hipLaunchKernelGGL((k_tile_quad<MODE, QT, kStgSmall, ARG>), dim3(gridx), dim3(kTileThreads), 0, s, arg);
  else hipLaunchKernelGGL((k_tile_quad<MODE, QT, kStgLarge, ARG>), dim3(gridx), dim3(kTileThreads), 0, s, arg);
  return cq::hip_check(hipGetLastError(), "class-tile kernel launch");
}

template <class ARG>
int launch_any(int mode, bool qt, int kind, const ARG &arg, int32_t n, int32_t wgs, hipStream_t s) {
  switch (mode * 2 + (qt ? 1 : 0)) {
    case 0: return launch_kind<0, false, ARG>(kind, arg, n, wgs, s);
    case 1: return launch_kind<0, true, ARG>(kind, arg, n, wgs, s);
    case 2: return launch_kind<1, false, ARG>(kind, arg, n, wgs, s);
    case 3: return launch_kind<1, true, ARG>(kind, arg, n, wgs, s);
    case 4: return launch_kind<2, false, ARG>(kind, arg, n, wgs, s);
    case 5: return launch_kind<2, true, ARG>(kind, arg, n, wgs, s);
    // Mode 3 has its own instance in the uniform-size kernel only.
    case 6: return kind >= kKindUniform ? launch_kind<3, false, ARG>(kind, arg, n, wgs, s) : launch_kind<2, false, ARG>(kind, arg, n, wgs, s);
    case 7: return kind >= kKindUniform ? launch_kind<3, true, ARG>(kind, arg, n, wgs, s) : launch_kind<2, true, ARG>(kind, arg, n, wgs, s);
    // Split-BF16 x3 (mode 4) runs in the class-tile kernel and the fixed-size
    // kernels of 9x9, 13x13 and 16x16 blocks.
    case 8:
    case 9:
      if (kind == kKindQuadSmall || kind == kKindQuadLarge)
        return qt ? launch_quad<4, true, ARG>(kind, arg, n, wgs, s) : launch_quad<4, false, ARG>(kind, arg, n, wgs, s);
      if (kind == kKindUniform + 9 || kind == kKindUniform + 13 || kind == kKindUniform + 16)
        return qt ? launch_uniform<4, true, ARG>(kind, arg, n, s) : launch_uniform<4, false, ARG>(kind, arg, n, s);
      return cq::set_error(CQMM_ERR_STATE, "split-BF16 products need blocks of 9 to 16 or a class-tile schedule");
    default: return cq::set_errorf(CQMM_ERR_ARG, "tile products: precision mode %d", mode);
  }
}

int tile_kind(const cq::TilePlan &tp) {
  if (tp.maxn) {
    if (tp.dense) return kKindDenseMx;
    if (tp.quad) return stage_elems(tp.maxn) == kStgSmall ? kKindQuadSmall : kKindQuadLarge;
    return stage_elems(tp.maxn) == kStgSmall ? kKindMixedSmall : kKindMixedLarge;
  }
  if (tp.dense) return kKindDense4;
  return kKindUniform + tp.nsf;
}

}  // namespace

namespace cq {

int launch_tile_products_f32(const TilePlan &tp, bool q_trans, const float *P, const float *Q, float *O, int mode,
                             hipStream_t s) {
  if (tp.n_tiles == 0) return CQMM_OK;
  return launch_any(mode, q_trans, tile_kind(tp), tile_launch_args(tp, P, Q, O), tp.n_tiles, tp.wgs, s);
}

TileArgs tile_launch_args(const TilePlan &tp, const void *P, const void *Q, void *O) {
  TileArgs a;
  std::memset(&a, 0, sizeof(a));
  a.tile_chunk = tp.tile_chunk.as<int32_t>();
  a.chunk_lbeg = tp.chunk_lbeg.as<int64_t>();
  a.list = tp.list.as<int32_t>();
  a.list64 = tp.list64.as<uint64_t>();
  a.listls = tp.listls.as<uint32_t>();
  a.pstart = tp.pstart.as<int64_t>();
  a.pcnt = tp.pcnt.as<uint64_t>();
  a.prod = tp.prod.as<uint32_t>();
  a.wave_out = tp.wave_out.as<int32_t>();
  a.wave_out64 = tp.wave_out64.as<int64_t>();
  a.wave_dim = tp.wave_dim.as<uint8_t>();
  a.chunk_entry = tp.chunk_entry.as<int32_t>();
  a.emask = tp.entry_mask.as<uint32_t>();
  a.ecoff = tp.entry_coff.as<int32_t>();
  a.rdim = tp.tile_rdim.as<uint8_t>();
  a.cdim = tp.tile_cdim.as<uint8_t>();
  a.P = P;
  a.Q = Q;
  a.O = O;
  a.n_tiles = tp.n_tiles;
  a.tile_base = 0;
  return a;
}

bool tiles_take_bf16x3(const TilePlan &tp) {
  if (tp.n_tiles == 0) return false;
  const int k = tile_kind(tp);
  return k == kKindQuadSmall || k == kKindQuadLarge || k == kKindUniform + 9 || k == kKindUniform + 13 ||
         k == kKindUniform + 16;
}

int tile_family(const TilePlan &tp) {
  const int k = tile_kind(tp);
  if (k == kKindQuadSmall || k == kKindQuadLarge) return CQMM_TILES_CLASS;
  if (k == kKindMixedSmall || k == kKindMixedLarge) return CQMM_TILES_MIXED;
  return k;  // dense 1, dense mixed 2, uniform 16 + n
}

int tile_group_key(const TilePlan &tp, bool q_trans, int mode) {
  return (tile_kind(tp) * 2 + (q_trans ? 1 : 0)) * 4 + mode;
}

int launch_tile_args(const TilePlan &tp, bool q_trans, int mode, const TileArgs &a, hipStream_t s) {
  if (a.n_tiles <= 0) return CQMM_OK;
  return launch_any(mode, q_trans, tile_kind(tp), a, a.n_tiles, tp.wgs, s);
}

int launch_tile_group(const TilePlan &tp, bool q_trans, int mode, const TileGroup &g, hipStream_t s) {
  if (g.ngrp <= 0 || g.total <= 0) return CQMM_OK;
  if (g.ngrp > kMaxGroup) return set_errorf(CQMM_ERR_ARG, "grouped tile launch: %d products (at most %d)", g.ngrp, kMaxGroup);
  return launch_any(mode, q_trans, tile_kind(tp), g, g.total, tp.wgs, s);
}

// CQMM_TILES=0 builds plans without the tile schedule (triplet kernels only).
bool tiles_enabled() {
  static const bool v = [] {
    const char *e = getenv("CQMM_TILES");
    return !(e && std::string(e) == "0");
  }();
  return v;
}

// Read at every plan build, so that one process can build both schedules.
bool class_tiles_enabled() {
  const char *e = getenv("CQMM_MIXED_TILES");
  return e && std::string(e) == "class";
}

// CQMM_TILE_ROWS=morton: plans built from a system order their tile rows like
// the tile columns (symmetric-result plans always do).
bool morton_rows_enabled() {
  const char *e = getenv("CQMM_TILE_ROWS");
  return e && std::string(e) == "morton";
}

// CQMM_TILE_ROWS=primary: symmetric-result plans keep primary-order tile rows.
bool primary_rows_forced() {
  const char *e = getenv("CQMM_TILE_ROWS");
  return e && std::string(e) == "primary";
}

bool uniform_class_tiles_enabled() {
  const char *e = getenv("CQMM_UNIFORM_TILES");
  return e && std::string(e) == "class";
}

std::vector<int32_t> group_ranks_by_size(const std::vector<int32_t> &rank, const int32_t *size) {
  const int32_t n = (int32_t)rank.size();
  std::vector<int32_t> idx(n), out(n);
  for (int32_t i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
    return size[a] != size[b] ? size[a] < size[b] : rank[a] < rank[b];
  });
  int32_t pos = 0, prev = -1;
  for (int32_t k = 0; k < n; ++k) {
    const int32_t i = idx[k];
    if (size[i] != prev) {
      pos = (pos + 7) & ~7;
      prev = size[i];
    }
    out[i] = pos++;
  }
  return out;
}

std::vector<int32_t> halo_morton_rank(int32_t np_in_halo, const int32_t *nh_part, const int32_t *lab_hcover,
                                      int32_t ncovery, int32_t ncoverz) {
  std::vector<uint64_t> code(np_in_halo);
  const int64_t nyz = (int64_t)ncovery * ncoverz;
  for (int32_t p = 0; p < np_in_halo; ++p) {
    const int64_t l = (int64_t)lab_hcover[p] - 1;
    const uint64_t x = (uint64_t)(l / nyz), y = (uint64_t)((l / ncoverz) % ncovery), z = (uint64_t)(l % ncoverz);
    uint64_t c = 0;
    for (int bit = 0; bit < 20; ++bit)
      c |= (((x >> bit) & 1u) << (3 * bit)) | (((y >> bit) & 1u) << (3 * bit + 1)) | (((z >> bit) & 1u) << (3 * bit + 2));
    code[p] = c;
  }
  std::vector<int32_t> order(np_in_halo);
  for (int32_t p = 0; p < np_in_halo; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
  std::vector<int64_t> first(np_in_halo + 1, 0);
  for (int32_t p = 0; p < np_in_halo; ++p) first[p + 1] = first[p] + nh_part[p];
  std::vector<int32_t> rank(first[np_in_halo]);
  int32_t next = 0;
  for (int32_t p : order)
    for (int64_t k = first[p]; k < first[p + 1]; ++k) rank[k] = next++;
  return rank;
}

// Output slots to wavefronts, per tile: a Latin square over (tile row,
// column) — with k = (row + perm[column]) % 8, wavefront k % kWaves owns the
// block as its slot (8 / kWaves) * row + k / kWaves, so every wavefront owns
// 8 / kWaves blocks of each tile row and column, which spreads every entry's
// rectangle of products evenly — chosen among kLatinTries column permutations
// (the first with the smallest largest wavefront product count).
// owner[tile*64 + row*8 + col] = wavefront * kSlots + slot, or 0xff where the
// tile has no such output block; wave_out[tile*64 + owner] = that block.
constexpr int kLatinTries = 24;
struct LatinPerms {
  int8_t p[kLatinTries][8];
};

__global__ void k_assign_slots(const int32_t *__restrict__ cnt, const int32_t *__restrict__ tile_out,
                               LatinPerms perms, int32_t n_tiles, uint8_t *__restrict__ owner,
                               int32_t *__restrict__ wave_out) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tiles) return;
  const int32_t *c = cnt + (int64_t)t * 64;
  int32_t cs[64];
  for (int sl = 0; sl < 64; ++sl) cs[sl] = c[sl];
  int best = 0;
  int64_t best_max = -1;
  for (int k = 0; k < kLatinTries; ++k) {
    int64_t load[kWaves] = {};
    for (int sl = 0; sl < 64; ++sl) load[(((sl >> 3) + perms.p[k][sl & 7]) & 7) % kWaves] += cs[sl];
    int64_t m = 0;
    for (int w = 0; w < kWaves; ++w) m = load[w] > m ? load[w] : m;
    if (best_max < 0 || m < best_max) {
      best_max = m;
      best = k;
    }
  }
  for (int u = 0; u < 8; ++u)
    for (int col = 0; col < 8; ++col) {
      const int64_t sl = (int64_t)t * 64 + u * 8 + col;
      const int32_t ob = tile_out[sl];
      if (ob < 0) {
        owner[sl] = 0xff;
        continue;
      }
      const int k = (u + perms.p[best][col]) & 7;
      const int o = (k % kWaves) * kSlots + (8 / kWaves) * u + k / kWaves;
      owner[sl] = (uint8_t)o;
      wave_out[(int64_t)t * 64 + o] = ob;
    }
}

// The kLatinTries column permutations (the first is the identity; the others
// a fixed pseudo-random sequence, so plans are reproducible).
LatinPerms latin_perms() {
  LatinPerms lp;
  uint64_t lcg = 0x9e3779b97f4a7c15ull;
  for (int k = 0; k < kLatinTries; ++k) {
    int q[8];
    for (int i = 0; i < 8; ++i) q[i] = i;
    for (int i = 7; k > 0 && i > 0; --i) {
      lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
      const int j = (int)((lcg >> 33) % (uint64_t)(i + 1));
      std::swap(q[i], q[j]);
    }
    for (int i = 0; i < 8; ++i) lp.p[k][i] = (int8_t)q[i];
  }
  return lp;
}

// Device owner table and wave_out (tile_out32: the tiles' output blocks on
// the device) from the products per output slot.
int assign_slots(const TileBuildInput &in, const int32_t *tile_out32, DevBuf &d_owner, DevBuf &wave_out,
                 hipStream_t s) {
  const int64_t n = in.n_trip;
  const size_t ns = (size_t)in.n_tiles * 64;
  int st;
  DevBuf d_cnt;
  if ((st = d_cnt.alloc(ns * 4)) || (st = d_owner.alloc(ns)) || (st = wave_out.alloc(ns * 4))) return st;
  if ((st = hip_check(hipMemsetAsync(d_cnt.p, 0, ns * 4, s), "memset slot counts"))) return st;
  hipLaunchKernelGGL(k_slot_counts, dim3(grid1(n, 256)), dim3(256), 0, s, in.key, in.slot, n, in.nx, d_cnt.as<int32_t>());
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1((int64_t)ns, 256)), dim3(256), 0, s, wave_out.as<int32_t>(), (int64_t)ns, -1);
  hipLaunchKernelGGL(k_assign_slots, dim3(grid1(in.n_tiles, 128)), dim3(128), 0, s, d_cnt.as<int32_t>(), tile_out32,
                     latin_perms(), in.n_tiles, d_owner.as<uint8_t>(), wave_out.as<int32_t>());
  if ((st = hip_check(hipGetLastError(), "slot assignment launch"))) return st;
  // d_cnt is freed on return: wait for the kernels that read it.
  return hip_check(hipStreamSynchronize(s), "slot assignment");
}

int build_tiles_mixed(const TileBuildInput &in, TilePlan &tp, hipStream_t s);

int build_tiles(const TileBuildInput &in, TilePlan &tp, hipStream_t s) {
  if (in.nsf == 0) return build_tiles_mixed(in, tp, s);
  cq::PhaseTimer pt;
  const int64_t n = in.n_trip;
  const int64_t ns2 = (int64_t)in.nsf * in.nsf;
  const int32_t cap = tile_cap(in.nsf);
  int st;
  tp = TilePlan();
  tp.nsf = in.nsf;
  tp.n_tiles = in.n_tiles;
  // Output blocks by block index (element offset / NSF^2).
  std::vector<int32_t> tout32((size_t)in.n_tiles * 64);
  parallel_for((int64_t)tout32.size(), [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k) tout32[k] = in.tile_out[k] < 0 ? -1 : (int32_t)(in.tile_out[k] / ns2);
  });
  if ((st = tp.tile_out.alloc((size_t)in.n_tiles * 64 * 4))) return st;
  if ((st = hip_check(hipMemcpyAsync(tp.tile_out.p, tout32.data(), (size_t)in.n_tiles * 64 * 4, hipMemcpyHostToDevice, s),
                      "H2D tile_out")) ||
      (st = hip_check(hipStreamSynchronize(s), "tile_out")))
    return st;
  DevBuf d_owner;
  tp.dense = in.nsf == kDenseNsf;
  if (tp.dense) {
    // Wavefront w computes quadrant (w / 2, w % 2) of the tile: rows
    // 4*(w/2)..+3, columns 4*(w%2)..+3.
    std::vector<int32_t> wave_out((size_t)in.n_tiles * 64, -1);
    for (int64_t t = 0; t < in.n_tiles; ++t)
      for (int u = 0; u < 8; ++u)
        for (int c = 0; c < 8; ++c)
          wave_out[(t * kWaves + (u >> 2) * 2 + (c >> 2)) * 16 + (u & 3) * 4 + (c & 3)] = tout32[t * 64 + u * 8 + c];
    if ((st = tp.wave_out.alloc(wave_out.size() * 4))) return st;
    if ((st = hip_check(hipMemcpyAsync(tp.wave_out.p, wave_out.data(), wave_out.size() * 4, hipMemcpyHostToDevice, s),
                        "H2D wave_out")) ||
        (st = hip_check(hipStreamSynchronize(s), "wave_out")))
      return st;
  } else {
    pt.mark("  tiles: tile_out");
    if ((st = assign_slots(in, tp.tile_out.as<int32_t>(), d_owner, tp.wave_out, s))) return st;
  }
  const int32_t max_entries = (tp.dense || tp.quad) ? kDenseEntries : kChunkEntries;
  pt.mark("  tiles: owners");
  // Sort the products by (tile, contraction atom).
  unsigned end_bit = 1;
  const uint64_t kmax = (uint64_t)in.n_tiles * (uint64_t)in.nx;
  while (end_bit < 64 && ((uint64_t)1 << end_bit) < kmax) ++end_bit;
  DevBuf d_perm, d_key2, d_perm2, d_tmp;
  if ((st = d_perm.alloc(n * 4)) || (st = d_key2.alloc(n * 8)) || (st = d_perm2.alloc(n * 4))) return st;
  hipLaunchKernelGGL(k_iota32, dim3(grid1(n, 256)), dim3(256), 0, s, d_perm.as<int32_t>(), n);
  size_t bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, in.key, d_key2.as<uint64_t>(), d_perm.as<int32_t>(),
                                           d_perm2.as<int32_t>(), (size_t)n, 0, end_bit, s);
  if ((st = hip_check(e, "radix_sort_pairs (tiles, size)"))) return st;
  if ((st = d_tmp.alloc(bytes))) return st;
  e = rocprim::radix_sort_pairs(d_tmp.p, bytes, in.key, d_key2.as<uint64_t>(), d_perm.as<int32_t>(),
                                d_perm2.as<int32_t>(), (size_t)n, 0, end_bit, s);
  if ((st = hip_check(e, "radix_sort_pairs (tiles)"))) return st;
  if ((st = hip_check(hipStreamSynchronize(s), "tile sort"))) return st;
  d_tmp.release();
  d_perm.release();
  pt.mark("  tiles: sort");
  // Entries = distinct keys.
  DevBuf d_flag, d_eid;
  if ((st = d_flag.alloc(n * 4)) || (st = d_eid.alloc(n * 4))) return st;
  hipLaunchKernelGGL(k_entry_flags, dim3(grid1(n, 256)), dim3(256), 0, s, d_key2.as<uint64_t>(), n, d_flag.as<int32_t>());
  if ((st = inclusive_scan(d_flag.as<int32_t>(), d_eid.as<int32_t>(), (size_t)n, s))) return st;
  d_flag.release();
  int32_t ne32 = 0;
  if ((st = hip_check(hipMemcpyAsync(&ne32, d_eid.as<int32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, s), "D2H entries")) ||
      (st = hip_check(hipStreamSynchronize(s), "entry count")))
    return st;
  const int64_t ne = ne32;
  tp.n_entries = ne;
  DevBuf d_ep, d_eq, d_etile, d_cnt, d_emask, d_elbeg, d_ecoff;
  if ((st = d_ep.alloc(ne * 8 * 4)) || (st = d_eq.alloc(ne * 8 * 4)) || (st = d_etile.alloc(ne * 4)) ||
      (st = d_cnt.alloc(ne * 8)))
    return st;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1(ne * 8, 256)), dim3(256), 0, s, d_ep.as<int32_t>(), ne * 8, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1(ne * 8, 256)), dim3(256), 0, s, d_eq.as<int32_t>(), ne * 8, -1);
  hipLaunchKernelGGL(k_entry_fill, dim3(grid1(n, 256)), dim3(256), 0, s, d_key2.as<uint64_t>(), d_perm2.as<int32_t>(),
                     d_eid.as<int32_t>(), in.slot, in.p, in.q, n, in.nx, ns2, d_ep.as<int32_t>(), d_eq.as<int32_t>(),
                     d_etile.as<int32_t>());
  d_eid.release();
  d_perm2.release();
  d_key2.release();
  if ((st = d_emask.alloc(ne * 4)) || (st = d_elbeg.alloc((ne + 1) * 8))) return st;
  hipLaunchKernelGGL(k_entry_mask, dim3(grid1(ne, 256)), dim3(256), 0, s, d_ep.as<int32_t>(), d_eq.as<int32_t>(), ne,
                     d_emask.as<uint32_t>(), d_cnt.as<int64_t>());
  if ((st = hip_check(hipMemsetAsync(d_elbeg.p, 0, 8, s), "memset lbeg"))) return st;
  if ((st = inclusive_scan(d_cnt.as<int64_t>(), d_elbeg.as<int64_t>() + 1, (size_t)ne, s))) return st;
  int64_t nl = 0;
  if ((st = hip_check(hipMemcpyAsync(&nl, d_elbeg.as<int64_t>() + ne, 8, hipMemcpyDeviceToHost, s), "D2H loads")) ||
      (st = hip_check(hipStreamSynchronize(s), "load count")))
    return st;
  tp.n_loads = nl;
  if ((st = tp.list.alloc(nl * 4))) return st;
  hipLaunchKernelGGL(k_list_fill, dim3(grid1(ne, 256)), dim3(256), 0, s, d_ep.as<int32_t>(), d_eq.as<int32_t>(),
                     d_elbeg.as<int64_t>(), ne, tp.list.as<int32_t>());
  // Chunks.
  DevBuf d_tebeg, d_nch;
  if ((st = d_tebeg.alloc(((size_t)in.n_tiles + 1) * 4)) || (st = d_nch.alloc((size_t)in.n_tiles * 4))) return st;
  hipLaunchKernelGGL(k_tile_ebeg, dim3(grid1((int64_t)in.n_tiles + 1, 256)), dim3(256), 0, s, d_etile.as<int32_t>(), ne,
                     in.n_tiles, d_tebeg.as<int32_t>());
  hipLaunchKernelGGL((k_chunks<false>), dim3(grid1(in.n_tiles, 256)), dim3(256), 0, s, d_tebeg.as<int32_t>(),
                     d_emask.as<uint32_t>(), d_owner.as<uint8_t>(), in.n_tiles, cap, max_entries, nullptr,
                     d_nch.as<int32_t>());
  if ((st = tp.tile_chunk.alloc(((size_t)in.n_tiles + 1) * 4))) return st;
  if ((st = hip_check(hipMemsetAsync(tp.tile_chunk.p, 0, 4, s), "memset tile_chunk"))) return st;
  if ((st = inclusive_scan(d_nch.as<int32_t>(), tp.tile_chunk.as<int32_t>() + 1, (size_t)in.n_tiles, s))) return st;
  int32_t nch = 0;
  if ((st = hip_check(hipMemcpyAsync(&nch, tp.tile_chunk.as<int32_t>() + in.n_tiles, 4, hipMemcpyDeviceToHost, s),
                      "D2H chunks")) ||
      (st = hip_check(hipStreamSynchronize(s), "chunk count")))
    return st;
  tp.n_chunks = nch;
  DevBuf d_centry;
  if ((st = d_centry.alloc(((size_t)nch + 1) * 4))) return st;
  hipLaunchKernelGGL((k_chunks<true>), dim3(grid1(in.n_tiles, 256)), dim3(256), 0, s, d_tebeg.as<int32_t>(),
                     d_emask.as<uint32_t>(), d_owner.as<uint8_t>(), in.n_tiles, cap, max_entries,
                     tp.tile_chunk.as<int32_t>(), d_centry.as<int32_t>());
  if ((st = hip_check(hipMemcpyAsync(d_centry.as<int32_t>() + nch, &ne32, 4, hipMemcpyHostToDevice, s),
                      "H2D chunk end")))
    return st;
  if ((st = tp.chunk_lbeg.alloc(((size_t)nch + 1) * 8))) return st;
  hipLaunchKernelGGL(k_chunk_lbeg, dim3(grid1((int64_t)nch + 1, 256)), dim3(256), 0, s, d_centry.as<int32_t>(),
                     d_elbeg.as<int64_t>(), (int64_t)nch + 1, tp.chunk_lbeg.as<int64_t>());
  if ((st = d_ecoff.alloc((size_t)ne * 4))) return st;
  hipLaunchKernelGGL(k_entry_coff, dim3(grid1((int64_t)nch, 256)), dim3(256), 0, s, d_centry.as<int32_t>(),
                     d_elbeg.as<int64_t>(), (int64_t)nch, d_ecoff.as<int32_t>());
  pt.mark("  tiles: entries + chunks");
  if (tp.dense) {
    // The dense kernel reads the entries themselves; no product lists.
    tp.chunk_entry = std::move(d_centry);
    tp.entry_mask = std::move(d_emask);
    tp.entry_coff = std::move(d_ecoff);
    tp.n_prod = n;
    if ((st = hip_check(hipGetLastError(), "tile builder launch"))) return st;
    return hip_check(hipStreamSynchronize(s), "tile build");
  }
  // Product lists: count per chunk, first product of each chunk, fill.
  DevBuf d_ctile, d_np, d_pbase;
  if ((st = d_ctile.alloc((size_t)std::max(nch, 1) * 4)) || (st = d_np.alloc((size_t)std::max(nch, 1) * 8)) ||
      (st = d_pbase.alloc(((size_t)nch + 1) * 8)))
    return st;
  hipLaunchKernelGGL(k_chunk_tile, dim3(grid1(in.n_tiles, 256)), dim3(256), 0, s, tp.tile_chunk.as<int32_t>(),
                     in.n_tiles, d_ctile.as<int32_t>());
  hipLaunchKernelGGL((k_chunk_prods<false>), dim3(grid1(std::max(nch, 1), 256)), dim3(256), 0, s, nch,
                     d_centry.as<int32_t>(), d_ctile.as<int32_t>(), d_emask.as<uint32_t>(), d_ecoff.as<int32_t>(),
                     d_owner.as<uint8_t>(), (int32_t)ns2, d_np.as<int64_t>(), nullptr, nullptr, nullptr, nullptr);
  if ((st = hip_check(hipMemsetAsync(d_pbase.p, 0, 8, s), "memset pbase"))) return st;
  if ((st = inclusive_scan(d_np.as<int64_t>(), d_pbase.as<int64_t>() + 1, (size_t)nch, s))) return st;
  int64_t nprod = 0;
  if ((st = hip_check(hipMemcpyAsync(&nprod, d_pbase.as<int64_t>() + nch, 8, hipMemcpyDeviceToHost, s), "D2H products")) ||
      (st = hip_check(hipStreamSynchronize(s), "product count")))
    return st;
  tp.n_prod = nprod;
  if (nprod != n)
    return set_errorf(CQMM_ERR_STATE, "tile plan: %lld listed products for %lld triplets", (long long)nprod,
                      (long long)n);
  // pstart/pcnt carry one extra chunk (the clamp target of the kernel's
  // per-chunk registers) and prod kWaveProducts extra entries (each
  // wavefront loads a full row of 64 list words per chunk).
  if ((st = tp.pstart.alloc(((size_t)nch + 1) * kWaves * 8)) || (st = tp.pcnt.alloc(((size_t)nch + 1) * kWaves * 8)) ||
      (st = tp.prod.alloc(((size_t)nprod + kWaveProducts) * 4)))
    return st;
  std::vector<int64_t> tail_start(kWaves, nprod);
  if ((st = hip_check(hipMemsetAsync(tp.pcnt.as<uint64_t>() + (size_t)nch * kWaves, 0, kWaves * 8, s), "memset pcnt")) ||
      (st = hip_check(hipMemcpyAsync(tp.pstart.as<int64_t>() + (size_t)nch * kWaves, tail_start.data(), kWaves * 8,
                                     hipMemcpyHostToDevice, s), "H2D pstart tail")) ||
      (st = hip_check(hipMemsetAsync(tp.prod.as<uint32_t>() + nprod, 0, kWaveProducts * 4, s), "memset prod tail")))
    return st;
  hipLaunchKernelGGL((k_chunk_prods<true>), dim3(grid1(std::max(nch, 1), 256)), dim3(256), 0, s, nch,
                     d_centry.as<int32_t>(), d_ctile.as<int32_t>(), d_emask.as<uint32_t>(), d_ecoff.as<int32_t>(),
                     d_owner.as<uint8_t>(), (int32_t)ns2, nullptr, d_pbase.as<int64_t>(), tp.pstart.as<int64_t>(),
                     tp.pcnt.as<uint64_t>(), tp.prod.as<uint32_t>());
  if ((st = hip_check(hipGetLastError(), "tile builder launch"))) return st;
  st = hip_check(hipStreamSynchronize(s), "tile build");
  pt.mark("  tiles: product lists");
  return st;
}

int launch_tile_products(const TilePlan &tp, bool q_trans, const double *P, const double *Q, double *O,
                         hipStream_t s) {
  if (tp.n_tiles == 0) return CQMM_OK;
  return launch_any(0, q_trans, tile_kind(tp), tile_launch_args(tp, P, Q, O), tp.n_tiles, tp.wgs, s);
}

// Mixed block sizes: the uniform builder's steps with element offsets and
// per-row/column/contraction dimensions (see the kernels above).
int build_tiles_mixed(const TileBuildInput &in, TilePlan &tp, hipStream_t s) {
  const int64_t n = in.n_trip;
  const int32_t cap = stage_elems(in.maxn);
  int st;
  tp = TilePlan();
  tp.nsf = 0;
  tp.maxn = in.maxn;
  tp.n_tiles = in.n_tiles;
  DevBuf d_owner, d_rdim, d_cdim, d_xdim;
  tp.dense = in.maxn <= kDenseNsf;
  tp.quad = !tp.dense && in.class_uniform;
  if (tp.dense || tp.quad) {
    // Blocks of at most 4x4: wavefront w computes quadrant (w / 2, w % 2) of
    // the tile (k_tile_dense_mx); wave_out64 holds its blocks' element offsets.
    const size_t ns = (size_t)in.n_tiles * 64;
    std::vector<int64_t> wave_out(ns, -1);
    parallel_for(in.n_tiles, [&](int64_t b, int64_t e) {
      for (int64_t t = b; t < e; ++t)
        for (int u = 0; u < 8; ++u)
          for (int c = 0; c < 8; ++c)
            wave_out[(t * kWaves + (u >> 2) * 2 + (c >> 2)) * 16 + (u & 3) * 4 + (c & 3)] = in.tile_out[t * 64 + u * 8 + c];
    });
    if ((st = tp.wave_out64.alloc(ns * 8)) || (st = d_rdim.alloc((size_t)in.n_tiles * 8)) ||
        (st = d_cdim.alloc((size_t)in.n_tiles * 8)) || (st = d_xdim.alloc((size_t)std::max<int64_t>(in.nx, 1))))
      return st;
    if ((st = hip_check(hipMemcpyAsync(tp.wave_out64.p, wave_out.data(), ns * 8, hipMemcpyHostToDevice, s),
                        "H2D wave_out")) ||
        (st = hip_check(hipMemcpyAsync(d_rdim.p, in.tile_rdim, (size_t)in.n_tiles * 8, hipMemcpyHostToDevice, s),
                        "H2D rdim")) ||
        (st = hip_check(hipMemcpyAsync(d_cdim.p, in.tile_cdim, (size_t)in.n_tiles * 8, hipMemcpyHostToDevice, s),
                        "H2D cdim")) ||
        (st = hip_check(hipMemcpyAsync(d_xdim.p, in.xdim, (size_t)in.nx, hipMemcpyHostToDevice, s), "H2D xdim")) ||
        (st = hip_check(hipStreamSynchronize(s), "dense tile data")))
      return st;
  } else {
    // Slot owners on the device (from which output blocks exist), then the
    // per-wavefront element offsets and block shapes on the host.
    const size_t ns = (size_t)in.n_tiles * 64;
    std::vector<int32_t> exists(ns);
    parallel_for((int64_t)ns, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; ++k) exists[k] = in.tile_out[k] < 0 ? -1 : 0;
    });
    std::vector<uint8_t> owner(ns);
    {
      DevBuf d_exists, d_wo;
      if ((st = d_exists.alloc(ns * 4))) return st;
      if ((st = hip_check(hipMemcpyAsync(d_exists.p, exists.data(), ns * 4, hipMemcpyHostToDevice, s), "H2D exists")))
        return st;
      if ((st = assign_slots(in, d_exists.as<int32_t>(), d_owner, d_wo, s))) return st;
      if ((st = hip_check(hipMemcpyAsync(owner.data(), d_owner.p, ns, hipMemcpyDeviceToHost, s), "D2H owner")) ||
          (st = hip_check(hipStreamSynchronize(s), "owner")))
        return st;
    }
    std::vector<int64_t> wave_out(ns, -1);
    std::vector<uint8_t> wave_dim(ns, 0);
    for (int64_t t = 0; t < in.n_tiles; ++t)
      for (int sl = 0; sl < 64; ++sl) {
        const uint8_t o = owner[t * 64 + sl];
        if (o == 0xff) continue;
        const int64_t k = t * 64 + o;
        wave_out[k] = in.tile_out[t * 64 + sl];
        const int nr = in.tile_rdim[t * 8 + (sl >> 3)], nc = in.tile_cdim[t * 8 + (sl & 7)];
        if (nr < 1 || nr > 16 || nc < 1 || nc > 16)
          return set_errorf(CQMM_ERR_STATE, "tile plan: output block %d x %d outside 1..16", nr, nc);
        wave_dim[k] = (uint8_t)((nr - 1) | (nc - 1) << 4);
      }
    if ((st = tp.wave_out64.alloc(wave_out.size() * 8)) ||
        (st = tp.wave_dim.alloc(wave_dim.size())) || (st = d_rdim.alloc((size_t)in.n_tiles * 8)) ||
        (st = d_cdim.alloc((size_t)in.n_tiles * 8)) || (st = d_xdim.alloc((size_t)std::max<int64_t>(in.nx, 1))))
      return st;
    if ((st = hip_check(hipMemcpyAsync(tp.wave_out64.p, wave_out.data(), wave_out.size() * 8, hipMemcpyHostToDevice, s),
                        "H2D wave_out")) ||
        (st = hip_check(hipMemcpyAsync(tp.wave_dim.p, wave_dim.data(), wave_dim.size(), hipMemcpyHostToDevice, s),
                        "H2D wave_dim")) ||
        (st = hip_check(hipMemcpyAsync(d_rdim.p, in.tile_rdim, (size_t)in.n_tiles * 8, hipMemcpyHostToDevice, s),
                        "H2D rdim")) ||
        (st = hip_check(hipMemcpyAsync(d_cdim.p, in.tile_cdim, (size_t)in.n_tiles * 8, hipMemcpyHostToDevice, s),
                        "H2D cdim")) ||
        (st = hip_check(hipMemcpyAsync(d_xdim.p, in.xdim, (size_t)in.nx, hipMemcpyHostToDevice, s), "H2D xdim")) ||
        (st = hip_check(hipStreamSynchronize(s), "owner")))
      return st;
  }
  // Sort the products by (tile, contraction atom).
  unsigned end_bit = 1;
  const uint64_t kmax = (uint64_t)in.n_tiles * (uint64_t)in.nx;
  while (end_bit < 64 && ((uint64_t)1 << end_bit) < kmax) ++end_bit;
  DevBuf d_perm, d_key2, d_perm2, d_tmp;
  if ((st = d_perm.alloc(n * 4)) || (st = d_key2.alloc(n * 8)) || (st = d_perm2.alloc(n * 4))) return st;
  hipLaunchKernelGGL(k_iota32, dim3(grid1(n, 256)), dim3(256), 0, s, d_perm.as<int32_t>(), n);
  size_t bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, in.key, d_key2.as<uint64_t>(), d_perm.as<int32_t>(),
                                           d_perm2.as<int32_t>(), (size_t)n, 0, end_bit, s);
  if ((st = hip_check(e, "radix_sort_pairs (tiles, size)"))) return st;
  if ((st = d_tmp.alloc(bytes))) return st;
  e = rocprim::radix_sort_pairs(d_tmp.p, bytes, in.key, d_key2.as<uint64_t>(), d_perm.as<int32_t>(),
                                d_perm2.as<int32_t>(), (size_t)n, 0, end_bit, s);
  if ((st = hip_check(e, "radix_sort_pairs (tiles)"))) return st;
  if ((st = hip_check(hipStreamSynchronize(s), "tile sort"))) return st;
  d_tmp.release();
  d_perm.release();
  DevBuf d_flag, d_eid;
  if ((st = d_flag.alloc(n * 4)) || (st = d_eid.alloc(n * 4))) return st;
  hipLaunchKernelGGL(k_entry_flags, dim3(grid1(n, 256)), dim3(256), 0, s, d_key2.as<uint64_t>(), n, d_flag.as<int32_t>());
  if ((st = inclusive_scan(d_flag.as<int32_t>(), d_eid.as<int32_t>(), (size_t)n, s))) return st;
  d_flag.release();
  int32_t ne32 = 0;
  if ((st = hip_check(hipMemcpyAsync(&ne32, d_eid.as<int32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, s), "D2H entries")) ||
      (st = hip_check(hipStreamSynchronize(s), "entry count")))
    return st;
  const int64_t ne = ne32;
  tp.n_entries = ne;
  DevBuf d_ep, d_eq, d_etile, d_ex, d_cnt, d_emask, d_esz, d_elbeg, d_ecoff;
  if ((st = d_ep.alloc(ne * 8 * 8)) || (st = d_eq.alloc(ne * 8 * 8)) || (st = d_etile.alloc(ne * 4)) ||
      (st = d_ex.alloc(ne * 4)) || (st = d_cnt.alloc(ne * 8)))
    return st;
  if ((st = hip_check(hipMemsetAsync(d_ep.p, 0xff, ne * 8 * 8, s), "memset ep")) ||
      (st = hip_check(hipMemsetAsync(d_eq.p, 0xff, ne * 8 * 8, s), "memset eq")))
    return st;
  hipLaunchKernelGGL(k_entry_fill_mx, dim3(grid1(n, 256)), dim3(256), 0, s, d_key2.as<uint64_t>(), d_perm2.as<int32_t>(),
                     d_eid.as<int32_t>(), in.slot, in.p, in.q, n, in.nx, d_ep.as<int64_t>(), d_eq.as<int64_t>(),
                     d_etile.as<int32_t>(), d_ex.as<int32_t>());
  d_eid.release();
  d_perm2.release();
  d_key2.release();
  if ((st = d_emask.alloc(ne * 4)) || (st = d_esz.alloc(ne * 4)) || (st = d_elbeg.alloc((ne + 1) * 8))) return st;
  hipLaunchKernelGGL(k_entry_mask_mx, dim3(grid1(ne, 256)), dim3(256), 0, s, d_ep.as<int64_t>(), d_eq.as<int64_t>(),
                     d_etile.as<int32_t>(), d_ex.as<int32_t>(), d_rdim.as<uint8_t>(), d_cdim.as<uint8_t>(),
                     d_xdim.as<uint8_t>(), ne, d_emask.as<uint32_t>(), d_esz.as<int32_t>(), d_cnt.as<int64_t>());
  if ((st = hip_check(hipMemsetAsync(d_elbeg.p, 0, 8, s), "memset lbeg"))) return st;
  if ((st = inclusive_scan(d_cnt.as<int64_t>(), d_elbeg.as<int64_t>() + 1, (size_t)ne, s))) return st;
  int64_t nl = 0;
  if ((st = hip_check(hipMemcpyAsync(&nl, d_elbeg.as<int64_t>() + ne, 8, hipMemcpyDeviceToHost, s), "D2H loads")) ||
      (st = hip_check(hipStreamSynchronize(s), "load count")))
    return st;
  tp.n_loads = nl;
  DevBuf d_lsz;
  if ((st = tp.list64.alloc(nl * 8)) || (st = d_lsz.alloc(nl * 2)) || (st = tp.listls.alloc(nl * 4))) return st;
  hipLaunchKernelGGL(k_list_fill_mx, dim3(grid1(ne, 256)), dim3(256), 0, s, d_ep.as<int64_t>(), d_eq.as<int64_t>(),
                     d_etile.as<int32_t>(), d_ex.as<int32_t>(), d_rdim.as<uint8_t>(), d_cdim.as<uint8_t>(),
                     d_xdim.as<uint8_t>(), d_elbeg.as<int64_t>(), ne, tp.list64.as<uint64_t>(), d_lsz.as<uint16_t>());
  d_ep.release();
  d_eq.release();
  // Chunks.
  DevBuf d_tebeg, d_nch;
  if ((st = d_tebeg.alloc(((size_t)in.n_tiles + 1) * 4)) || (st = d_nch.alloc((size_t)in.n_tiles * 4))) return st;
  hipLaunchKernelGGL(k_tile_ebeg, dim3(grid1((int64_t)in.n_tiles + 1, 256)), dim3(256), 0, s, d_etile.as<int32_t>(), ne,
                     in.n_tiles, d_tebeg.as<int32_t>());
  const int32_t max_entries = tp.dense ? kDenseEntries : kChunkEntries;
  hipLaunchKernelGGL((k_chunks_mx<false>), dim3(grid1(in.n_tiles, 256)), dim3(256), 0, s, d_tebeg.as<int32_t>(),
                     d_emask.as<uint32_t>(), d_esz.as<int32_t>(), d_owner.as<uint8_t>(), in.n_tiles, cap, max_entries,
                     nullptr, d_nch.as<int32_t>());
  if ((st = tp.tile_chunk.alloc(((size_t)in.n_tiles + 1) * 4))) return st;
  if ((st = hip_check(hipMemsetAsync(tp.tile_chunk.p, 0, 4, s), "memset tile_chunk"))) return st;
  if ((st = inclusive_scan(d_nch.as<int32_t>(), tp.tile_chunk.as<int32_t>() + 1, (size_t)in.n_tiles, s))) return st;
  int32_t nch = 0;
  if ((st = hip_check(hipMemcpyAsync(&nch, tp.tile_chunk.as<int32_t>() + in.n_tiles, 4, hipMemcpyDeviceToHost, s),
                      "D2H chunks")) ||
      (st = hip_check(hipStreamSynchronize(s), "chunk count")))
    return st;
  tp.n_chunks = nch;
  DevBuf d_centry;
  if ((st = d_centry.alloc(((size_t)nch + 1) * 4))) return st;
  hipLaunchKernelGGL((k_chunks_mx<true>), dim3(grid1(in.n_tiles, 256)), dim3(256), 0, s, d_tebeg.as<int32_t>(),
                     d_emask.as<uint32_t>(), d_esz.as<int32_t>(), d_owner.as<uint8_t>(), in.n_tiles, cap, max_entries,
                     tp.tile_chunk.as<int32_t>(), d_centry.as<int32_t>());
  if ((st = hip_check(hipMemcpyAsync(d_centry.as<int32_t>() + nch, &ne32, 4, hipMemcpyHostToDevice, s),
                      "H2D chunk end")))
    return st;
  if ((st = tp.chunk_lbeg.alloc(((size_t)nch + 1) * 8))) return st;
  hipLaunchKernelGGL(k_chunk_lbeg, dim3(grid1((int64_t)nch + 1, 256)), dim3(256), 0, s, d_centry.as<int32_t>(),
                     d_elbeg.as<int64_t>(), (int64_t)nch + 1, tp.chunk_lbeg.as<int64_t>());
  if ((st = d_ecoff.alloc((size_t)ne * 4))) return st;
  hipLaunchKernelGGL(k_entry_coff_mx, dim3(grid1((int64_t)nch, 256)), dim3(256), 0, s, d_centry.as<int32_t>(),
                     d_esz.as<int32_t>(), (int64_t)nch, d_ecoff.as<int32_t>());
  hipLaunchKernelGGL(k_list_lds_mx, dim3(grid1((int64_t)nch, 256)), dim3(256), 0, s, tp.chunk_lbeg.as<int64_t>(),
                     d_lsz.as<uint16_t>(), (int64_t)nch, tp.listls.as<uint32_t>());
  if (tp.dense || tp.quad) {
    // The dense and class-tile kernels read the entries and the tile block
    // shapes; no product lists.
    tp.chunk_entry = std::move(d_centry);
    tp.entry_mask = std::move(d_emask);
    tp.entry_coff = std::move(d_ecoff);
    tp.tile_rdim = std::move(d_rdim);
    tp.tile_cdim = std::move(d_cdim);
    tp.n_prod = n;
    if ((st = hip_check(hipGetLastError(), "tile builder launch"))) return st;
    return hip_check(hipStreamSynchronize(s), "tile build");
  }
  // Product lists.
  DevBuf d_ctile, d_np, d_pbase;
  if ((st = d_ctile.alloc((size_t)std::max(nch, 1) * 4)) || (st = d_np.alloc((size_t)std::max(nch, 1) * 8)) ||
      (st = d_pbase.alloc(((size_t)nch + 1) * 8)))
    return st;
  hipLaunchKernelGGL(k_chunk_tile, dim3(grid1(in.n_tiles, 256)), dim3(256), 0, s, tp.tile_chunk.as<int32_t>(),
                     in.n_tiles, d_ctile.as<int32_t>());
  hipLaunchKernelGGL((k_chunk_prods_mx<false>), dim3(grid1(std::max(nch, 1), 256)), dim3(256), 0, s, nch,
                     d_centry.as<int32_t>(), d_ctile.as<int32_t>(), d_emask.as<uint32_t>(), d_ecoff.as<int32_t>(),
                     d_ex.as<int32_t>(), d_owner.as<uint8_t>(), d_rdim.as<uint8_t>(), d_cdim.as<uint8_t>(),
                     d_xdim.as<uint8_t>(), d_np.as<int64_t>(), nullptr, nullptr, nullptr, nullptr);
  if ((st = hip_check(hipMemsetAsync(d_pbase.p, 0, 8, s), "memset pbase"))) return st;
  if ((st = inclusive_scan(d_np.as<int64_t>(), d_pbase.as<int64_t>() + 1, (size_t)nch, s))) return st;
  int64_t nprod = 0;
  if ((st = hip_check(hipMemcpyAsync(&nprod, d_pbase.as<int64_t>() + nch, 8, hipMemcpyDeviceToHost, s), "D2H products")) ||
      (st = hip_check(hipStreamSynchronize(s), "product count")))
    return st;
  tp.n_prod = nprod;
  if (nprod != n)
    return set_errorf(CQMM_ERR_STATE, "tile plan: %lld listed products for %lld triplets", (long long)nprod,
                      (long long)n);
  if ((st = tp.pstart.alloc(((size_t)nch + 1) * kWaves * 8)) || (st = tp.pcnt.alloc(((size_t)nch + 1) * kWaves * 8)) ||
      (st = tp.prod.alloc(((size_t)nprod + kWaveProducts) * 4)))
    return st;
  std::vector<int64_t> tail_start(kWaves, nprod);
  if ((st = hip_check(hipMemsetAsync(tp.pcnt.as<uint64_t>() + (size_t)nch * kWaves, 0, kWaves * 8, s), "memset pcnt")) ||
      (st = hip_check(hipMemcpyAsync(tp.pstart.as<int64_t>() + (size_t)nch * kWaves, tail_start.data(), kWaves * 8,
                                     hipMemcpyHostToDevice, s), "H2D pstart tail")) ||
      (st = hip_check(hipMemsetAsync(tp.prod.as<uint32_t>() + nprod, 0, kWaveProducts * 4, s), "memset prod tail")))
    return st;
  hipLaunchKernelGGL((k_chunk_prods_mx<true>), dim3(grid1(std::max(nch, 1), 256)), dim3(256), 0, s, nch,
                     d_centry.as<int32_t>(), d_ctile.as<int32_t>(), d_emask.as<uint32_t>(), d_ecoff.as<int32_t>(),
                     d_ex.as<int32_t>(), d_owner.as<uint8_t>(), d_rdim.as<uint8_t>(), d_cdim.as<uint8_t>(),
                     d_xdim.as<uint8_t>(), nullptr, d_pbase.as<int64_t>(), tp.pstart.as<int64_t>(),
                     tp.pcnt.as<uint64_t>(), tp.prod.as<uint32_t>());
  if ((st = hip_check(hipGetLastError(), "mixed tile builder launch"))) return st;
  return hip_check(hipStreamSynchronize(s), "mixed tile build");
}

}  // namespace cq
