// Device-side declarations shared by the HIP sources of libcqmm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cq {

struct Runtime {
  bool initialised = false;
  int device = -1;
  int n_cu = 0;
  hipStream_t stream = nullptr;
};

Runtime &runtime();
int ensure_init();
// Device address of the no-allocation scratch array on `device` (the device
// current), looked up once at cqmm_init (cq_noalloc.hip).
int noalloc_setup(int device);
// Stream for the calling thread's library work: the thread's own stream when
// one was set with set_thread_stream, else the library stream.
hipStream_t stream();
void set_thread_stream(hipStream_t s);
int hip_check(hipError_t e, const char *what);

// Owning device buffer (freed with the plan that holds it).
// Device bytes held through DevBuf by the calling thread, and their high-water
// mark since the last hbm_peak_reset() (per-request HBM accounting in serving).
void hbm_track(int64_t delta);
int64_t hbm_current();
int64_t hbm_peak();
void hbm_peak_reset();

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      hbm_track(-(int64_t)bytes);
    }
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n);
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Output-tile schedule of one product with a uniform block size (NSF x NSF
// blocks stored contiguously, so every offset is a multiple of NSF^2 and is kept
// as a block index). A tile is kTileRows primary atoms x kTileCols column atoms
// of the output; its entries are the contraction atoms x with at least one
// product in the tile, each carrying which of the tile's left-operand blocks
// (row, x) and right-operand blocks (column, x) exist. Entries are grouped into
// chunks whose operand blocks fit the kernel's LDS stage.
constexpr int kTileRows = 8, kTileCols = 8;
struct TilePlan {
  int nsf = 0;
  int32_t n_tiles = 0, n_chunks = 0;
  int64_t n_entries = 0, n_loads = 0, n_prod = 0;
  DevBuf tile_chunk;   // int32 [n_tiles+1]  first chunk of each tile
  DevBuf chunk_lbeg;   // int64 [n_chunks+1] first operand of each chunk in `list`
  DevBuf list;         // int32 [n_loads] operand block index; bit 31 set: right operand
  DevBuf pstart;       // int64 [(n_chunks+1)*4] first product of (chunk, wavefront) in `prod`
  DevBuf pcnt;         // uint64 [(n_chunks+1)*4] products of (chunk, wavefront) per slot, 4 bits each
  DevBuf prod;         // uint32 [n_prod+64] LDS element offsets of the two operands (left | right << 16)
  DevBuf tile_out;     // int32 [n_tiles*64] output block of slot row*8+col, or -1
  DevBuf wave_out;     // int32 [n_tiles*4*16] output block of (tile, wavefront, slot), or -1
  // 4x4 blocks (dense): each wavefront computes a 4x4-block quadrant of the
  // tile per entry with one MFMA; no product lists. wave_out then holds the
  // quadrant's blocks (row * 4 + column within the quadrant).
  bool dense = false;
  bool quad = false;   // class tiles (k_tile_quad): one row and one column block size per tile
  int32_t wgs = 0;     // class tiles: workgroups of a persistent launch (0: one per tile)
  DevBuf chunk_entry;  // int32 [n_chunks+1] first entry of each chunk
  DevBuf entry_mask;   // uint32 [n_entries] left-operand rows | right-operand columns << 8 present
  DevBuf entry_coff;   // int32 [n_entries] first operand block (mixed: element) of the entry in its chunk's stage
  DevBuf tile_rdim;    // uint8 [n_tiles*8] rows of each tile row's blocks (mixed dense: blocks up to 4x4)
  DevBuf tile_cdim;    // uint8 [n_tiles*8] columns of each tile column's blocks
  // Mixed block sizes (maxn > 0): operands and outputs by element offset.
  int maxn = 0;        // largest block dimension; 0: uniform schedule above
  DevBuf list64;       // uint64 [n_loads] operand element offset; bit 63 set: right operand
  DevBuf listls;       // uint32 [n_loads] LDS element offset in the chunk | elements << 16
  DevBuf wave_out64;   // int64 [n_tiles*4*16] output element offset of (tile, wavefront, slot), or -1
  DevBuf wave_dim;     // uint8 [n_tiles*4*16] (rows - 1) | (columns - 1) << 4 of that output block
};

// Arguments of one tile-kernel launch over one product's schedule (the fields
// a kernel family does not use stay null). tile_base: first tile of this
// product in a grouped launch.
struct TileArgs {
  const int32_t *tile_chunk;
  const int64_t *chunk_lbeg;
  const int32_t *list;          // uniform sizes: operand block indices
  const uint64_t *list64;       // mixed sizes: operand element offsets
  const uint32_t *listls;       // mixed sizes: LDS offsets and sizes
  const int64_t *pstart;
  const uint64_t *pcnt;
  const uint32_t *prod;
  const int32_t *wave_out;      // uniform sizes
  const int64_t *wave_out64;    // mixed sizes
  const uint8_t *wave_dim;
  const int32_t *chunk_entry;   // dense-quadrant kernels
  const uint32_t *emask;
  const int32_t *ecoff;
  const uint8_t *rdim, *cdim;
  const void *P, *Q;
  void *O;
  int32_t n_tiles, tile_base;
};
// Up to kMaxGroup products in one grouped launch, passed by value (kernel
// argument memory, 2.7 KB).
constexpr int kMaxGroup = 16;
struct TileGroup {
  TileArgs a[kMaxGroup];
  int32_t ngrp, total;
};

// Triplet list of one product, grouped by output block (segment), each group in
// the reference's summation order. Offsets are 0-based element offsets.
struct TripletPlan {
  int mult_type = 1;
  bool q_trans = false;     // right operand used transposed (type 2)
  bool off64 = false;       // operand offsets stored as int64
  int uniform = 0;          // support functions per atom when uniform, else 0
  int maxn = 0;             // largest block dimension (up to 64; above 16: sub-block kernels)
  int64_t n_seg = 0;        // output blocks (every block of the output matrix)
  int64_t n_trip = 0;
  int64_t len_out = 0, len_p = 0, len_q = 0;
  double flops = 0.0;
  DevBuf seg_beg;           // int64 [n_seg+1]
  DevBuf seg_out;           // int64 [n_seg]
  DevBuf seg_mn;            // uint8 [2*n_seg] (non-uniform only)
  DevBuf tp, tq;            // int32 or int64 [n_trip]
  DevBuf tk;                // uint8 [n_trip] inner dimension (non-uniform only)
  TilePlan tiles;           // FP64 tile schedule (uniform block sizes), n_tiles == 0 if absent
  int kernel = 0;           // CQMM_KERNEL_* selection for the FP64 path
  // Symmetric results (cqmm_plan_create_symmetric): of each transpose pair of
  // result blocks (p, q), p < q, only p is computed; the launchers then write
  // block q as the transpose of block p (n_mirror pairs, 0 for other plans).
  int64_t n_mirror = 0;
  DevBuf mir_p, mir_q, mir_n1, mir_n2;  // int64 element offsets, int64, uint8 rows, uint8 columns of p
};

// Tile schedule construction from the builder's per-triplet tile keys.
struct TileBuildInput {
  int nsf = 0;
  int64_t n_trip = 0;
  const uint64_t *key = nullptr;   // device [n_trip]: tile * nx + contraction index
  const uint8_t *slot = nullptr;   // device [n_trip]: row%8 * 8 + column rank%8
  const int64_t *p = nullptr, *q = nullptr;  // device [n_trip]: operand element offsets
  int64_t nx = 0;
  int32_t n_tiles = 0;
  const int64_t *tile_out = nullptr;  // host [n_tiles*64]: output element offset, or -1
  int maxn = 0;                        // largest block dimension
  const uint8_t *tile_rdim = nullptr;  // host [n_tiles*8]: block rows of each tile row (0: none)
  const uint8_t *tile_cdim = nullptr;  // host [n_tiles*8]: block columns of each tile column
  const uint8_t *xdim = nullptr;       // host [nx]: contraction dimension of each contraction atom
  bool class_uniform = false;          // every tile has one row and one column block size
};
bool tiles_enabled();
// Species mixes: tiles grouped by block size and run by the class-tile kernel
// (CQMM_MIXED_TILES=class; default: primary/Morton order and the per-product
// mixed-size kernel, which keeps tiles spatially compact).
bool class_tiles_enabled();
// Tile rows of plans built from a system in the column (Morton) order of the
// atoms' own copies (CQMM_TILE_ROWS=morton; symmetric-result plans always).
bool morton_rows_enabled();
bool primary_rows_forced();
// Uniform 9, 13, 16 blocks as one class of the class-tile schedule
// (CQMM_UNIFORM_TILES=class; default: the fixed-size tile kernel).
bool uniform_class_tiles_enabled();
// Ranks of n items in the order (size[i], rank[i]), each size group starting
// at a multiple of 8.
std::vector<int32_t> group_ranks_by_size(const std::vector<int32_t> &rank, const int32_t *size);
// Column order of a halo for tiling: Morton order of the halo partitions'
// cover-set coordinates (label = (x*ncovery + y)*ncoverz + z + 1), atoms of a
// partition in halo order. Returns rank[halo atom].
std::vector<int32_t> halo_morton_rank(int32_t np_in_halo, const int32_t *nh_part, const int32_t *lab_hcover,
                                      int32_t ncovery, int32_t ncoverz);
int build_tiles(const TileBuildInput &in, TilePlan &tp, hipStream_t s);
int launch_tile_products(const TilePlan &tp, bool q_trans, const double *P, const double *Q, double *O,
                         hipStream_t s);
// FP32 precision modes of the product kernels (operands FP32, FP32 MFMA):
// 1 accumulates each output block in FP32; 2 adds each product's FP32 result
// into FP64 registers; 3 sums a block's products within one chunk of the tile
// schedule in FP32 (at most 15 products, consecutive in the block's list) and
// adds that partial sum into FP64 (the per-output-block kernels treat 3 as 2).
int launch_tile_products_f32(const TilePlan &tp, bool q_trans, const float *P, const float *Q, float *O, int mode,
                             hipStream_t s);

// Grouped launches: the tile schedules of up to kMaxGroup products whose
// tile_group_key agrees run in one launch (a[j].tile_base = product j's first
// tile in the group's numbering, total = all tiles). Each product's tiles are computed exactly as in its own
// launch (same summation order). mode: 0 FP64, else an FP32 mode (1, 2, 3).
TileArgs tile_launch_args(const TilePlan &tp, const void *P, const void *Q, void *O);
int tile_group_key(const TilePlan &tp, bool q_trans, int mode);
int launch_tile_group(const TilePlan &tp, bool q_trans, int mode, const TileGroup &g, hipStream_t s);
// One product's tile launch from its TileArgs (the launch cqmm_mat_mult makes).
int launch_tile_args(const TilePlan &tp, bool q_trans, int mode, const TileArgs &a, hipStream_t s);

// True when a plan's products run through its tile schedule (else the
// per-output-block kernels).
bool uses_tiles(const TripletPlan &tp);
// Tile kernel family of a schedule (CQMM_TILES_* of cqmm_plan_info).
int tile_family(const TilePlan &tp);
// True when the schedule's kernel has the split-BF16 x3 mode (class tiles,
// uniform 9, 13, 16).
bool tiles_take_bf16x3(const TilePlan &tp);
// Launch the FP64 (or FP32) block-product kernel over a plan.
int launch_block_products(const TripletPlan &tp, const double *P, const double *Q, double *O,
                          bool fp32, hipStream_t stream);
// Same products on FP32 copies of the operands with FP32 MFMA in precision mode
// `mode` (1, 2 or 3, as for launch_tile_products_f32).
int launch_block_products_f32(const TripletPlan &tp, const float *P, const float *Q, float *O, int mode,
                              hipStream_t stream);
// Symmetric-result plans: write each mirrored block as the transpose of its
// computed partner (FP64 or FP32 result); nothing for other plans.
// launch_block_products(_f32) call it; callers that launch a plan's tile
// schedule themselves call it after the products.
int mirror_result(const TripletPlan &tp, double *O, hipStream_t stream);
int mirror_result_f32(const TripletPlan &tp, float *O, hipStream_t stream);

}  // namespace cq

struct cqmm_plan {
  cq::TripletPlan tp;
  int64_t lena = 0, lenb = 0, lenc = 0;
  // Staging buffers for host-pointer calls (allocated on first use).
  cq::DevBuf sa, sb, sc;
  // FP32 copies of the two operands and the result for CQMM_FP32 calls.
  cq::DevBuf fp, fq, fo;
};
