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
// Stream for the calling thread's library work: the thread's own stream when
// one was set with set_thread_stream, else the library stream.
hipStream_t stream();
void set_thread_stream(hipStream_t s);
int hip_check(hipError_t e, const char *what);

// Owning device buffer (freed with the plan that holds it).
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
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n);
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Triplet list of one product, grouped by output block (segment), each group in
// the reference's summation order. Offsets are 0-based element offsets.
struct TripletPlan {
  int mult_type = 1;
  bool q_trans = false;     // right operand used transposed (type 2)
  bool off64 = false;       // operand offsets stored as int64
  int uniform = 0;          // support functions per atom when uniform, else 0
  int64_t n_seg = 0;        // output blocks (every block of the output matrix)
  int64_t n_trip = 0;
  int64_t len_out = 0, len_p = 0, len_q = 0;
  double flops = 0.0;
  DevBuf seg_beg;           // int64 [n_seg+1]
  DevBuf seg_out;           // int64 [n_seg]
  DevBuf seg_mn;            // uint8 [2*n_seg] (non-uniform only)
  DevBuf tp, tq;            // int32 or int64 [n_trip]
  DevBuf tk;                // uint8 [n_trip] inner dimension (non-uniform only)
  // Host copies of the segment starts for reports (optional).
};

// Launch the FP64 (or FP32) block-product kernel over a plan.
int launch_block_products(const TripletPlan &tp, const double *P, const double *Q, double *O,
                          bool fp32, hipStream_t stream);
// Same products on FP32 copies of the operands with FP32 MFMA; acc64 sums the
// per-product results of each block in FP64 registers, else in the FP32 accumulator.
int launch_block_products_f32(const TripletPlan &tp, const float *P, const float *Q, float *O, bool acc64,
                              hipStream_t stream);

}  // namespace cq

struct cqmm_plan {
  cq::TripletPlan tp;
  int64_t lena = 0, lenb = 0, lenc = 0;
  // Staging buffers for host-pointer calls (allocated on first use).
  cq::DevBuf sa, sb, sc;
};
