// Maps and launchers for the streaming matrix kernels (cq_ops.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cq_device.h"

namespace cq {

// Matched element runs between two patterns: y[dst+e] op= x[src+e], e < len.
struct RunMap {
  int64_t nruns = 0;
  int64_t nelem = 0;
  DevBuf dst, src, len;  // int64, int64, int32
};

// Onsite blocks of a pattern (for matrix_trace).
struct DiagMap {
  int64_t nblk = 0;
  DevBuf off, nsf;  // int64, int32
};

// Transpose pairs (p <= q) of a pattern (for symmetrise).
struct PairMap {
  int64_t npairs = 0;
  DevBuf poff, qoff, n1, n2;  // int64, int64, uint8, uint8
};

int op_axpby(int64_t n, double alpha, double *y, double beta, const double *x, hipStream_t s);
int op_scale(int64_t n, double alpha, double *y, hipStream_t s);
int op_copy(int64_t n, double *y, const double *x, hipStream_t s);
int op_runs_axpby(const RunMap &m, double alpha, double *y, double beta, const double *x, hipStream_t s);
// The same with coefficients read on the device where the pointer is non-null
// (the value argument is used otherwise).
int op_axpby_dev(int64_t n, double alpha, const double *alpha_p, double *y, double beta, const double *beta_p,
                 const double *x, hipStream_t s);
int op_scale_dev(int64_t n, double alpha, const double *alpha_p, double *y, hipStream_t s);
int op_runs_axpby_dev(const RunMap &m, double alpha, const double *alpha_p, double *y, double beta,
                      const double *beta_p, const double *x, hipStream_t s);
int op_dot(int64_t n, const double *x, const double *y, double *d_partials, double *d_out, hipStream_t s);
int op_trace(const DiagMap &m, const double *a, double *d_partials, double *d_out, hipStream_t s);
int op_symmetrise(const PairMap &m, double *a, hipStream_t s);
int op_transpose(const PairMap &m, const double *a, double *b, hipStream_t s);
int op_to_f32(int64_t n, const double *x, float *y, hipStream_t s);
int op_to_f64(int64_t n, const float *x, double *y, hipStream_t s);
// Split-BF16 operand words for the BF16x3 products (bits 0-15: x rounded to
// BF16, bits 16-31: the remainder rounded to BF16); the second form converts
// FP32 values in place.
int op_to_bf16x2(int64_t n, const double *x, uint32_t *y, hipStream_t s);
int op_f32_to_bf16x2(int64_t n, uint32_t *y, hipStream_t s);
int op_set_diag(const DiagMap &m, double v, double *a, hipStream_t s);
int reduction_partials();

}  // namespace cq
