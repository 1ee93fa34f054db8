// Streaming kernels of the matrix API that surrounds the product (mult_module):
// same-range sums and scaling, range-mapped sums, deterministic FP64
// reductions, onsite traces, the in-place symmetrisation L <- (L + L^T)/2, and
// FP64 <-> FP32 conversion. All are HBM-bound; each thread handles consecutive
// elements so loads and stores coalesce.
#include <hip/hip_runtime.h>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"
#include "cq_ops.h"

namespace {

constexpr int kBS = 256;
constexpr int kRedBlocks = 1024;  // fixed partial count: the reduction order never depends on the launch

inline unsigned grid_for(int64_t n, int per_thread = 1) {
  int64_t g = (n + (int64_t)kBS * per_thread - 1) / ((int64_t)kBS * per_thread);
  if (g < 1) g = 1;
  if (g > 65535LL * 16) g = 65535LL * 16;
  return (unsigned)g;
}

// Symmetric results: block q (n x m) = transpose of block p (m x n), one
// wavefront per pair; blocks are column-major at their element offsets.
template <class T>
__global__ void k_mirror(int64_t npairs, const int64_t *__restrict__ poff, const int64_t *__restrict__ qoff,
                         const uint8_t *__restrict__ n1, const uint8_t *__restrict__ n2, T *__restrict__ A) {
  const int lane = threadIdx.x & 63;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < npairs;
       r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t p = poff[r], q = qoff[r];
    const int m = n1[r], n = n2[r];
    for (int e = lane; e < m * n; e += 64) {
      const int a = e % m, b = e / m;
      A[q + b + (int64_t)n * a] = A[p + a + (int64_t)m * b];
    }
  }
}

// A coefficient given by value, or (pointer non-null) read from device memory
// where an earlier kernel on the stream computed it (no host round trip).
__device__ inline double coef(double v, const double *p) { return p ? *p : v; }

// y = alpha*y + beta*x (alpha == 0 stores beta*x without reading y).
__global__ void k_axpby(int64_t n, double av, const double *__restrict__ ap, double *__restrict__ y, double bv,
                        const double *__restrict__ bp, const double *__restrict__ x) {
  const double alpha = coef(av, ap), beta = coef(bv, bp);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = alpha == 0.0 ? beta * x[t] : alpha * y[t] + beta * x[t];
}

__global__ void k_scale(int64_t n, double av, const double *__restrict__ ap, double *__restrict__ y) {
  const double alpha = coef(av, ap);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = alpha == 0.0 ? 0.0 : alpha * y[t];
}

__global__ void k_copy(int64_t n, double *__restrict__ y, const double *__restrict__ x) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = x[t];
}

// Range-mapped sum: runs of matched elements (dst_off, src_off, len).
__global__ void k_runs_axpby(int64_t nruns, const int64_t *__restrict__ dofs, const int64_t *__restrict__ sofs,
                             const int32_t *__restrict__ lens, double av, const double *__restrict__ ap,
                             double *__restrict__ y, double bv, const double *__restrict__ bp,
                             const double *__restrict__ x) {
  const double alpha = coef(av, ap), beta = coef(bv, bp);
  const int lane = threadIdx.x & 63;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < nruns;
       r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t d = dofs[r], s = sofs[r];
    const int32_t len = lens[r];
    for (int32_t e = lane; e < len; e += 64)
      y[d + e] = alpha == 0.0 ? beta * x[s + e] : alpha * y[d + e] + beta * x[s + e];
  }
}

// First pass of a deterministic dot product: block b sums a fixed strided set.
__global__ void __launch_bounds__(kBS) k_dot_partial(int64_t n, const double *__restrict__ x,
                                                     const double *__restrict__ y, double *__restrict__ part) {
  __shared__ double sh[kBS];
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x; t < n; t += (int64_t)kRedBlocks * kBS) s += x[t] * y[t];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Second pass: one block sums the kRedBlocks partials (plus an optional offset slot).
__global__ void __launch_bounds__(kBS) k_reduce_final(const double *__restrict__ part, int n, double *__restrict__ out) {
  __shared__ double sh[kBS];
  double s = 0.0;
  for (int t = threadIdx.x; t < n; t += kBS) s += part[t];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// Onsite trace partials: diag[t] = offset of onsite block t, nsf[t] its size.
__global__ void __launch_bounds__(kBS) k_trace_partial(int64_t nblk, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ nsf, const double *__restrict__ a,
                                                       double *__restrict__ part) {
  __shared__ double sh[kBS];
  double s = 0.0;
  for (int64_t b = (int64_t)blockIdx.x * kBS + threadIdx.x; b < nblk; b += (int64_t)kRedBlocks * kBS) {
    const int n = nsf[b];
    const double *blk = a + off[b];
    for (int d = 0; d < n; ++d) s += blk[d + (int64_t)n * d];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// In-place symmetrisation over transpose pairs (p, q) with p <= q:
// A_p(a,b) = A_q(b,a) = 0.5*A_p(a,b) + 0.5*A_q(b,a). For p == q only a <= b.
__global__ void k_symmetrise(int64_t npairs, const int64_t *__restrict__ poff, const int64_t *__restrict__ qoff,
                             const uint8_t *__restrict__ n1, const uint8_t *__restrict__ n2, double *__restrict__ A) {
  const int lane = threadIdx.x & 63;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < npairs;
       r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t p = poff[r], q = qoff[r];
    const int m = n1[r], n = n2[r];
    for (int e = lane; e < m * n; e += 64) {
      const int a = e % m, b = e / m;  // element (a,b) of the m x n block p
      if (p == q && a > b) continue;
      const double vp = A[p + a + (int64_t)m * b];
      const double vq = A[q + b + (int64_t)n * a];
      const double v = 0.5 * vp + 0.5 * vq;
      A[p + a + (int64_t)m * b] = v;
      A[q + b + (int64_t)n * a] = v;
    }
  }
}

// B = A^T on a pattern closed under transposition: for each pair (p, q) of
// blocks (i,j,s) and (j,i,-s), B_q = (A_p)^T and B_p = (A_q)^T.
__global__ void k_transpose(int64_t npairs, const int64_t *__restrict__ poff, const int64_t *__restrict__ qoff,
                            const uint8_t *__restrict__ n1, const uint8_t *__restrict__ n2,
                            const double *__restrict__ A, double *__restrict__ B) {
  const int lane = threadIdx.x & 63;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < npairs;
       r += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int64_t p = poff[r], q = qoff[r];
    const int m = n1[r], n = n2[r];
    for (int e = lane; e < m * n; e += 64) {
      const int a = e % m, b = e / m;  // element (a,b) of the m x n block p
      B[q + b + (int64_t)n * a] = A[p + a + (int64_t)m * b];
      B[p + a + (int64_t)m * b] = A[q + b + (int64_t)n * a];
    }
  }
}

// Onsite diagonals set to `v` (matrix must be zero elsewhere beforehand).
__global__ void k_set_diag(int64_t nblk, const int64_t *__restrict__ off, const int32_t *__restrict__ nsf, double v,
                           double *__restrict__ a) {
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += (int64_t)gridDim.x * blockDim.x) {
    const int n = nsf[b];
    for (int d = 0; d < n; ++d) a[off[b] + d + (int64_t)n * d] = v;
  }
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = (float)x[t];
}

__global__ void k_to_f64(int64_t n, const float *__restrict__ x, double *__restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = (double)x[t];
}

// Split-BF16 operand words: the high 16 bits of the float nearest to x
// rounded to BF16 (round to nearest even) in bits 0-15, the BF16 rounding of
// the remainder x - hi in bits 16-31, so that hi + lo carries 16 significant
// bits of x. NaN and infinity keep their high half and a zero remainder.
__device__ inline uint32_t bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return u >> 16;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ inline uint32_t split_bf16(double x) {
  const uint32_t hi = bf16_rne((float)x);
  if ((hi & 0x7f80u) == 0x7f80u) return hi;
  const double rest = x - (double)__uint_as_float(hi << 16);
  return hi | (bf16_rne((float)rest) << 16);
}

__global__ void k_to_bf16x2(int64_t n, const double *__restrict__ x, uint32_t *__restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = split_bf16(x[t]);
}

// In place: FP32 values (a product's FP32 result) -> split-BF16 words.
__global__ void k_f32_to_bf16x2(int64_t n, uint32_t *__restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    y[t] = split_bf16((double)__uint_as_float(y[t]));
}

}  // namespace

namespace cq {

int op_axpby(int64_t n, double alpha, double *y, double beta, const double *x, hipStream_t s) {
  return op_axpby_dev(n, alpha, nullptr, y, beta, nullptr, x, s);
}

int op_axpby_dev(int64_t n, double alpha, const double *alpha_p, double *y, double beta, const double *beta_p,
                 const double *x, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, alpha, alpha_p, y, beta, beta_p, x);
  return hip_check(hipGetLastError(), "axpby launch");
}

int op_scale(int64_t n, double alpha, double *y, hipStream_t s) { return op_scale_dev(n, alpha, nullptr, y, s); }

int op_scale_dev(int64_t n, double alpha, const double *alpha_p, double *y, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, alpha, alpha_p, y);
  return hip_check(hipGetLastError(), "scale launch");
}

int op_copy(int64_t n, double *y, const double *x, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, y, x);
  return hip_check(hipGetLastError(), "copy launch");
}

int op_runs_axpby(const RunMap &m, double alpha, double *y, double beta, const double *x, hipStream_t s) {
  return op_runs_axpby_dev(m, alpha, nullptr, y, beta, nullptr, x, s);
}

int op_runs_axpby_dev(const RunMap &m, double alpha, const double *alpha_p, double *y, double beta,
                      const double *beta_p, const double *x, hipStream_t s) {
  if (m.nruns <= 0) return CQMM_OK;
  const unsigned g = grid_for(m.nruns * 64, 1);
  hipLaunchKernelGGL(k_runs_axpby, dim3(g), dim3(kBS), 0, s, m.nruns, m.dst.as<int64_t>(), m.src.as<int64_t>(),
                     m.len.as<int32_t>(), alpha, alpha_p, y, beta, beta_p, x);
  return hip_check(hipGetLastError(), "mapped sum launch");
}

int op_dot(int64_t n, const double *x, const double *y, double *d_partials, double *d_out, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(kBS), 0, s, n, x, y, d_partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBS), 0, s, d_partials, kRedBlocks, d_out);
  return hip_check(hipGetLastError(), "dot launch");
}

int op_trace(const DiagMap &m, const double *a, double *d_partials, double *d_out, hipStream_t s) {
  hipLaunchKernelGGL(k_trace_partial, dim3(kRedBlocks), dim3(kBS), 0, s, m.nblk, m.off.as<int64_t>(),
                     m.nsf.as<int32_t>(), a, d_partials);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBS), 0, s, d_partials, kRedBlocks, d_out);
  return hip_check(hipGetLastError(), "trace launch");
}

int op_symmetrise(const PairMap &m, double *a, hipStream_t s) {
  if (m.npairs <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_symmetrise, dim3(grid_for(m.npairs * 64, 1)), dim3(kBS), 0, s, m.npairs,
                     m.poff.as<int64_t>(), m.qoff.as<int64_t>(), m.n1.as<uint8_t>(), m.n2.as<uint8_t>(), a);
  return hip_check(hipGetLastError(), "symmetrise launch");
}

template <class T>
int mirror_launch(const TripletPlan &tp, T *O, hipStream_t s) {
  if (tp.n_mirror <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_mirror<T>, dim3(grid_for(tp.n_mirror * 64, 1)), dim3(kBS), 0, s, tp.n_mirror,
                     tp.mir_p.as<int64_t>(), tp.mir_q.as<int64_t>(), tp.mir_n1.as<uint8_t>(), tp.mir_n2.as<uint8_t>(), O);
  return hip_check(hipGetLastError(), "mirror launch");
}

int mirror_result(const TripletPlan &tp, double *O, hipStream_t s) { return mirror_launch(tp, O, s); }
int mirror_result_f32(const TripletPlan &tp, float *O, hipStream_t s) { return mirror_launch(tp, O, s); }

int op_transpose(const PairMap &m, const double *a, double *b, hipStream_t s) {
  if (m.npairs <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_transpose, dim3(grid_for(m.npairs * 64, 1)), dim3(kBS), 0, s, m.npairs,
                     m.poff.as<int64_t>(), m.qoff.as<int64_t>(), m.n1.as<uint8_t>(), m.n2.as<uint8_t>(), a, b);
  return hip_check(hipGetLastError(), "transpose launch");
}

int op_to_f32(int64_t n, const double *x, float *y, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, x, y);
  return hip_check(hipGetLastError(), "to_f32 launch");
}

int op_to_f64(int64_t n, const float *x, double *y, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_to_f64, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, x, y);
  return hip_check(hipGetLastError(), "to_f64 launch");
}

int op_to_bf16x2(int64_t n, const double *x, uint32_t *y, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_to_bf16x2, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, x, y);
  return hip_check(hipGetLastError(), "to_bf16x2 launch");
}

int op_f32_to_bf16x2(int64_t n, uint32_t *y, hipStream_t s) {
  if (n <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_f32_to_bf16x2, dim3(grid_for(n, 4)), dim3(kBS), 0, s, n, y);
  return hip_check(hipGetLastError(), "f32_to_bf16x2 launch");
}

int op_set_diag(const DiagMap &m, double v, double *a, hipStream_t s) {
  if (m.nblk <= 0) return CQMM_OK;
  hipLaunchKernelGGL(k_set_diag, dim3(grid_for(m.nblk, 1)), dim3(kBS), 0, s, m.nblk, m.off.as<int64_t>(),
                     m.nsf.as<int32_t>(), v, a);
  return hip_check(hipGetLastError(), "set_diag launch");
}

int reduction_partials() { return kRedBlocks; }

}  // namespace cq
