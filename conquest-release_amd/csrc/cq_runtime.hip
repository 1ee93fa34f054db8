// Library state, error reporting and device-memory helpers of libcqmm.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/cqmm.h"
#include "cq_common.h"
#include "cq_device.h"

namespace cq {

static thread_local std::string g_last_error;

int set_error(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

int set_errorf(int status, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

static Runtime g_rt;

Runtime &runtime() { return g_rt; }

static thread_local hipStream_t t_stream = nullptr;
hipStream_t stream() { return t_stream ? t_stream : g_rt.stream; }
void set_thread_stream(hipStream_t s) { t_stream = s; }

int ensure_init() {
  if (g_rt.initialised) return CQMM_OK;
  return cqmm_init(-1);
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return CQMM_OK;
  return set_errorf(e == hipErrorOutOfMemory ? CQMM_ERR_NOMEM : CQMM_ERR_HIP, "%s: %s", what,
                    hipGetErrorString(e));
}

}  // namespace cq

extern "C" {

const char *cqmm_last_error(void) { return cq::g_last_error.c_str(); }

int cqmm_init(int device) {
  return cq::guarded(__func__, [&]() -> int {
    cq::Runtime &rt = cq::g_rt;
    if (rt.initialised) {
      if (device < 0 || device == rt.device) return CQMM_OK;
      return cq::set_error(CQMM_ERR_STATE, "cqmm_init: already initialised on another device");
    }
    int n = 0;
    int st = cq::hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (st) return st;
    if (n <= 0) return cq::set_error(CQMM_ERR_HIP, "cqmm_init: no HIP device visible");
    int dev = device;
    if (dev < 0) {
      st = cq::hip_check(hipGetDevice(&dev), "hipGetDevice");
      if (st) return st;
    }
    if (dev >= n) return cq::set_errorf(CQMM_ERR_ARG, "cqmm_init: device %d of %d", dev, n);
    st = cq::hip_check(hipSetDevice(dev), "hipSetDevice");
    if (st) return st;
    st = cq::noalloc_setup(dev);
    if (st) return st;
    // Default-priority stream of the library (no priority other than the default).
    st = cq::hip_check(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking), "hipStreamCreate");
    if (st) return st;
    hipDeviceProp_t prop;
    st = cq::hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (st) return st;
    rt.n_cu = prop.multiProcessorCount;
    rt.device = dev;
    rt.initialised = true;
    return CQMM_OK;
  });
}

int cqmm_finalize(void) {
  return cq::guarded(__func__, [&]() -> int {
    cq::Runtime &rt = cq::g_rt;
    if (!rt.initialised) return CQMM_OK;
    (void)hipStreamSynchronize(rt.stream);
    (void)hipStreamDestroy(rt.stream);
    rt.stream = nullptr;
    rt.initialised = false;
    return CQMM_OK;
  });
}

int cqmm_device_alloc(void **ptr, int64_t bytes) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ptr || bytes < 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_device_alloc: bad argument");
    int st = cq::ensure_init();
    if (st) return st;
    *ptr = nullptr;
    if (bytes == 0) return CQMM_OK;
    return cq::hip_check(hipMalloc(ptr, (size_t)bytes), "hipMalloc");
  });
}

int cqmm_device_free(void *ptr) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ptr) return CQMM_OK;
    return cq::hip_check(hipFree(ptr), "hipFree");
  });
}

int cqmm_memcpy_h2d(void *dst, const void *src, int64_t bytes) {
  return cq::guarded(__func__, [&]() -> int {
    int st = cq::ensure_init();
    if (st) return st;
    if (bytes <= 0) return CQMM_OK;
    st = cq::hip_check(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, cq::stream()), "hipMemcpy H2D");
    if (st) return st;
    return cq::hip_check(hipStreamSynchronize(cq::stream()), "hipStreamSynchronize");
  });
}

int cqmm_memcpy_d2h(void *dst, const void *src, int64_t bytes) {
  return cq::guarded(__func__, [&]() -> int {
    int st = cq::ensure_init();
    if (st) return st;
    if (bytes <= 0) return CQMM_OK;
    st = cq::hip_check(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, cq::stream()), "hipMemcpy D2H");
    if (st) return st;
    return cq::hip_check(hipStreamSynchronize(cq::stream()), "hipStreamSynchronize");
  });
}

namespace {
// Grid-stride copy, 16 bytes per lane per step.
__global__ void k_stream_copy(double2 *__restrict__ dst, const double2 *__restrict__ src, int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
}
__global__ void k_stream_copy_tail(double *__restrict__ dst, const double *__restrict__ src, int64_t i) {
  dst[i] = src[i];
}
}  // namespace

int cqmm_device_copy(double *dst, const double *src, int64_t n) {
  return cq::guarded(__func__, [&]() -> int {
    if (n < 0 || (n > 0 && (!dst || !src))) return cq::set_error(CQMM_ERR_ARG, "cqmm_device_copy: bad argument");
    if (((uintptr_t)dst | (uintptr_t)src) & 15u)
      return cq::set_error(CQMM_ERR_ARG, "cqmm_device_copy: pointers must be 16-byte aligned");
    int st = cq::ensure_init();
    if (st) return st;
    if (n == 0) return CQMM_OK;
    const int64_t n2 = n / 2;
    hipStream_t s = cq::stream();
    if (n2 > 0) {
      const int64_t want = (n2 + 255) / 256;
      const unsigned grid = (unsigned)(want < 8192 ? want : 8192);
      hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (double2 *)dst, (const double2 *)src, n2);
    }
    if (n & 1) hipLaunchKernelGGL(k_stream_copy_tail, dim3(1), dim3(1), 0, s, dst, src, n - 1);
    return cq::hip_check(hipGetLastError(), "cqmm_device_copy launch");
  });
}

static hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;

int cqmm_timer_start(void) {
  return cq::guarded(__func__, [&]() -> int {
    int st = cq::ensure_init();
    if (st) return st;
    if (!g_ev0 && (st = cq::hip_check(hipEventCreate(&g_ev0), "hipEventCreate"))) return st;
    if (!g_ev1 && (st = cq::hip_check(hipEventCreate(&g_ev1), "hipEventCreate"))) return st;
    return cq::hip_check(hipEventRecord(g_ev0, cq::stream()), "hipEventRecord");
  });
}

int cqmm_timer_stop(float *ms) {
  return cq::guarded(__func__, [&]() -> int {
    if (!g_ev0 || !g_ev1 || !ms) return cq::set_error(CQMM_ERR_STATE, "cqmm_timer_stop: timer not started");
    int st = cq::hip_check(hipEventRecord(g_ev1, cq::stream()), "hipEventRecord");
    if (st) return st;
    if ((st = cq::hip_check(hipEventSynchronize(g_ev1), "hipEventSynchronize"))) return st;
    return cq::hip_check(hipEventElapsedTime(ms, g_ev0, g_ev1), "hipEventElapsedTime");
  });
}

int cqmm_synchronize(void) {
  return cq::guarded(__func__, [&]() -> int {
    int st = cq::ensure_init();
    if (st) return st;
    return cq::hip_check(hipStreamSynchronize(cq::stream()), "hipStreamSynchronize");
  });
}

}  // extern "C"
