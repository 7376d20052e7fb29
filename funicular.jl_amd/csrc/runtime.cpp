// Runtime half of libfunicular_hip.so: errors, counters, memory, streams,
// events and copies. The device contract this implements is
// src/device_api.jl:12-138 of Funicular.jl; the CUDA extension
// (ext/FunicularCUDAExt.jl) shows the behaviour each call has to match.
#include "fhip_internal.hpp"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

namespace fhip {

static thread_local std::string g_last_error;

static std::atomic<uint64_t> c_device_allocs{0}, c_device_bytes{0}, c_host_allocs{0},
    c_host_bytes{0}, c_h2d{0}, c_d2h{0}, c_d2d{0}, c_copies{0}, c_launches{0}, c_syncs{0};

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t err, const char *what) {
  // Clear the sticky per-thread error so the next call starts clean.
  (void)hipGetLastError();
  return fail(FHIP_E_HIP, "%s: %s (%s)", what, hipGetErrorString(err), hipGetErrorName(err));
}

void count_launch() { c_launches.fetch_add(1, std::memory_order_relaxed); }

int check_launch(const char *what) {
  count_launch();
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return hip_fail(err, what);
  return FHIP_OK;
}

size_t eltype_size(int type) {
  switch (type) {
    case FHIP_F32: return 4;
    case FHIP_F64: return 8;
    case FHIP_C32: return 8;
    case FHIP_C64: return 16;
    default: return 0;
  }
}

}  // namespace fhip

using namespace fhip;

#define HIP_TRY(expr, what)                        \
  do {                                             \
    hipError_t _e = (expr);                        \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

extern "C" {

const char *fhip_last_error(void) { return g_last_error.c_str(); }

const char *fhip_version(void) { return "funicular_hip 0.1.0 (gfx950)"; }

int fhip_device_count(int *count) {
  if (!count) return fail(FHIP_E_ARGUMENT, "fhip_device_count: null output pointer");
  HIP_TRY(hipGetDeviceCount(count), "hipGetDeviceCount");
  return FHIP_OK;
}

int fhip_set_device(int device) {
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  return FHIP_OK;
}

int fhip_device_arch(char *buf, size_t len) {
  if (!buf || len == 0) return fail(FHIP_E_ARGUMENT, "fhip_device_arch: empty buffer");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  snprintf(buf, len, "%s", prop.gcnArchName);
  return FHIP_OK;
}

int fhip_device_sync(void) {
  c_syncs.fetch_add(1, std::memory_order_relaxed);
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return FHIP_OK;
}

int fhip_malloc(void **ptr, size_t nbytes) {
  if (!ptr) return fail(FHIP_E_ARGUMENT, "fhip_malloc: null output pointer");
  *ptr = nullptr;
  if (nbytes == 0) nbytes = 256;  // a valid, distinct pointer for empty buffers
  HIP_TRY(hipMalloc(ptr, nbytes), "hipMalloc");
  // The allocation is used from streams other than the one that made it, so
  // it has happened before this returns (ext/FunicularCUDAExt.jl:32-40).
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize after hipMalloc");
  c_device_allocs.fetch_add(1, std::memory_order_relaxed);
  c_device_bytes.fetch_add(nbytes, std::memory_order_relaxed);
  return FHIP_OK;
}

int fhip_free(void *ptr) {
  if (!ptr) return FHIP_OK;
  size_t nbytes = 0;
  void *base = nullptr;
  if (hipMemGetAddressRange((hipDeviceptr_t *)&base, &nbytes, (hipDeviceptr_t)ptr) == hipSuccess &&
      base == ptr)
    c_device_bytes.fetch_sub(nbytes <= c_device_bytes.load() ? nbytes : c_device_bytes.load(),
                             std::memory_order_relaxed);
  (void)hipGetLastError();
  HIP_TRY(hipFree(ptr), "hipFree");
  return FHIP_OK;
}

int fhip_host_alloc(void **ptr, size_t nbytes) {
  if (!ptr) return fail(FHIP_E_ARGUMENT, "fhip_host_alloc: null output pointer");
  *ptr = nullptr;
  if (nbytes == 0) nbytes = 64;
  HIP_TRY(hipHostMalloc(ptr, nbytes, hipHostMallocDefault), "hipHostMalloc");
  c_host_allocs.fetch_add(1, std::memory_order_relaxed);
  c_host_bytes.fetch_add(nbytes, std::memory_order_relaxed);
  return FHIP_OK;
}

int fhip_host_free(void *ptr) {
  if (!ptr) return FHIP_OK;
  HIP_TRY(hipHostFree(ptr), "hipHostFree");
  return FHIP_OK;
}

int fhip_host_is_pinned(int *pinned, const void *ptr) {
  if (!pinned) return fail(FHIP_E_ARGUMENT, "fhip_host_is_pinned: null output pointer");
  *pinned = 0;
  if (!ptr) return FHIP_OK;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, ptr);
  if (e == hipErrorInvalidValue) {
    // Pageable memory the runtime has never seen: not page-locked. The query
    // leaves that error behind; clear it so the next call does not report it.
    (void)hipGetLastError();
    return FHIP_OK;
  }
  HIP_TRY(e, "hipPointerGetAttributes");
  *pinned = attr.type == hipMemoryTypeHost ? 1 : 0;
  return FHIP_OK;
}

int fhip_stream_create(void **stream) {
  if (!stream) return fail(FHIP_E_ARGUMENT, "fhip_stream_create: null output pointer");
  hipStream_t s;
  // Non-blocking: a default-flag stream would serialize against the null
  // stream (ext/FunicularCUDAExt.jl:68-71). Default priority only.
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  *stream = (void *)s;
  return FHIP_OK;
}

int fhip_stream_destroy(void *stream) {
  if (!stream) return FHIP_OK;
  HIP_TRY(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
  return FHIP_OK;
}

int fhip_stream_sync(void *stream) {
  c_syncs.fetch_add(1, std::memory_order_relaxed);
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return FHIP_OK;
}

int fhip_event_create(void **event, int disable_timing) {
  if (!event) return fail(FHIP_E_ARGUMENT, "fhip_event_create: null output pointer");
  hipEvent_t e;
  HIP_TRY(hipEventCreateWithFlags(&e, disable_timing ? hipEventDisableTiming : hipEventDefault),
          "hipEventCreateWithFlags");
  *event = (void *)e;
  return FHIP_OK;
}

int fhip_event_destroy(void *event) {
  if (!event) return FHIP_OK;
  HIP_TRY(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
  return FHIP_OK;
}

int fhip_event_record(void *event, void *stream) {
  if (!event) return fail(FHIP_E_ARGUMENT, "fhip_event_record: null event");
  HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
  return FHIP_OK;
}

int fhip_stream_wait_event(void *stream, void *event) {
  if (!event) return fail(FHIP_E_ARGUMENT, "fhip_stream_wait_event: null event");
  HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0), "hipStreamWaitEvent");
  return FHIP_OK;
}

int fhip_event_sync(void *event) {
  if (!event) return fail(FHIP_E_ARGUMENT, "fhip_event_sync: null event");
  c_syncs.fetch_add(1, std::memory_order_relaxed);
  HIP_TRY(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
  return FHIP_OK;
}

int fhip_event_query(void *event, int *done) {
  if (!event || !done) return fail(FHIP_E_ARGUMENT, "fhip_event_query: null argument");
  hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipSuccess) {
    *done = 1;
    return FHIP_OK;
  }
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = 0;
    return FHIP_OK;
  }
  return hip_fail(e, "hipEventQuery");
}

int fhip_event_elapsed(float *ms, void *start, void *stop) {
  if (!ms || !start || !stop) return fail(FHIP_E_ARGUMENT, "fhip_event_elapsed: null argument");
  HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
  return FHIP_OK;
}

int fhip_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                        size_t width_bytes, size_t rows, int kind, void *stream) {
  if (width_bytes == 0 || rows == 0) return FHIP_OK;
  if (!dst || !src) return fail(FHIP_E_ARGUMENT, "fhip_memcpy2d_async: null pointer");
  if (width_bytes > dpitch || width_bytes > spitch)
    return fail(FHIP_E_ARGUMENT,
                "fhip_memcpy2d_async: a line of %zu bytes does not fit pitches %zu (dst) / %zu (src)",
                width_bytes, dpitch, spitch);
  hipMemcpyKind k;
  std::atomic<uint64_t> *counter;
  switch (kind) {
    case FHIP_H2D: k = hipMemcpyHostToDevice; counter = &c_h2d; break;
    case FHIP_D2H: k = hipMemcpyDeviceToHost; counter = &c_d2h; break;
    case FHIP_D2D: k = hipMemcpyDeviceToDevice; counter = &c_d2d; break;
    case FHIP_H2H: k = hipMemcpyHostToHost; counter = &c_d2d; break;
    default: return fail(FHIP_E_ARGUMENT, "fhip_memcpy2d_async: unknown copy kind %d", kind);
  }
  c_copies.fetch_add(1, std::memory_order_relaxed);
  counter->fetch_add(width_bytes * rows, std::memory_order_relaxed);
  if (rows == 1 || (dpitch == width_bytes && spitch == width_bytes)) {
    HIP_TRY(hipMemcpyAsync(dst, src, width_bytes * rows, k, (hipStream_t)stream), "hipMemcpyAsync");
  } else {
    HIP_TRY(hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, rows, k, (hipStream_t)stream),
            "hipMemcpy2DAsync");
  }
  return FHIP_OK;
}

int fhip_get_counters(fhip_counters *out) {
  if (!out) return fail(FHIP_E_ARGUMENT, "fhip_get_counters: null output pointer");
  out->device_allocs = c_device_allocs.load();
  out->device_bytes = c_device_bytes.load();
  out->host_allocs = c_host_allocs.load();
  out->host_bytes = c_host_bytes.load();
  out->h2d_bytes = c_h2d.load();
  out->d2h_bytes = c_d2h.load();
  out->d2d_bytes = c_d2d.load();
  out->copies = c_copies.load();
  out->kernel_launches = c_launches.load();
  out->host_syncs = c_syncs.load();
  return FHIP_OK;
}

int fhip_reset_counters(void) {
  c_h2d = 0;
  c_d2h = 0;
  c_d2d = 0;
  c_copies = 0;
  c_launches = 0;
  c_syncs = 0;
  c_device_allocs = 0;
  c_host_allocs = 0;
  return FHIP_OK;
}

}  // extern "C"
