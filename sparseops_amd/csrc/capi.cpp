// C ABI: status strings, thread-local error detail, device memory / stream / event helpers.
#include <cstdio>

#include "common.hpp"

namespace spops {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

const char *get_error() { return g_error.c_str(); }

void throw_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

}  // namespace spops

using namespace spops;

extern "C" const char *spops_status_string(int s) {
  switch (s) {
    case SPOPS_SUCCESS: return "success";
    case SPOPS_ERROR_INVALID_VALUE: return "invalid value";
    case SPOPS_ERROR_OUT_OF_MEMORY: return "out of memory";
    case SPOPS_ERROR_HIP: return "HIP runtime error";
    case SPOPS_ERROR_NOT_SUPPORTED: return "not supported";
    case SPOPS_ERROR_LIMIT: return "size limit exceeded";
    default: return "unknown status";
  }
}

extern "C" const char *spops_last_error(void) { return get_error(); }

extern "C" int spops_version(int *major, int *minor, int *patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return SPOPS_SUCCESS;
}

extern "C" int spops_device_count(int *count) {
  return guarded([&]() {
    SPOPS_REQUIRE(count, SPOPS_ERROR_INVALID_VALUE, "null count");
    SPOPS_HIP_CHECK(hipGetDeviceCount(count));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_set_device(int device) {
  return guarded([&]() {
    SPOPS_HIP_CHECK(hipSetDevice(device));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_malloc(void **p, size_t bytes) {
  return guarded([&]() {
    SPOPS_REQUIRE(p, SPOPS_ERROR_INVALID_VALUE, "null pointer");
    SPOPS_HIP_CHECK(hipMalloc(p, bytes ? bytes : 16));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_free(void *p) {
  return guarded([&]() {
    if (p) SPOPS_HIP_CHECK(hipFree(p));
    return SPOPS_SUCCESS;
  });
}

static int copy(void *dst, const void *src, size_t n, hipMemcpyKind k, void *stream) {
  return guarded([&]() {
    if (n == 0) return SPOPS_SUCCESS;
    SPOPS_REQUIRE(dst && src, SPOPS_ERROR_INVALID_VALUE, "null pointer");
    SPOPS_HIP_CHECK(hipMemcpyAsync(dst, src, n, k, static_cast<hipStream_t>(stream)));
    if (k != hipMemcpyDeviceToDevice) SPOPS_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_memcpy_htod(void *d, const void *s, size_t n, void *st) { return copy(d, s, n, hipMemcpyHostToDevice, st); }
extern "C" int spops_memcpy_dtoh(void *d, const void *s, size_t n, void *st) { return copy(d, s, n, hipMemcpyDeviceToHost, st); }
extern "C" int spops_memcpy_dtod(void *d, const void *s, size_t n, void *st) { return copy(d, s, n, hipMemcpyDeviceToDevice, st); }

extern "C" int spops_memset(void *d, int v, size_t n, void *stream) {
  return guarded([&]() {
    if (n) SPOPS_HIP_CHECK(hipMemsetAsync(d, v, n, static_cast<hipStream_t>(stream)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_stream_synchronize(void *stream) {
  return guarded([&]() {
    SPOPS_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_stream_create(void **stream) {
  return guarded([&]() {
    SPOPS_REQUIRE(stream, SPOPS_ERROR_INVALID_VALUE, "null pointer");
    hipStream_t s;
    SPOPS_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_stream_destroy(void *stream) {
  return guarded([&]() {
    if (stream) SPOPS_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_event_create(void **ev) {
  return guarded([&]() {
    SPOPS_REQUIRE(ev, SPOPS_ERROR_INVALID_VALUE, "null pointer");
    hipEvent_t e;
    SPOPS_HIP_CHECK(hipEventCreate(&e));
    *ev = e;
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_event_destroy(void *ev) {
  return guarded([&]() {
    if (ev) SPOPS_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(ev)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_event_record(void *ev, void *stream) {
  return guarded([&]() {
    SPOPS_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream)));
    return SPOPS_SUCCESS;
  });
}

extern "C" int spops_event_elapsed_ms(float *ms, void *start, void *stop) {
  return guarded([&]() {
    SPOPS_REQUIRE(ms, SPOPS_ERROR_INVALID_VALUE, "null pointer");
    SPOPS_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    SPOPS_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return SPOPS_SUCCESS;
  });
}
