// Shared host/device helpers for libsparseops_hip (gfx950 only: wave64, CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "spops_hip.h"

namespace spops {

constexpr int kWave = 64;      // CDNA wavefront width
constexpr int kBlock = 256;    // threads per workgroup for every kernel here (4 waves)

// ---- error reporting ----------------------------------------------------------------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *get_error();

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define SPOPS_HIP_CHECK(expr)                                                                   \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      ::spops::throw_error(e_ == hipErrorOutOfMemory ? SPOPS_ERROR_OUT_OF_MEMORY : SPOPS_ERROR_HIP, \
                           "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define SPOPS_REQUIRE(cond, code, ...)                                                          \
  do {                                                                                          \
    if (!(cond)) ::spops::throw_error(code, __VA_ARGS__);                                        \
  } while (0)

// Runs `f` and maps exceptions to status codes (no exception crosses the C ABI).
template <typename F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    set_error("%s", e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return SPOPS_ERROR_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    set_error("%s", e.what());
    return SPOPS_ERROR_INVALID_VALUE;
  }
}

// ---- device memory ----------------------------------------------------------------------------
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
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    SPOPS_HIP_CHECK(hipMalloc(&p, n));
    bytes = n;
  }
  template <typename T>
  void upload(const std::vector<T> &v) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) SPOPS_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload_raw(const void *src, size_t n) {
    alloc(n);
    if (n) SPOPS_HIP_CHECK(hipMemcpy(p, src, n, hipMemcpyHostToDevice));
  }
  template <typename T>
  T *as() const { return static_cast<T *>(p); }
};

// ---- value conversion (host) --------------------------------------------------------------------
// double -> float with round-to-odd, so a following float -> bf16 round-to-nearest-even
// equals a direct double -> bf16 round-to-nearest-even (float keeps > 2 extra bits).
inline float to_float_round_odd(double d) {
  float f = (float)d;
  if ((double)f != d && f != 0.0f && f == f && f - f == 0.0f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 1u) == 0u) {
      bool toward_up = ((double)f < d) == (f > 0.0f);
      u = toward_up ? u + 1 : u - 1;
    }
    std::memcpy(&f, &u, 4);
  }
  return f;
}

inline uint16_t float_to_bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

inline uint16_t double_to_bf16_rne(double d) { return float_to_bf16_rne(to_float_round_odd(d)); }

inline size_t value_bytes(int vt) { return vt == SPOPS_VALUE_F64 ? 8 : vt == SPOPS_VALUE_F32 ? 4 : 2; }
inline size_t xy_bytes(int vt) { return vt == SPOPS_VALUE_F64 ? 8 : 4; }

// ---- device helpers -------------------------------------------------------------------------------
__device__ __forceinline__ float bf16_to_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Raw buffer loads: an offset past the resource's size returns 0 and touches no memory, so
// masked lanes cost no traffic and the loads need no branches (all loads of an unrolled
// loop issue back to back).  Offsets are bytes, 32-bit; resources are kept below 2 GiB.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr uint32_t kOOB = 0x7FFFFFF0u;           // an offset no resource reaches
constexpr uint64_t kMaxBufBytes = 0x7FFFFFF0ull;
constexpr int kAuxNT = 2;                        // non-temporal (streamed once)
typedef unsigned bu2 __attribute__((ext_vector_type(2)));
typedef unsigned bu4 __attribute__((ext_vector_type(4)));

// p and bytes must be the same in every lane of the wave: they are moved to scalar registers
// (a resource in vector registers would make the compiler loop over lanes for every load).
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, uint64_t bytes) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a), hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  const int n = __builtin_amdgcn_readfirstlane((int)(bytes < kMaxBufBytes ? bytes : kMaxBufBytes));
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(((uint64_t)hi << 32) | lo), (short)0, n,
                                           0x00020000);
}

template <typename T, int aux = 0>
__device__ __forceinline__ T bload(rsrc_t r, uint32_t off) {
  static_assert(sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "load size");
  T out;
  if constexpr (sizeof(T) == 2) {
    unsigned short v = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, aux);
    __builtin_memcpy(&out, &v, 2);
  } else if constexpr (sizeof(T) == 4) {
    unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, aux);
    __builtin_memcpy(&out, &v, 4);
  } else if constexpr (sizeof(T) == 8) {
    bu2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, aux);
    __builtin_memcpy(&out, &v, 8);
  } else {
    bu4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, aux);
    __builtin_memcpy(&out, &v, 16);
  }
  return out;
}

// Workgroups for `items` work items at `per` items per workgroup of kBlock threads.  One
// dispatch holds fewer than 2^32 work items; a launch past that is refused (kernels with more
// items than that stride over them with a capped grid instead).
inline unsigned grid_blocks(int64_t items, int64_t per) {
  const int64_t b = (items + per - 1) / per;
  SPOPS_REQUIRE(b * kBlock < ((int64_t)1 << 32), SPOPS_ERROR_LIMIT,
                "a launch of %lld workgroups would exceed 2^32 work items", (long long)b);
  return (unsigned)b;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x >> 6); }

// Error-free transformation a + b = s + e (Knuth TwoSum), used for compensated FP32 sums.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float z = s - a;
  e = (a - (s - z)) + (b - z);
}

// Accumulator: plain for FP64, compensated (Neumaier-style with TwoSum) for FP32.
template <typename Acc, bool Comp>
struct Accum;

template <>
struct Accum<double, false> {
  double s = 0.0;
  __device__ __forceinline__ void add(double v) { s += v; }
  __device__ __forceinline__ void fma(double a, double b) { s = __builtin_fma(a, b, s); }
  __device__ __forceinline__ double value() const { return s; }
  __device__ __forceinline__ void merge_xor(int mask) { s += __shfl_xor(s, mask, kWave); }
};

template <>
struct Accum<float, false> {
  float s = 0.0f;
  __device__ __forceinline__ void add(float v) { s += v; }
  __device__ __forceinline__ void fma(float a, float b) { s = __builtin_fmaf(a, b, s); }
  __device__ __forceinline__ float value() const { return s; }
  __device__ __forceinline__ void merge_xor(int mask) { s += __shfl_xor(s, mask, kWave); }
};

template <>
struct Accum<float, true> {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    float t, e;
    two_sum(s, v, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void fma(float a, float b) {
    float p = a * b;
    float pe = __builtin_fmaf(a, b, -p);   // exact product error (TwoProd)
    add(p);
    c += pe;
  }
  __device__ __forceinline__ float value() const { return s + c; }
  __device__ __forceinline__ void merge_xor(int mask) {
    float os = __shfl_xor(s, mask, kWave), oc = __shfl_xor(c, mask, kWave);
    float t, e;
    two_sum(s, os, t, e);
    s = t;
    c = c + oc + e;
  }
};

}  // namespace spops
