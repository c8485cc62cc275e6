/* Glue between the reference-shaped C++ API (SpOps.h, spgemm.h, ...) and the C ABI of
 * libsparseops_hip.so (spops_hip.h).  Header-only; needs no HIP headers.
 *
 * Device residency: the first SpOps_<fmt> call on a host matrix uploads it (converted to the
 * device layout) and keeps the device copy keyed by the address of its first index array;
 * later calls reuse it.  The host matrix must not change while cached -- call
 * spops_dropin_release(ptr) (or spops_dropin_release_all()) after modifying or freeing it.
 * x and y are staged through grow-only device buffers on every call (synchronous, as the
 * reference's functions are).  Status codes from the library become std::runtime_error. */
#ifndef SPOPS_DROPIN_GLUE_H
#define SPOPS_DROPIN_GLUE_H

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "spops_hip.h"

namespace spops_dropin {

inline void check(int status, const char *what) {
  if (status != SPOPS_SUCCESS)
    throw std::runtime_error(std::string(what) + ": " + spops_status_string(status) + ": " + spops_last_error());
}

template <typename I>
constexpr int index_code() {
  static_assert(std::is_integral<I>::value && (sizeof(I) == 4 || sizeof(I) == 8), "index type must be 32/64-bit");
  return sizeof(I) == 4 ? SPOPS_INDEX_I32 : SPOPS_INDEX_I64;
}

template <typename V>
constexpr int value_code() {
  static_assert(std::is_same<V, float>::value || std::is_same<V, double>::value, "value type must be float/double");
  return std::is_same<V, double>::value ? SPOPS_VALUE_F64 : SPOPS_VALUE_F32;
}

// Grow-only device buffer.
struct DeviceBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t n) {
    if (n > bytes) {
      if (p) spops_free(p);
      p = nullptr;
      check(spops_malloc(&p, n), "spops_malloc");
      bytes = n;
    }
    return p;
  }
  ~DeviceBuffer() {
    if (p) spops_free(p);
  }
};

struct State {
  std::mutex mu;
  std::unordered_map<const void *, spops_matrix> matrices;
  DeviceBuffer x, y, scratch[4];
};

// Never destroyed: device memory still held at exit is returned with the process, and freeing
// it from a static destructor could run after the HIP runtime has shut down.
inline State &state() {
  static State *s = new State;
  return *s;
}

// Device matrix for `key`, created by `make` on first use.
template <typename Make>
spops_matrix cached(const void *key, Make &&make) {
  State &s = state();
  auto it = s.matrices.find(key);
  if (it != s.matrices.end()) return it->second;
  spops_matrix h = nullptr;
  check(make(&h), "create device matrix");
  s.matrices.emplace(key, h);
  return h;
}

// y = alpha*A*x + beta*y with host x, y (caller holds state().mu).
template <typename V>
void spmv_host(spops_matrix A, int64_t n_rows, int64_t n_cols, V alpha, const V *x, V beta, V *y) {
  State &s = state();
  void *dx = s.x.get((size_t)(n_cols ? n_cols : 1) * sizeof(V));
  void *dy = s.y.get((size_t)(n_rows ? n_rows : 1) * sizeof(V));
  check(spops_memcpy_htod(dx, x, (size_t)n_cols * sizeof(V), nullptr), "copy x");
  if (beta != V(0)) check(spops_memcpy_htod(dy, y, (size_t)n_rows * sizeof(V), nullptr), "copy y");
  check(spops_spmv(A, (double)alpha, dx, (double)beta, dy, nullptr), "spops_spmv");
  check(spops_memcpy_dtoh(y, dy, (size_t)n_rows * sizeof(V), nullptr), "copy y back");
}

// Upload a host array to a scratch device buffer (slot 0..3), widening indices to int64.
template <typename I>
const int64_t *upload_index64(int slot, const I *a, int64_t n) {
  std::vector<int64_t> w(a, a + n);
  void *d = state().scratch[slot].get((size_t)(n ? n : 1) * 8);
  check(spops_memcpy_htod(d, w.data(), (size_t)n * 8, nullptr), "upload");
  return static_cast<const int64_t *>(d);
}

}  // namespace spops_dropin

/* Drop the device copy cached for a host matrix (key: its first index array). */
inline void spops_dropin_release(const void *key) {
  auto &s = spops_dropin::state();
  std::lock_guard<std::mutex> g(s.mu);
  auto it = s.matrices.find(key);
  if (it != s.matrices.end()) {
    spops_matrix_destroy(it->second);
    s.matrices.erase(it);
  }
}

inline void spops_dropin_release_all() {
  auto &s = spops_dropin::state();
  std::lock_guard<std::mutex> g(s.mu);
  for (auto &kv : s.matrices) spops_matrix_destroy(kv.second);
  s.matrices.clear();
}

#endif
