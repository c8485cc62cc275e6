/* Glue between the reference-shaped C++ API (SpOps.h, spgemm.h, ...) and the C ABI of
 * libsparseops_hip.so (spops_hip.h).  Header-only; needs no HIP headers.
 *
 * Device residency.  The first call on a host matrix uploads it (converted to the device
 * layout) and keeps the device copy; later calls on the same host matrix reuse it.  A cache
 * entry is found by the address of the matrix's first index array and is reused only when the
 * matrix still matches what the copy was made from: every dimension and count, the type
 * sizes, the addresses of all its arrays, and a 64-bit digest of their contents.  The digest
 * mode comes from the environment variable SPOPS_DROPIN_VERIFY, read once, or from
 * spops_dropin_set_verify():
 *   "sampled" (default)  65 entries of each array at an even stride (first and last
 *                        included): a freed matrix whose address is reused by another one,
 *                        or one re-read from a different file, is caught; an in-place edit of
 *                        an entry the sample misses is not -- call spops_dropin_release(ptr)
 *                        after editing a cached matrix in place;
 *   "full"               every byte of every array on every call (one host pass over the
 *                        matrix per call);
 *   "off"                no caching: each call uploads the matrix and frees the copy.
 * A stale entry is destroyed and rebuilt.  x and y are staged through grow-only device
 * buffers on every call (synchronous, as the reference's functions are).  Status codes from
 * the library become std::runtime_error. */
#ifndef SPOPS_DROPIN_GLUE_H
#define SPOPS_DROPIN_GLUE_H

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "spops_hip.h"

#define SPOPS_DROPIN_VERIFY_SAMPLED 0
#define SPOPS_DROPIN_VERIFY_FULL 1
#define SPOPS_DROPIN_VERIFY_OFF 2

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
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  void *get(size_t n) {
    if (n > bytes) {
      if (p) spops_free(p);
      p = nullptr;
      bytes = 0;
      check(spops_malloc(&p, n), "spops_malloc");
      bytes = n;
    }
    return p;
  }
  ~DeviceBuffer() {
    if (p) spops_free(p);
  }
};

// Content digest of host arrays (a multiply-rotate mix over 64-bit words).
struct Digest {
  uint64_t h = 0x9e3779b97f4a7c15ull;
  void word(uint64_t w) {
    h ^= w * 0xbf58476d1ce4e5b9ull;
    h = ((h << 31) | (h >> 33)) * 0x94d049bb133111ebull;
  }
  void bytes(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, b + i, 8);
      word(w);
    }
    uint64_t t = 0;
    if (i < n) std::memcpy(&t, b + i, n - i);
    word(t ^ ((uint64_t)n << 56));
  }
  // n elements of `elem` bytes: all of them (full) or 65 at an even stride, first and last
  // included (sampled)
  void array(const void *p, size_t elem, int64_t n, int mode) {
    word((uint64_t)n);
    if (p == nullptr || n <= 0) return;
    const unsigned char *b = static_cast<const unsigned char *>(p);
    if (mode == SPOPS_DROPIN_VERIFY_FULL || n <= 65) {
      bytes(b, (size_t)n * elem);
      return;
    }
    for (int s = 0; s <= 64; ++s) bytes(b + (size_t)((n - 1) * (int64_t)s / 64) * elem, elem);
  }
};

// What a cached device copy was made from.
struct Source {
  int kind = 0;
  int64_t dims[8] = {0};
  const void *arrays[6] = {nullptr};
  uint64_t digest = 0;
  bool operator==(const Source &o) const {
    if (kind != o.kind || digest != o.digest) return false;
    for (int i = 0; i < 8; ++i)
      if (dims[i] != o.dims[i]) return false;
    for (int i = 0; i < 6; ++i)
      if (arrays[i] != o.arrays[i]) return false;
    return true;
  }
};

// A cached device copy: a library matrix or cluster handle, or raw device arrays (LeSpGEMM's B).
struct Entry {
  Source src;
  spops_matrix m = nullptr;
  spops_cluster cl = nullptr;
  DeviceBuffer buf[3];
  ~Entry() {
    if (m) spops_matrix_destroy(m);
    if (cl) spops_cluster_destroy(cl);
  }
};

inline int verify_from_env() {
  const char *e = std::getenv("SPOPS_DROPIN_VERIFY");
  if (e == nullptr) return SPOPS_DROPIN_VERIFY_SAMPLED;
  if (std::strcmp(e, "full") == 0) return SPOPS_DROPIN_VERIFY_FULL;
  if (std::strcmp(e, "off") == 0) return SPOPS_DROPIN_VERIFY_OFF;
  return SPOPS_DROPIN_VERIFY_SAMPLED;
}

struct State {
  std::mutex mu;
  int verify = verify_from_env();
  std::unordered_map<const void *, std::unique_ptr<Entry>> entries;
  std::unique_ptr<Entry> transient;   // "off" mode: the last call's copy, freed by the next call
  DeviceBuffer x, y, scratch[4];
};

// Never destroyed: device memory still held at exit is returned with the process, and freeing
// it from a static destructor could run after the HIP runtime has shut down.
inline State &state() {
  static State *s = new State;
  return *s;
}

// The cached entry for host matrix `key` described by `src` (whose digest is filled here),
// rebuilt by `make(Entry&)` when missing or stale.  Caller holds state().mu.
template <typename Fill, typename Make>
Entry &cached_entry(const void *key, Source src, Fill &&fill_digest, Make &&make) {
  State &s = state();
  Digest d;
  fill_digest(d, s.verify);
  src.digest = d.h;
  if (s.verify != SPOPS_DROPIN_VERIFY_OFF) {
    auto it = s.entries.find(key);
    if (it != s.entries.end()) {
      if (it->second->src == src) return *it->second;
      s.entries.erase(it);   // stale: the host matrix behind this address has changed
    }
  }
  std::unique_ptr<Entry> e(new Entry);
  e->src = src;
  make(*e);
  Entry &ref = *e;
  if (s.verify == SPOPS_DROPIN_VERIFY_OFF) s.transient = std::move(e);
  else s.entries[key] = std::move(e);
  return ref;
}

// Device matrix for a host matrix, created by `make(spops_matrix*)` when missing or stale.
template <typename Fill, typename Make>
spops_matrix cached(const void *key, const Source &src, Fill &&fill_digest, Make &&make) {
  return cached_entry(key, src, fill_digest, [&](Entry &e) { check(make(&e.m), "create device matrix"); }).m;
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

// Upload a host index array to device buffer `d`, widened to int64.
template <typename I>
const int64_t *upload_index64(DeviceBuffer &d, const I *a, int64_t n) {
  std::vector<int64_t> w(a, a + n);
  void *p = d.get((size_t)(n ? n : 1) * 8);
  check(spops_memcpy_htod(p, w.data(), (size_t)n * 8, nullptr), "upload");
  return static_cast<const int64_t *>(p);
}

// Upload to a scratch device buffer (slot 0..3), widening indices to int64.
template <typename I>
const int64_t *upload_index64(int slot, const I *a, int64_t n) {
  return upload_index64(state().scratch[slot], a, n);
}

// CSR host arrays as int64 device arrays plus device values (the B operand of LeSpGEMM and
// the MinHash input), cached like a matrix.  Caller holds state().mu.
template <typename V>
struct DeviceCsr {
  const int64_t *rowptr;
  const int64_t *col;
  const V *values;
};

template <typename I, typename V>
DeviceCsr<V> cached_csr64(const I *row_offset, const I *col_index, const V *values, int64_t n_rows, int64_t n_cols,
                             int64_t nnz) {
  Source src;
  src.kind = 100 + (int)sizeof(I) * 10 + (int)sizeof(V);
  src.dims[0] = n_rows; src.dims[1] = n_cols; src.dims[2] = nnz;
  src.arrays[0] = row_offset; src.arrays[1] = col_index; src.arrays[2] = values;
  Entry &e = cached_entry(
      row_offset, src,
      [&](Digest &d, int mode) {
        d.array(row_offset, sizeof(I), n_rows + 1, mode);
        d.array(col_index, sizeof(I), nnz, mode);
        if (values) d.array(values, sizeof(V), nnz, mode);
      },
      [&](Entry &en) {
        upload_index64(en.buf[0], row_offset, n_rows + 1);
        upload_index64(en.buf[1], col_index, nnz);
        void *dv = en.buf[2].get((size_t)(nnz ? nnz : 1) * sizeof(V));
        if (values) check(spops_memcpy_htod(dv, values, (size_t)nnz * sizeof(V), nullptr), "upload values");
      });
  return {static_cast<const int64_t *>(e.buf[0].p), static_cast<const int64_t *>(e.buf[1].p),
          static_cast<const V *>(e.buf[2].p)};
}

}  // namespace spops_dropin

/* Digest mode of the cache (SPOPS_DROPIN_VERIFY_SAMPLED / _FULL / _OFF); returns the previous
 * one.  Changing it drops every cached copy. */
inline int spops_dropin_set_verify(int mode) {
  auto &s = spops_dropin::state();
  std::lock_guard<std::mutex> g(s.mu);
  const int old = s.verify;
  if (mode >= SPOPS_DROPIN_VERIFY_SAMPLED && mode <= SPOPS_DROPIN_VERIFY_OFF) {
    s.verify = mode;
    s.entries.clear();
    s.transient.reset();
  }
  return old;
}

/* Drop the device copy cached for a host matrix (key: its first index array). */
inline void spops_dropin_release(const void *key) {
  auto &s = spops_dropin::state();
  std::lock_guard<std::mutex> g(s.mu);
  s.entries.erase(key);
}

inline void spops_dropin_release_all() {
  auto &s = spops_dropin::state();
  std::lock_guard<std::mutex> g(s.mu);
  s.entries.clear();
  s.transient.reset();
}

#endif
