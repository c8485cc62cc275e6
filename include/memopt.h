/* Host array helpers with the reference's names and semantics (include/memopt.h:38-75):
 * new_array<T>(N) returns 64-byte aligned memory (release with delete_array), copy_array
 * duplicates an array.  Used by the drop-in format structs below. */
#ifndef SPOPS_DROPIN_MEMOPT_H
#define SPOPS_DROPIN_MEMOPT_H

#include <cstdint>
#include <cstdlib>
#include <cstring>

#ifndef CACHE_LINE
#define CACHE_LINE 64
#endif

template <typename T>
T *new_array(const size_t N) {
  void *p = nullptr;
  const size_t bytes = (N ? N : 1) * sizeof(T);
  if (posix_memalign(&p, CACHE_LINE, bytes) != 0) return nullptr;
  return static_cast<T *>(p);
}

template <typename T>
void delete_array(T *p) {
  free(p);
}

template <typename T>
void memcpy_array(T *dst, const T *src, const size_t length) {
  std::memcpy(dst, src, sizeof(T) * length);
}

template <typename T>
T *copy_array(const T *src, const size_t N) {
  T *dst = new_array<T>(N);
  memcpy_array<T>(dst, src, N);
  return dst;
}

#endif
