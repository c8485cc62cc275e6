/* Row partition helpers with the reference's names (include/sparse_partition.h:1-45,
 * utils/partition.cpp:29-80), kept for callers that compute their own host-side splits.  The
 * GPU kernels behind SpOps_* balance their work themselves and ignore any partition.
 *
 *   lower_bound_int(t, l, r, v)   first index m in [l, r] with t[m] >= v (r when none before it)
 *   balanced_partition_row_by_nnz(row_ptr, rows, T, part)
 *       part[0] = 0, part[T] = rows and, for 0 < i < T, part[i] = the first row r (at most
 *       rows - 1) whose inclusive entry count row_ptr[r + 1] reaches i * max(1, nnz / T). */
#ifndef SPOPS_DROPIN_SPARSE_PARTITION_H
#define SPOPS_DROPIN_SPARSE_PARTITION_H

#include <algorithm>

template <typename IndexType>
IndexType lower_bound_int(const IndexType *t, IndexType l, IndexType r, IndexType value) {
  return (IndexType)(std::lower_bound(t + l, t + r, value) - t);
}

template <typename IndexType>
void balanced_partition_row_by_nnz(const IndexType *row_ptr, IndexType rows, IndexType num_threads,
                                   IndexType *partition) {
  const IndexType *incl = row_ptr + 1;             // entries up to and including row r
  const IndexType nnz = rows > 0 ? incl[rows - 1] : 0;
  const IndexType per = std::max<IndexType>(1, nnz / std::max<IndexType>(num_threads, 1));
  partition[0] = 0;
  for (IndexType i = 1; i < num_threads; i++)
    partition[i] = rows > 0 ? lower_bound_int(incl, (IndexType)0, (IndexType)(rows - 1), (IndexType)(per * i)) : 0;
  partition[num_threads] = rows;
}

#endif
