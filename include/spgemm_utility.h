/* Row clustering of the LeSpGEMM path (drop-in for the reference's
 * include/spgemm_utility.h): hierarchical_clustering_v0_fast :743-915 and
 * hierarchical_clustering_v1 :609-733 with the same signatures, bit-identical permutation
 * and offsets (host code in libsparseops_hip.so, spops_hierarchical_clustering), plus the
 * Jaccard helpers :362-429 and CumulativeSum :334-343. */
#ifndef SPOPS_DROPIN_SPGEMM_UTILITY_H
#define SPOPS_DROPIN_SPGEMM_UTILITY_H

#include <set>
#include <type_traits>
#include <vector>

#include "spops_dropin.h"

template <typename T>
inline T CumulativeSum(T *arr, T size) {
  T total = 0;
  for (T i = 0; i < size; ++i) {
    T v = arr[i];
    arr[i] = total;
    total += v;
  }
  return total;
}

template <typename IndexType, typename ValueType>
inline double jaccard_similarity(const IndexType *rowptr, const IndexType *colids, IndexType row_a, IndexType row_b) {
  std::set<IndexType> sa(colids + rowptr[row_a], colids + rowptr[row_a + 1]);
  IndexType c = 0;
  for (IndexType j = rowptr[row_b]; j < rowptr[row_b + 1]; ++j) c += sa.count(colids[j]) ? 1 : 0;
  IndexType u = (rowptr[row_a + 1] - rowptr[row_a]) + (rowptr[row_b + 1] - rowptr[row_b]) - c;
  return (u == 0) ? 0.0 : (1.0 * c / u);
}

namespace spops_dropin {
template <typename IndexType, typename ValueType, typename PairLike>
void hierarchical_clustering(int version, const IndexType *rowptr, const IndexType *colids, IndexType num_rows,
                             const std::vector<PairLike> &pairs, IndexType cluster_size,
                             std::vector<IndexType> &permutation_out, std::vector<IndexType> &offsets_out) {
  const int64_t n = num_rows, nnz = num_rows > 0 ? (int64_t)rowptr[num_rows] : 0;
  std::vector<int64_t> rp(rowptr, rowptr + n + 1), ci(colids, colids + nnz);
  std::vector<int64_t> pi(pairs.size()), pj(pairs.size());
  std::vector<ValueType> sc(pairs.size());
  for (size_t t = 0; t < pairs.size(); ++t) {
    pi[t] = pairs[t].i;
    pj[t] = pairs[t].j;
    sc[t] = static_cast<ValueType>(pairs[t].score);
  }
  std::vector<int64_t> perm(n), offsets(n + 1);
  int64_t nc = 0;
  check(spops_hierarchical_clustering(version, n, rp.data(), ci.data(), (int64_t)pairs.size(), pi.data(), pj.data(),
                                      sc.data(), value_code<ValueType>(), cluster_size, perm.data(), offsets.data(), &nc),
        "spops_hierarchical_clustering");
  permutation_out.assign(perm.begin(), perm.end());
  offsets_out.assign(offsets.begin(), offsets.begin() + nc + 1);
}
}  // namespace spops_dropin

template <typename IndexType, typename ValueType, typename PairLike>
inline void hierarchical_clustering_v0_fast(const IndexType *rowptr, const IndexType *colids, IndexType num_rows,
                                            const std::vector<PairLike> &pairs, IndexType cluster_size,
                                            std::vector<IndexType> &permutation_out,
                                            std::vector<IndexType> &offsets_out) {
  spops_dropin::hierarchical_clustering<IndexType, ValueType>(0, rowptr, colids, num_rows, pairs, cluster_size,
                                                              permutation_out, offsets_out);
}

template <typename IndexType, typename ValueType, typename PairLike>
inline void hierarchical_clustering_v1(const IndexType *rowptr, const IndexType *colids, IndexType num_rows,
                                       const std::vector<PairLike> &pairs, IndexType cluster_size,
                                       std::vector<IndexType> &permutation_out, std::vector<IndexType> &offsets_out) {
  spops_dropin::hierarchical_clustering<IndexType, ValueType>(1, rowptr, colids, num_rows, pairs, cluster_size,
                                                              permutation_out, offsets_out);
}

#endif
