/* Row clustering of the LeSpGEMM path (drop-in for the reference's
 * include/spgemm_utility.h): hierarchical_clustering_v0_fast :743-915,
 * hierarchical_clustering_v1 :609-733 and the map-based hierarchical_clustering_v0 :517-606
 * with reordered_dict_to_permutation_and_offsets :921-939, with the same signatures and
 * bit-identical permutation and offsets (host code in libsparseops_hip.so,
 * spops_hierarchical_clustering / spops_hierarchical_clustering_v0), plus the Jaccard helpers
 * :362-429 and CumulativeSum :334-343.  hierarchical_clustering_v0 does not add the pairs it
 * scores on the way to the caller's map (the reference does; its callers do not read them). */
#ifndef SPOPS_DROPIN_SPGEMM_UTILITY_H
#define SPOPS_DROPIN_SPGEMM_UTILITY_H

#include <map>
#include <set>
#include <type_traits>
#include <utility>
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

template <typename IndexType, typename ValueType>
inline std::map<IndexType, std::vector<IndexType>> hierarchical_clustering_v0(
    const IndexType *rowptr, const IndexType *colids, IndexType num_rows,
    std::map<std::pair<IndexType, IndexType>, ValueType> &close_pairs, IndexType cluster_size) {
  using namespace spops_dropin;
  const int64_t n = num_rows, nnz = num_rows > 0 ? (int64_t)rowptr[num_rows] : 0;
  std::vector<int64_t> rp(rowptr, rowptr + n + 1), ci(colids, colids + nnz);
  std::vector<int64_t> pi, pj;
  std::vector<ValueType> sc;
  pi.reserve(close_pairs.size());
  pj.reserve(close_pairs.size());
  sc.reserve(close_pairs.size());
  for (const auto &kv : close_pairs) {
    pi.push_back((int64_t)kv.first.first);
    pj.push_back((int64_t)kv.first.second);
    sc.push_back(kv.second);
  }
  std::vector<int64_t> perm((size_t)n), offsets((size_t)n + 1), roots((size_t)(n > 0 ? n : 1));
  int64_t nc = 0;
  check(spops_hierarchical_clustering_v0(n, rp.data(), ci.data(), (int64_t)pi.size(), pi.data(), pj.data(), sc.data(),
                                         value_code<ValueType>(), (int64_t)cluster_size, perm.data(), offsets.data(),
                                         roots.data(), &nc),
        "spops_hierarchical_clustering_v0");
  std::map<IndexType, std::vector<IndexType>> dict;
  for (int64_t c = 0; c < nc; ++c)
    dict[(IndexType)roots[(size_t)c]].assign(perm.begin() + offsets[(size_t)c], perm.begin() + offsets[(size_t)c + 1]);
  return dict;
}

// permutation[new_row] = original row, clusters in the map's (root) order; offsets [0, len0, ...]
template <typename IndexType>
inline void reordered_dict_to_permutation_and_offsets(const std::map<IndexType, std::vector<IndexType>> &reordered_dict,
                                                      IndexType num_rows, std::vector<IndexType> &permutation_out,
                                                      std::vector<IndexType> &offsets_out) {
  (void)num_rows;
  permutation_out.clear();
  offsets_out.assign(1, 0);
  for (const auto &kv : reordered_dict) {
    permutation_out.insert(permutation_out.end(), kv.second.begin(), kv.second.end());
    offsets_out.push_back((IndexType)permutation_out.size());
  }
}

#endif
