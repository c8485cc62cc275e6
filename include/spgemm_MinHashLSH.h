/* MinHash + LSH candidate pairs (drop-in for the vector API of the reference's
 * include/spgemm_MinHashLSH.h): CandidatePair :225-230, minhash_signatures_flat :119-180,
 * minhash_estimated_jaccard :183-204, band_hash :277-283, lsh_candidate_pairs_vector
 * :777-791.  Signatures and pairs are computed on the GPU (spops_minhash_signatures,
 * spops_lsh_candidate_pairs) and are bit-identical to the reference's. */
#ifndef SPOPS_DROPIN_MINHASH_LSH_H
#define SPOPS_DROPIN_MINHASH_LSH_H

#include <cstdint>
#include <type_traits>
#include <vector>

#include "spops_dropin.h"

template <typename IndexType, typename ValueType>
struct CandidatePair {
  IndexType i;
  IndexType j;
  ValueType score;
};

namespace lsh_internal {
inline uint64_t band_hash(const uint64_t *band_sig, int r) {
  uint64_t h = 0;
  for (int i = 0; i < r; ++i) h = h * 0x9e3779b97f4a7c15ULL + band_sig[i];
  return h;
}
}  // namespace lsh_internal

inline double minhash_estimated_jaccard(const uint64_t *sig_a, const uint64_t *sig_b, int k) {
  if (k <= 0) return 0.0;
  int match = 0;
  for (int d = 0; d < k; ++d) match += sig_a[d] == sig_b[d];
  return static_cast<double>(match) / static_cast<double>(k);
}

template <typename IndexType>
std::vector<uint64_t> minhash_signatures_flat(const IndexType *rowptr, const IndexType *col_index, IndexType num_rows,
                                              int k, uint64_t seed = 12345) {
  using namespace spops_dropin;
  std::vector<uint64_t> sigs((size_t)num_rows * (size_t)(k > 0 ? k : 0), ~0ull);
  if (k <= 0 || num_rows <= 0) return sigs;
  std::lock_guard<std::mutex> g(state().mu);
  const int64_t nnz = (int64_t)rowptr[num_rows];
  const int64_t *drp = upload_index64(0, rowptr, (int64_t)num_rows + 1);
  const int64_t *dcol = upload_index64(1, col_index, nnz);
  void *ds = state().scratch[2].get(sigs.size() * 8);
  check(spops_minhash_signatures(drp, dcol, SPOPS_INDEX_I64, num_rows, k, seed, static_cast<uint64_t *>(ds), nullptr),
        "spops_minhash_signatures");
  check(spops_memcpy_dtoh(sigs.data(), ds, sigs.size() * 8, nullptr), "download signatures");
  return sigs;
}

template <typename IndexType, typename ValueType>
std::vector<CandidatePair<IndexType, ValueType>> lsh_candidate_pairs_vector(const IndexType *rowptr,
                                                                             const IndexType *col_index,
                                                                             IndexType num_rows, int k, int num_bands,
                                                                             uint64_t seed = 12345,
                                                                             bool constrain_large_buckets = true) {
  using namespace spops_dropin;
  std::vector<CandidatePair<IndexType, ValueType>> out;
  if (num_rows <= 0 || k <= 0 || num_bands <= 0) return out;
  std::lock_guard<std::mutex> g(state().mu);
  const int64_t nnz = (int64_t)rowptr[num_rows];
  const int64_t *drp = upload_index64(0, rowptr, (int64_t)num_rows + 1);
  const int64_t *dcol = upload_index64(1, col_index, nnz);
  spops_pairs P = nullptr;
  check(spops_lsh_candidate_pairs(drp, dcol, SPOPS_INDEX_I64, num_rows, k, num_bands, seed, constrain_large_buckets ? 1 : 0,
                                  value_code<ValueType>(), &P, nullptr),
        "spops_lsh_candidate_pairs");
  int64_t n = 0;
  spops_pairs_count(P, &n);
  std::vector<int64_t> pi(n), pj(n);
  std::vector<ValueType> sc(n);
  int st = spops_pairs_download(P, pi.data(), pj.data(), sc.data());
  spops_pairs_destroy(P);
  check(st, "spops_pairs_download");
  out.resize(n);
  for (int64_t t = 0; t < n; ++t) out[t] = {(IndexType)pi[t], (IndexType)pj[t], sc[t]};
  return out;
}

#endif
