/* SpGEMM of the LeSpGEMM path (drop-in for the reference's include/spgemm.h):
 *   get_spgemm_flop   src/spgemm.cpp:24-42
 *   LeSpGEMM_VLength  src/spgemm.cpp:662-680 (template <bool sortOutput, I, V>)
 *   LeSpGEMM          src/spgemm.cpp:684-705, row-wise, and its kernels LeSpGEMM_hash_rowwise
 *                     :45-104, LeSpGEMM_array_rowwise_new :110-169, LeSpGEMM_spa_rowwise :175-236
 *                     (spops_spgemm_csr: one C for every flag, columns ascending, zero sums
 *                     kept; the reference's hash kernel leaves hash order when sortOutput = 0)
 *   HashSpGEMMTopK    src/spgemm.cpp:707-842 (spops_spgemm_topk_jaccard)
 * C = A_cluster * B on the GPU (spops_lespgemm, the hybrid LDS-hash / HBM-bitmap
 * accumulator).  kernel_flag 1, 2 and 3 all produce the same C here (they differ only in
 * the reference's CPU accumulator); for kernel 3 the caller-allocated acc_flag, min_ccol and
 * max_ccol arrays are filled as the reference does (src/spgemm.cpp:598-622, the
 * classification using the reference's L2 budget so dense_cluster_count is comparable) and
 * the phase timings land in mixed_acc_*_ms.  C's columns ascend inside each cluster (the
 * reference's sortOutput=1 order; with sortOutput=0 it leaves hash-slot order). */
#ifndef SPOPS_DROPIN_SPGEMM_H
#define SPOPS_DROPIN_SPGEMM_H

#include <stdexcept>
#include <vector>

#include "sparse_format.h"
#include "spgemm_cluster.h"
#include "spgemm_MinHashLSH.h"
#include "spgemm_utility.h"
#include "spops_dropin.h"

#ifndef MIXED_DENSITY_THRESHOLD
#define MIXED_DENSITY_THRESHOLD 0.15
#endif
#ifndef MIXED_L2_FRACTION
#define MIXED_L2_FRACTION 0.75
#endif
#include "plat_config.h"   /* CPU_L2CACHE_SIZE: the kernel-3 budget */

template <typename IndexType, typename ValueType>
long long int get_spgemm_flop(const CSR_Matrix<IndexType, ValueType> &A, const CSR_Matrix<IndexType, ValueType> &B) {
  long long int flops = 0;
  for (IndexType i = 0; i < A.num_rows; i++)
    for (IndexType j = A.row_offset[i]; j < A.row_offset[i + 1]; j++) {
      const IndexType a = A.col_index[j];
      if (a < B.num_rows) flops += B.row_offset[a + 1] - B.row_offset[a];
    }
  return flops * 2;
}

template <bool sortOutput, typename IndexType, typename ValueType>
void LeSpGEMM_VLength(const CSR_VlengthCluster<IndexType, ValueType> &A, const CSR_Matrix<IndexType, ValueType> &B,
                      CSR_VlengthCluster<IndexType, ValueType> &C, int kernel_flag = 1,
                      double mixed_l2_fraction = -1.0) {
  using namespace spops_dropin;
  typedef IndexType I;
  typedef ValueType V;
  if (A.cols != B.num_rows) throw std::invalid_argument("LeSpGEMM_VLength: A.cols != B.num_rows");
  std::lock_guard<std::mutex> g(state().mu);
  const int64_t nc = A.rows;
  std::vector<int64_t> csz(A.cluster_sz, A.cluster_sz + nc), rp(A.rowptr, A.rowptr + nc + 1);
  // A stays on the device between calls (cached and checked like the SpMV matrices)
  Source src;
  src.kind = 200 + (int)sizeof(I) * 10 + (int)sizeof(V);
  src.dims[0] = A.csr_rows; src.dims[1] = A.cols; src.dims[2] = nc; src.dims[3] = A.nnzc; src.dims[4] = A.nnzv;
  src.arrays[0] = A.cluster_sz; src.arrays[1] = A.rowptr; src.arrays[2] = A.rowptr_val; src.arrays[3] = A.colids;
  src.arrays[4] = A.values;
  spops_cluster hA = cached_entry(
                         A.rowptr, src,
                         [&](Digest &d, int mode) {
                           d.array(A.cluster_sz, sizeof(I), nc, mode);
                           d.array(A.rowptr, sizeof(I), nc + 1, mode);
                           d.array(A.rowptr_val, sizeof(I), nc + 1, mode);
                           d.array(A.colids, sizeof(I), A.nnzc, mode);
                           d.array(A.values, sizeof(V), A.nnzv, mode);
                         },
                         [&](Entry &e) {
                           std::vector<int64_t> rpv(A.rowptr_val, A.rowptr_val + nc + 1),
                               col(A.colids, A.colids + A.nnzc);
                           check(spops_cluster_create_from_host(&e.cl, A.csr_rows, A.cols, nc, csz.data(), rp.data(),
                                                                rpv.data(), col.data(), A.values, value_code<V>()),
                                 "spops_cluster_create_from_host");
                         })
                         .cl;
  spops_cluster hC = nullptr;
  int64_t stats[16] = {0};
  int st;
  {
    // B stays on the device between calls (cached and checked like the SpMV matrices)
    DeviceCsr<V> dB = cached_csr64<I, V>(B.row_offset, B.col_index, B.values, B.num_rows, B.num_cols, B.num_nnzs);
    st = spops_lespgemm(hA, B.num_rows, B.num_cols, B.num_nnzs, dB.rowptr, dB.col, dB.values, &hC, stats, nullptr);
  }
  check(st, "spops_lespgemm");
  int64_t info[8];
  spops_cluster_info(hC, info);
  C.csr_rows = A.csr_rows;
  C.rows = A.rows;
  C.cols = B.num_cols;
  C.nnzc = (I)info[3];
  C.nnzv = (I)info[4];
  C.max_cluster_sz = A.max_cluster_sz;
  std::vector<int64_t> ccsz(nc), crp(nc + 1), crpv(nc + 1), ccol(C.nnzc);
  C.values = new_array<V>(C.nnzv);
  st = spops_cluster_download(hC, ccsz.data(), crp.data(), crpv.data(), ccol.data(), C.values);
  std::vector<int64_t> nz(nc), mn(nc), mx(nc);
  if (st == SPOPS_SUCCESS) st = spops_cluster_symbolic_download(hC, nz.data(), mn.data(), mx.data());
  spops_cluster_destroy(hC);
  check(st, "download C");
  C.cluster_sz = new_array<I>(nc);
  C.rowptr = new_array<I>(nc + 1);
  C.rowptr_val = new_array<I>(nc + 1);
  C.colids = new_array<I>(C.nnzc);
  std::copy(ccsz.begin(), ccsz.end(), C.cluster_sz);
  std::copy(crp.begin(), crp.end(), C.rowptr);
  std::copy(crpv.begin(), crpv.end(), C.rowptr_val);
  std::copy(ccol.begin(), ccol.end(), C.colids);
  C.num_rows = C.csr_rows;
  C.num_cols = C.cols;
  C.mixed_acc_setup_ms = C.mixed_acc_malloc_ms = C.mixed_acc_symbolic_ms = 0.0;
  C.mixed_acc_dispatcher_ms = C.mixed_acc_numeric_ms = 0.0;
  C.dense_cluster_count = 0;
  if (kernel_flag == 3) {
    C.mixed_acc_setup_ms = stats[5] / 1e3;
    C.mixed_acc_symbolic_ms = stats[6] / 1e3;
    C.mixed_acc_malloc_ms = stats[7] / 1e3;
    C.mixed_acc_numeric_ms = stats[8] / 1e3;
    // classify_clusters (src/spgemm_Vlength_mixed.cpp:268-319): dense iff the B rows plus
    // the dense block fit the L2 budget and nz / col_range >= 0.15
    const double l2f = (mixed_l2_fraction >= 0.0 && mixed_l2_fraction <= 2.0) ? mixed_l2_fraction : MIXED_L2_FRACTION;
    const size_t budget = (size_t)((double)CPU_L2CACHE_SIZE * l2f);
    const double b_rowdense = B.num_rows ? (double)B.num_nnzs / (double)B.num_rows : 0.0;
    for (int64_t c = 0; c < nc; ++c) {
      if (C.min_ccol) C.min_ccol[c] = (I)mn[c];
      if (C.max_ccol) C.max_ccol[c] = (I)mx[c];
      char dense = 0;
      if (nz[c] > 0) {
        const int64_t range = mx[c] - mn[c] + 1;
        const double density = (double)nz[c] / (double)range;
        const size_t bbytes = (size_t)((double)(rp[c + 1] - rp[c]) * b_rowdense * (double)(sizeof(I) + sizeof(V)));
        const size_t dbytes = (size_t)range * (size_t)csz[c] * sizeof(V);
        dense = (bbytes + dbytes <= budget && density >= MIXED_DENSITY_THRESHOLD) ? 1 : 0;
      }
      if (C.acc_flag) C.acc_flag[c] = dense;
      C.dense_cluster_count += dense;
    }
  }
  (void)sortOutput;
}

namespace spops_dropin {
// Host CSR from three device arrays the library allocated (then freed)
template <typename I, typename V>
void take_device_csr(CSR_Matrix<I, V> &C, int64_t rows, int64_t cols, int64_t nnz, int64_t *drp, int64_t *dcol,
                     void *dval) {
  std::vector<int64_t> rp((size_t)rows + 1), col((size_t)nnz);
  std::vector<V> val((size_t)nnz);
  int st = spops_memcpy_dtoh(rp.data(), drp, rp.size() * 8, nullptr);
  if (st == SPOPS_SUCCESS && nnz) st = spops_memcpy_dtoh(col.data(), dcol, col.size() * 8, nullptr);
  if (st == SPOPS_SUCCESS && nnz) st = spops_memcpy_dtoh(val.data(), dval, val.size() * sizeof(V), nullptr);
  spops_free(drp);
  spops_free(dcol);
  spops_free(dval);
  check(st, "download C");
  C.num_rows = (I)rows;
  C.num_cols = (I)cols;
  C.num_nnzs = (I)nnz;
  C.tag = 0;
  C.partition = nullptr;
  C.row_offset = new_array<I>((size_t)rows + 1);
  C.col_index = new_array<I>(nnz ? (size_t)nnz : 1);
  C.values = new_array<V>(nnz ? (size_t)nnz : 1);
  std::copy(rp.begin(), rp.end(), C.row_offset);
  std::copy(col.begin(), col.end(), C.col_index);
  std::copy(val.begin(), val.end(), C.values);
}

template <typename I, typename V>
void rowwise_product(const CSR_Matrix<I, V> &A, const CSR_Matrix<I, V> &B, CSR_Matrix<I, V> &C) {
  if (A.num_cols != B.num_rows) throw std::invalid_argument("LeSpGEMM: A.num_cols != B.num_rows");
  std::lock_guard<std::mutex> g(state().mu);
  const DeviceCsr<V> dA = cached_csr64(A.row_offset, A.col_index, A.values, A.num_rows, A.num_cols, A.num_nnzs);
  const DeviceCsr<V> dB = cached_csr64(B.row_offset, B.col_index, B.values, B.num_rows, B.num_cols, B.num_nnzs);
  int64_t nnz = 0, *drp = nullptr, *dcol = nullptr;
  void *dval = nullptr;
  check(spops_spgemm_csr(A.num_rows, A.num_cols, A.num_nnzs, dA.rowptr, dA.col, dA.values, B.num_cols, B.num_nnzs,
                         dB.rowptr, dB.col, dB.values, value_code<V>(), &nnz, &drp, &dcol, &dval, nullptr, nullptr),
        "spops_spgemm_csr");
  take_device_csr(C, A.num_rows, B.num_cols, nnz, drp, dcol, dval);
}
}  // namespace spops_dropin

template <bool sortOutput = true, typename IndexType, typename ValueType>
void LeSpGEMM_hash_rowwise(const CSR_Matrix<IndexType, ValueType> &A, const CSR_Matrix<IndexType, ValueType> &B,
                           CSR_Matrix<IndexType, ValueType> &C, int kernel_flag) {
  spops_dropin::rowwise_product(A, B, C);
  C.kernel_flag = kernel_flag;
}

template <bool sortOutput = true, typename IndexType, typename ValueType>
void LeSpGEMM_array_rowwise_new(const CSR_Matrix<IndexType, ValueType> &A, const CSR_Matrix<IndexType, ValueType> &B,
                                CSR_Matrix<IndexType, ValueType> &C) {
  spops_dropin::rowwise_product(A, B, C);
  C.kernel_flag = 2;
}

template <bool sortOutput = true, typename IndexType, typename ValueType>
void LeSpGEMM_spa_rowwise(const CSR_Matrix<IndexType, ValueType> &A, const CSR_Matrix<IndexType, ValueType> &B,
                          CSR_Matrix<IndexType, ValueType> &C) {
  spops_dropin::rowwise_product(A, B, C);
  C.kernel_flag = 3;
}

template <bool sortOutput = true, typename IndexType, typename ValueType>
void LeSpGEMM(const CSR_Matrix<IndexType, ValueType> &A, const CSR_Matrix<IndexType, ValueType> &B,
              CSR_Matrix<IndexType, ValueType> &C, int kernel_flag = 1) {
  if (kernel_flag == 2) LeSpGEMM_array_rowwise_new<sortOutput>(A, B, C);
  else if (kernel_flag == 3) LeSpGEMM_spa_rowwise<sortOutput>(A, B, C);
  else LeSpGEMM_hash_rowwise<sortOutput>(A, B, C, kernel_flag);
}

template <typename IndexType, typename ValueType>
void HashSpGEMMTopK(const CSR_Matrix<IndexType, ValueType> &A, CSR_Matrix<IndexType, ValueType> &C,
                    IndexType top_k = 7) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  const DeviceCsr<ValueType> dA = cached_csr64(A.row_offset, A.col_index, A.values, A.num_rows, A.num_cols, A.num_nnzs);
  int64_t nnz = 0, *drp = nullptr, *dcol = nullptr;
  void *dval = nullptr;
  check(spops_spgemm_topk_jaccard(A.num_rows, A.num_cols, A.num_nnzs, dA.rowptr, dA.col, value_code<ValueType>(),
                                  (int64_t)top_k, 0, &nnz, &drp, &dcol, &dval, nullptr),
        "spops_spgemm_topk_jaccard");
  take_device_csr(C, A.num_rows, A.num_rows, nnz, drp, dcol, dval);
  C.kernel_flag = 0;
}

#endif
