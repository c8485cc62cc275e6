/* SpMV comparison helpers with the reference's names and output (include/sparse_operation.h:
 * 75-196): the test routines of the callers check one SpMV against another with them.
 *
 *   maximum_relative_error(A, B, N)   max |a - b| / (|a| + |b| + sqrt(eps)) over the entries
 *                                     that differ (0 when all are equal)
 *   compare_spmv_kernels(m1, spmv1, m2, spmv2)
 *       y = 0.8 A x + 0.7 y with x uniform in [0, 1) from rand() and y = 0, once with each
 *       routine; prints " [max error %9f]" and the reference's warnings: " POSSIBLE small
 *       Round-Error" when 5 sqrt(eps) < error < 0.01, " POSSIBLE FAILURE" when error >= 0.005
 *   test_spmv_kernel(m1, spmv1, m2, spmv2, name)   the same, framed by the routine's name
 * sum_csr_duplicates lives in sparse_conversion.h (coo_to_csr's compact step). */
#ifndef SPOPS_DROPIN_SPARSE_OPERATION_H
#define SPOPS_DROPIN_SPARSE_OPERATION_H

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "memopt.h"
#include "sparse_conversion.h"
#include "sparse_format.h"

template <typename T>
T maximum_relative_error(const T *A, const T *B, const size_t N) {
  const T eps = std::sqrt(std::numeric_limits<T>::epsilon());
  T worst = 0;
  for (size_t i = 0; i < N; i++) {
    const T d = std::abs(A[i] - B[i]);
    if (d != 0) worst = std::max(worst, d / (std::abs(A[i]) + std::abs(B[i]) + eps));
  }
  return worst;
}

template <typename SparseMatrix1, typename SpMV1, typename SparseMatrix2, typename SpMV2>
void compare_spmv_kernels(const SparseMatrix1 &sm1_host, SpMV1 spmv1, const SparseMatrix2 &sm2_host, SpMV2 spmv2) {
  assert(sm1_host.num_rows == sm2_host.num_rows);
  assert(sm1_host.num_cols == sm2_host.num_cols);
  assert(sm1_host.num_nnzs == sm2_host.num_nnzs);
  typedef typename SparseMatrix1::index_type IndexType;
  typedef typename SparseMatrix2::value_type ValueType;
  const ValueType alpha = 0.8, beta = 0.7;
  const IndexType rows = sm1_host.num_rows, cols = sm1_host.num_cols;
  ValueType *x1 = new_array<ValueType>(cols), *y1 = new_array<ValueType>(rows);
  for (IndexType i = 0; i < cols; i++) x1[i] = (ValueType)(std::rand() / (RAND_MAX + 1.0));
  std::fill(y1, y1 + rows, ValueType(0));
  ValueType *x2 = copy_array(x1, cols), *y2 = copy_array(y1, rows);
  spmv1(alpha, sm1_host, x1, beta, y1);
  spmv2(alpha, sm2_host, x2, beta, y2);
  const ValueType err = maximum_relative_error(y1, y2, (size_t)rows);
  std::printf(" [max error %9f]", (double)err);
  if (err > 5 * std::sqrt(std::numeric_limits<ValueType>::epsilon()) && err < 0.01)
    std::printf(" POSSIBLE small Round-Error");
  else if (err >= 0.005)
    std::printf(" POSSIBLE FAILURE");
  delete_array(x1);
  delete_array(y1);
  delete_array(x2);
  delete_array(y2);
}

template <typename SparseMatrix1, typename SpMV1, typename SparseMatrix2, typename SpMV2>
void test_spmv_kernel(const SparseMatrix1 &sm1_host, SpMV1 spmv1, const SparseMatrix2 &sm2_host, SpMV2 spmv2,
                      const char *method_name) {
  std::printf("\ttesting %-26s", method_name);
  std::printf("[gpu]:");
  compare_spmv_kernels(sm1_host, spmv1, sm2_host, spmv2);
  std::printf("\n");
}

#endif
