/* SpMV test routines with the reference's names and signatures (include/spmv_testroutine.h:
 * 7-53; utils/test_spmv_csr.cpp:13-96, test_spmv_ell.cpp:25-102, test_spmv_sell_c_sigma.cpp:
 * 14-92, test_spmv_sell_c_R.cpp:14-86, test_spmv_bsr.cpp:15-90) for the formats of this
 * library.  Each builds its format from csr_ref with the reference's parameters (SELL:
 * sigma = SELL_SIGMA, C = CHUNK_SIZE, alignment SIMD_WIDTH / 8 / sizeof(V); BSR: BSR_BlockDimRow
 * x alignment blocks; ELL: the given leading dimension), sets kernel_flag = kernel_tag (the GPU
 * kernel runs for every tag), checks SpOps_<format> against a host CSR product of csr_ref
 * (test_spmv_kernel's printout), benchmarks it (benchmark_spmv_on_host) and returns ms per call.
 * convert_time receives the conversion's ms.  The device copy the drop-in cached for the test
 * matrix is released before the routine frees it.  The reference's routines also truncate
 * MAT_FEATURES (a file their conversions no longer write); these do not touch it.  COO, DIA,
 * S-ELL and CSR5 are not part of this library. */
#ifndef SPOPS_DROPIN_SPMV_TESTROUTINE_H
#define SPOPS_DROPIN_SPMV_TESTROUTINE_H

#include <cstdio>
#include <iostream>

#include "SpOps.h"
#include "general_config.h"
#include "plat_config.h"
#include "sparse_conversion.h"
#include "sparse_format.h"
#include "sparse_operation.h"
#include "sparse_partition.h"
#include "spmv_benchmark.h"
#include "thread.h"
#include "timer.h"

namespace spops_dropin {
// y = alpha A x + beta y on the host with long double sums: the checker the routines compare
// the GPU result with (beta == 0 does not read y)
template <typename I, typename V>
void host_csr_spmv(const V alpha, const CSR_Matrix<I, V> &A, const V *x, const V beta, V *y) {
  for (I r = 0; r < A.num_rows; r++) {
    long double s = 0;
    for (I j = A.row_offset[r]; j < A.row_offset[r + 1]; j++) s += (long double)A.values[j] * x[A.col_index[j]];
    y[r] = (V)((long double)alpha * s + (beta == 0 ? 0.0L : (long double)beta * y[r]));
  }
}

inline const char *tag_suffix(int kernel_tag) {
  return kernel_tag == 0 ? "serial_simple" : kernel_tag == 1 ? "omp_simple" : "omp_ld";
}

template <typename I, typename V, typename M, typename F>
double check_and_time(const CSR_Matrix<I, V> &csr_ref, M &m, F spmv, const std::string &format, int kernel_tag) {
  const std::string name = format + "_" + tag_suffix(kernel_tag);
  std::cout << "\n===  Compared " << format << " (" << name << ") with a host CSR product  ===" << std::endl;
  test_spmv_kernel(csr_ref, host_csr_spmv<I, V>, m, spmv, name.c_str());
  std::cout << "\n===  Performance of " << name << "  ===" << std::endl;
  return benchmark_spmv_on_host(m, spmv, name);
}
}  // namespace spops_dropin

template <typename IndexType, typename ValueType>
double test_csr_matrix_kernels(const CSR_Matrix<IndexType, ValueType> &csr_ref, int kernel_tag, int schedule_mod) {
  std::cout << "=====  Testing CSR Kernels  =====" << std::endl;
  CSR_Matrix<IndexType, ValueType> csr_test = csr_ref;
  csr_test.row_offset = copy_array(csr_ref.row_offset, (size_t)csr_ref.num_rows + 1);
  csr_test.col_index = copy_array(csr_ref.col_index, (size_t)csr_ref.num_nnzs);
  csr_test.values = copy_array(csr_ref.values, (size_t)csr_ref.num_nnzs);
  csr_test.partition = nullptr;
  csr_test.kernel_flag = kernel_tag;
  if (kernel_tag == 1) {
    const IndexType t = (IndexType)Le_get_thread_num();
    set_omp_schedule(schedule_mod, (int)std::max<IndexType>(OMP_ROWS_SIZE, csr_test.num_rows / std::max<IndexType>(t, 1)));
  } else if (kernel_tag == 2) {
    const IndexType t = (IndexType)Le_get_thread_num();
    csr_test.partition = new_array<IndexType>((size_t)t + 1);
    balanced_partition_row_by_nnz(csr_test.row_offset, csr_test.num_rows, t, csr_test.partition);
  }
  const double ms = spops_dropin::check_and_time(csr_ref, csr_test, SpOps_csr<IndexType, ValueType>, "csr", kernel_tag);
  spops_dropin_release(csr_test.row_offset);   // the device copy of the matrix
  delete_csr_matrix(csr_test);
  return ms;
}

template <typename IndexType, typename ValueType>
double test_ell_matrix_kernels(const CSR_Matrix<IndexType, ValueType> &csr_ref, int kernel_tag, LeadingDimension ld,
                               int schedule_mod, double &convert_time) {
  (void)schedule_mod;
  std::cout << "=====  Testing ELL Kernels  =====" << std::endl;
  timer t;
  ELL_Matrix<IndexType, ValueType> ell = csr_to_ell(csr_ref, ld);
  convert_time = t.milliseconds_elapsed();
  ell.kernel_flag = kernel_tag;
  const double ms = spops_dropin::check_and_time(csr_ref, ell, SpOps_ell<IndexType, ValueType>, "ell", kernel_tag);
  spops_dropin_release(ell.col_index);
  delete_host_matrix(ell);
  return ms;
}

template <typename IndexType, typename ValueType>
double test_sell_c_sigma_matrix_kernels(const CSR_Matrix<IndexType, ValueType> &csr_ref, int kernel_tag,
                                        int schedule_mod, double &convert_time) {
  (void)schedule_mod;
  std::cout << "=====  Testing SELL-c-sigma Kernels  =====" << std::endl;
  const IndexType alignment = (IndexType)(SIMD_WIDTH / 8 / sizeof(ValueType));
  timer t;
  SELL_C_Sigma_Matrix<IndexType, ValueType> s = csr_to_sell_c_sigma(csr_ref, nullptr, SELL_SIGMA, CHUNK_SIZE, alignment);
  convert_time = t.milliseconds_elapsed();
  s.kernel_flag = kernel_tag;
  const double ms =
      spops_dropin::check_and_time(csr_ref, s, SpOps_sell_c_sigma<IndexType, ValueType>, "sell_c_sigma", kernel_tag);
  spops_dropin_release(s.reorder);
  delete_host_matrix(s);
  return ms;
}

template <typename IndexType, typename ValueType>
double test_sell_c_R_matrix_kernels(const CSR_Matrix<IndexType, ValueType> &csr_ref, int kernel_tag, int schedule_mod) {
  (void)schedule_mod;
  std::cout << "=====  Testing SELL-c-R Kernels  =====" << std::endl;
  const IndexType alignment = (IndexType)(SIMD_WIDTH / 8 / sizeof(ValueType));
  SELL_C_R_Matrix<IndexType, ValueType> s = csr_to_sell_c_R(csr_ref, nullptr, CHUNK_SIZE, alignment);
  s.kernel_flag = kernel_tag;
  const double ms = spops_dropin::check_and_time(csr_ref, s, SpOps_sell_c_R<IndexType, ValueType>, "sell_c_R", kernel_tag);
  spops_dropin_release(s.reorder);
  delete_host_matrix(s);
  return ms;
}

template <typename IndexType, typename ValueType>
double test_bsr_matrix_kernels(const CSR_Matrix<IndexType, ValueType> &csr_ref, int kernel_tag, int schedule_mod,
                               double &convert_time) {
  (void)schedule_mod;
  std::cout << "=====  Testing BSR Kernels  =====" << std::endl;
  const IndexType alignment = (IndexType)(SIMD_WIDTH / 8 / sizeof(ValueType));
  timer t;
  BSR_Matrix<IndexType, ValueType> b = csr_to_bsr(csr_ref, (IndexType)BSR_BlockDimRow, alignment);
  convert_time = t.milliseconds_elapsed();
  b.kernel_flag = kernel_tag;
  const double ms = spops_dropin::check_and_time(csr_ref, b, SpOps_bsr<IndexType, ValueType>, "bsr", kernel_tag);
  spops_dropin_release(b.row_ptr);
  delete_host_matrix(b);
  return ms;
}

#endif
