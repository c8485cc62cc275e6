/* SpMV benchmark helpers with the reference's names, loop and printout
 * (include/spmv_benchmark.h:26-215):
 *
 *   bytes_per_spmv(A)            the reference's traffic model per format (below)
 *   benchmark_spmv(A, spmv, min_iter, max_iter, seconds, name)
 *       x uniform in [0, 1) from rand(), y = 0; one timed warm-up call; then max_iter calls
 *       when the warm-up took under 20 ms, else min(max_iter, max(min_iter, seconds / warm-up))
 *       calls, timed together; sets A.gflops (2 nnz / time), A.gbytes (bytes_per_spmv / time)
 *       and A.time (ms per call); prints the reference's "\tbenchmarking ..." line; returns ms
 *   benchmark_spmv_on_host(A, spmv, name)   the same with MIN_ITER, MAX_ITER, TIME_LIMIT
 *
 * With the drop-in SpOps_* the timed calls are whole host-array calls: x goes to the device,
 * the kernel runs, y comes back (the matrix stays cached on the device after the first call).
 *
 * The traffic model is the reference's, kept formula for formula so that GB/s figures are
 * comparable with the reference's printouts (it is not the algorithmic byte count this
 * library reports elsewhere):
 *   CSR   2 sizeof(I) rows + sizeof(I) nnz + 2 sizeof(V) nnz + 2 sizeof(V) rows
 *   COO   2 sizeof(I) nnz + 2 sizeof(V) nnz + 2 sizeof(V) per row holding an entry
 *   BSR   2 sizeof(I) mb + sizeof(I) nnzb + 2 sizeof(V) nnzb bR bC + 2 sizeof(V) rows
 *   ELL   (sizeof(I) + sizeof(V)) rows width + sizeof(V) nnz + 2 sizeof(V) rows
 *   SELL-C-sigma, SELL-C-R   sizeof(I) chunks + sizeof(I) rows + 2 sizeof(I) C sum(chunk_len)
 *                            + sizeof(V) nnz + 2 sizeof(V) rows
 *         (the chunk values are counted at sizeof(I), as the reference does) */
#ifndef SPOPS_DROPIN_SPMV_BENCHMARK_H
#define SPOPS_DROPIN_SPMV_BENCHMARK_H

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "general_config.h"
#include "memopt.h"
#include "sparse_format.h"
#include "timer.h"

template <typename IndexType, typename ValueType>
size_t bytes_per_spmv(const COO_Matrix<IndexType, ValueType> &mtx) {
  std::vector<char> used((size_t)mtx.num_rows, 0);
  for (size_t n = 0; n < (size_t)mtx.num_nnzs; n++) used[(size_t)mtx.row_index[n]] = 1;
  const size_t rows_used = (size_t)std::count(used.begin(), used.end(), 1);
  return 2 * sizeof(IndexType) * (size_t)mtx.num_nnzs + 2 * sizeof(ValueType) * (size_t)mtx.num_nnzs +
         2 * sizeof(ValueType) * rows_used;
}

template <typename IndexType, typename ValueType>
size_t bytes_per_spmv(const CSR_Matrix<IndexType, ValueType> &mtx) {
  const size_t r = (size_t)mtx.num_rows, z = (size_t)mtx.num_nnzs;
  return 2 * sizeof(IndexType) * r + sizeof(IndexType) * z + 2 * sizeof(ValueType) * z + 2 * sizeof(ValueType) * r;
}

template <typename IndexType, typename ValueType>
size_t bytes_per_spmv(const BSR_Matrix<IndexType, ValueType> &mtx) {
  const size_t blocks = (size_t)mtx.nnzb;
  return 2 * sizeof(IndexType) * (size_t)mtx.mb + sizeof(IndexType) * blocks +
         2 * sizeof(ValueType) * blocks * (size_t)mtx.blockDim_c * (size_t)mtx.blockDim_r +
         2 * sizeof(ValueType) * (size_t)mtx.num_rows;
}

template <typename IndexType, typename ValueType>
size_t bytes_per_spmv(const ELL_Matrix<IndexType, ValueType> &mtx) {
  const size_t slots = (size_t)mtx.num_rows * (size_t)mtx.max_row_width;
  return (sizeof(IndexType) + sizeof(ValueType)) * slots + sizeof(ValueType) * (size_t)mtx.num_nnzs +
         2 * sizeof(ValueType) * (size_t)mtx.num_rows;
}

namespace spops_dropin {
template <typename IndexType, typename ValueType, typename M>
size_t sell_model_bytes(const M &mtx) {
  size_t slots = 0;
  for (IndexType c = 0; c < mtx.validchunkNum; ++c) slots += (size_t)mtx.chunk_len[c] * (size_t)mtx.chunkWidth_C;
  return sizeof(IndexType) * (size_t)mtx.validchunkNum + sizeof(IndexType) * (size_t)mtx.num_rows +
         2 * sizeof(IndexType) * slots + sizeof(ValueType) * (size_t)mtx.num_nnzs +
         2 * sizeof(ValueType) * (size_t)mtx.num_rows;
}
}  // namespace spops_dropin

template <typename IndexType, typename ValueType>
size_t bytes_per_spmv(const SELL_C_Sigma_Matrix<IndexType, ValueType> &mtx) {
  return spops_dropin::sell_model_bytes<IndexType, ValueType>(mtx);
}

template <typename IndexType, typename ValueType>
size_t bytes_per_spmv(const SELL_C_R_Matrix<IndexType, ValueType> &mtx) {
  return spops_dropin::sell_model_bytes<IndexType, ValueType>(mtx);
}

template <typename SparseMatrix, typename SpMV>
double benchmark_spmv(SparseMatrix &sp_host, SpMV spmv, const int min_iterations, const int max_iterations,
                      const double seconds, const std::string method_name) {
  typedef typename SparseMatrix::value_type ValueType;
  typedef typename SparseMatrix::index_type IndexType;
  ValueType *x = new_array<ValueType>(sp_host.num_cols);
  ValueType *y = new_array<ValueType>(sp_host.num_rows);
  for (IndexType i = 0; i < sp_host.num_cols; i++) x[i] = (ValueType)(std::rand() / (RAND_MAX + 1.0));
  std::fill(y, y + sp_host.num_rows, ValueType(0));

  timer warm;
  spmv((ValueType)1, sp_host, x, (ValueType)0, y);
  const double est_ms = warm.milliseconds_elapsed();
  int iters = max_iterations;
  if (est_ms >= 20)
    iters = std::min(max_iterations, std::max(min_iterations, (int)(seconds * 1000 / est_ms)));
  std::printf("\tPerforming %d iterations\n", iters);

  timer t;
  for (int i = 0; i < iters; i++) spmv((ValueType)1, sp_host, x, (ValueType)0, y);
  const double ms = t.milliseconds_elapsed() / (double)iters;
  const double sec = ms / 1000.0;
  const double gflops = sec == 0 ? 0 : (2.0 * (double)sp_host.num_nnzs / sec) / 1e9;
  const double gbytes = sec == 0 ? 0 : ((double)bytes_per_spmv(sp_host) / sec) / 1e9;
  sp_host.gflops = gflops;
  sp_host.gbytes = gbytes;
  sp_host.time = ms;
  std::printf("\tbenchmarking %-20s [%s]: %8.4f ms ( %5.4f GFLOP/s %5.4f GB/s)\n", method_name.c_str(), "gpu", ms,
              gflops, gbytes);
  delete_array(x);
  delete_array(y);
  return ms;
}

template <typename SparseMatrix, typename SpMV>
double benchmark_spmv_on_host(SparseMatrix &sp_host, SpMV spmv, std::string method_name) {
  return benchmark_spmv<SparseMatrix, SpMV>(sp_host, spmv, MIN_ITER, MAX_ITER, TIME_LIMIT, method_name);
}

#endif
