/* Drop-in replacement of the reference's SpOps.h for the SpMV path on MI355X.
 *
 *   template <I, V> void SpOps_csr(alpha, const CSR_Matrix<I,V>&, const V* x, beta, V* y)
 *   ... SpOps_ell, SpOps_sell_c_sigma, SpOps_sell_c_R, SpOps_bsr
 *
 * Same names, argument meaning and template parameters as src/spmv_csr.cpp:165,
 * src/spmv_ell.cpp:225, src/spmv_sell_c_sigma.cpp:190, src/spmv_sell_c_R.cpp:190 and
 * src/spmv_bsr.cpp:215 (I in {int, long long, int64_t}, V in {float, double}):
 * y = alpha*A*x + beta*y with host x and y, computed by the gfx950 kernels of
 * libsparseops_hip.so.  kernel_flag (serial / OMP / balanced) is accepted and ignored: every
 * flag runs the GPU kernel.  beta == 0 does not read y.  The matrix is uploaded on first
 * use and cached (spops_dropin.h).  Link with -lsparseops_hip. */
#ifndef SPOPS_DROPIN_SPOPS_H
#define SPOPS_DROPIN_SPOPS_H

#include <vector>

#include "memopt.h"
#include "sparse_conversion.h"
#include "sparse_format.h"
#include "spops_dropin.h"

template <typename I, typename V>
void SpOps_csr(const V alpha, const CSR_Matrix<I, V> &csr, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  spops_matrix A = cached(csr.row_offset, [&](spops_matrix *h) {
    return spops_csr_create(h, csr.num_rows, csr.num_cols, csr.num_nnzs, index_code<I>(), csr.row_offset, csr.col_index,
                            csr.values, value_code<V>(), value_code<V>());
  });
  spmv_host<V>(A, csr.num_rows, csr.num_cols, alpha, x, beta, y);
}

template <typename I, typename V>
void SpOps_ell(const V alpha, const ELL_Matrix<I, V> &ell, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  spops_matrix A = cached(ell.col_index, [&](spops_matrix *h) {
    // entries of a row are the slots before its first -1 (csr_to_ell fills from slot 0)
    std::vector<I> rp(ell.num_rows + 1, 0), ci;
    std::vector<V> va;
    for (I r = 0; r < ell.num_rows; r++) {
      for (I t = 0; t < ell.max_row_width; t++) {
        const size_t pos = ell.ld == ColMajor ? (size_t)r + (size_t)t * ell.num_rows : (size_t)r * ell.max_row_width + t;
        if (ell.col_index[pos] < 0) continue;
        ci.push_back(ell.col_index[pos]);
        va.push_back(ell.values[pos]);
      }
      rp[r + 1] = (I)ci.size();
    }
    return spops_ell_create(h, ell.num_rows, ell.num_cols, (int64_t)ci.size(), index_code<I>(), rp.data(), ci.data(),
                            va.data(), value_code<V>(), value_code<V>());
  });
  spmv_host<V>(A, ell.num_rows, ell.num_cols, alpha, x, beta, y);
}

template <typename I, typename V>
void SpOps_sell_c_sigma(const V alpha, const SELL_C_Sigma_Matrix<I, V> &s, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  spops_matrix A = cached(s.reorder, [&](spops_matrix *h) {
    return spops_sell_create_from_chunks(h, s.num_rows, s.num_cols, s.num_nnzs, s.chunkWidth_C, s.validchunkNum,
                                         index_code<I>(), s.reorder, s.chunk_len, (const void *const *)s.col_index,
                                         (const void *const *)s.values, value_code<V>(), value_code<V>());
  });
  spmv_host<V>(A, s.num_rows, s.num_cols, alpha, x, beta, y);
}

template <typename I, typename V>
void SpOps_sell_c_R(const V alpha, const SELL_C_R_Matrix<I, V> &s, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  spops_matrix A = cached(s.reorder, [&](spops_matrix *h) {
    return spops_sell_create_from_chunks(h, s.num_rows, s.num_cols, s.num_nnzs, s.chunkWidth_C, s.validchunkNum,
                                         index_code<I>(), s.reorder, s.chunk_len, (const void *const *)s.col_index,
                                         (const void *const *)s.values, value_code<V>(), value_code<V>());
  });
  spmv_host<V>(A, s.num_rows, s.num_cols, alpha, x, beta, y);
}

template <typename I, typename V>
void SpOps_bsr(const V alpha, const BSR_Matrix<I, V> &b, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  spops_matrix A = cached(b.row_ptr, [&](spops_matrix *h) {
    return spops_bsr_create_from_blocks(h, b.num_rows, b.num_cols, b.blockDim_r, b.blockDim_c, b.mb, b.nnzb,
                                        index_code<I>(), b.row_ptr, b.block_colindex, b.block_data, value_code<V>(),
                                        value_code<V>());
  });
  spmv_host<V>(A, b.num_rows, b.num_cols, alpha, x, beta, y);
}

#include "spgemm.h"

#endif
