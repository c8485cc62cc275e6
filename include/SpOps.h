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
 * use and cached (spops_dropin.h).  Link with -lsparseops_hip.
 *
 * Like the reference's umbrella header (include/SpOps.h:4-41) it also brings in the
 * configuration, I/O, conversion, benchmark, test-routine, feature and SpGEMM headers, so a
 * caller that includes only this file finds every name the reference's umbrella provides for
 * these formats. */
#ifndef SPOPS_DROPIN_SPOPS_H
#define SPOPS_DROPIN_SPOPS_H

#include <vector>

#include "general_config.h"
#include "plat_config.h"
#include "memopt.h"
#include "mmio.h"
#include "thread.h"
#include "sparse_io.h"
#include "sparse_format.h"
#include "sparse_operation.h"
#include "sparse_partition.h"
#include "sparse_conversion.h"
#include "spmv_benchmark.h"
#include "spops_dropin.h"

namespace spops_dropin {
// Cache description of a SELL-C-sigma / SELL-C-R host matrix: its counts, its reorder and
// chunk_len arrays, the per-chunk pointer tables and (digest) the chunks they point to.
template <typename I, typename V, typename M>
Source sell_source(const M &s, int kind) {
  Source src;
  src.kind = kind + (int)sizeof(I) * 10 + (int)sizeof(V) * 100;
  src.dims[0] = s.num_rows; src.dims[1] = s.num_cols; src.dims[2] = s.num_nnzs;
  src.dims[3] = s.chunkWidth_C; src.dims[4] = s.validchunkNum;
  src.arrays[0] = s.reorder; src.arrays[1] = s.chunk_len; src.arrays[2] = s.col_index; src.arrays[3] = s.values;
  return src;
}

template <typename I, typename V, typename M>
auto sell_digest(const M &s) {
  return [&s](Digest &d, int mode) {
    const int64_t nch = s.validchunkNum;
    d.array(s.reorder, sizeof(I), s.num_rows, mode);
    d.array(s.chunk_len, sizeof(I), nch, mode);
    d.array(s.col_index, sizeof(void *), nch, mode);
    d.array(s.values, sizeof(void *), nch, mode);
    if (nch <= 0 || s.chunk_len == nullptr) return;
    // the chunks themselves: every chunk (full) or 65 chunks at an even stride (sampled)
    const int64_t step = (mode == SPOPS_DROPIN_VERIFY_FULL || nch <= 65) ? 1 : 0;
    for (int64_t t = 0; t <= (step ? nch - 1 : 64); ++t) {
      const int64_t c = step ? t : (nch - 1) * t / 64;
      const int64_t len = (int64_t)s.chunk_len[c] * s.chunkWidth_C;
      d.array(s.col_index[c], sizeof(I), len, mode);
      d.array(s.values[c], sizeof(V), len, mode);
    }
  };
}

// Cache description of a CSR host matrix: counts, array addresses and (digest) contents.
template <typename I, typename V>
Source csr_source(const CSR_Matrix<I, V> &csr) {
  Source src;
  src.kind = 1 + (int)sizeof(I) * 10 + (int)sizeof(V) * 100;
  src.dims[0] = csr.num_rows; src.dims[1] = csr.num_cols; src.dims[2] = csr.num_nnzs;
  src.arrays[0] = csr.row_offset; src.arrays[1] = csr.col_index; src.arrays[2] = csr.values;
  return src;
}

template <typename I, typename V>
auto csr_digest(const CSR_Matrix<I, V> &csr) {
  return [&csr](Digest &d, int mode) {
    d.array(csr.row_offset, sizeof(I), (int64_t)csr.num_rows + 1, mode);
    d.array(csr.col_index, sizeof(I), csr.num_nnzs, mode);
    d.array(csr.values, sizeof(V), csr.num_nnzs, mode);
  };
}
}  // namespace spops_dropin

template <typename I, typename V>
void SpOps_csr(const V alpha, const CSR_Matrix<I, V> &csr, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  spops_matrix A = cached(csr.row_offset, csr_source(csr), csr_digest(csr), [&](spops_matrix *h) {
    return spops_csr_create(h, csr.num_rows, csr.num_cols, csr.num_nnzs, index_code<I>(), csr.row_offset, csr.col_index,
                            csr.values, value_code<V>(), value_code<V>());
  });
  spmv_host<V>(A, csr.num_rows, csr.num_cols, alpha, x, beta, y);
}

template <typename I, typename V>
void SpOps_ell(const V alpha, const ELL_Matrix<I, V> &ell, const V *x, const V beta, V *y) {
  using namespace spops_dropin;
  std::lock_guard<std::mutex> g(state().mu);
  Source src;
  src.kind = 2 + (int)sizeof(I) * 10 + (int)sizeof(V) * 100;
  src.dims[0] = ell.num_rows; src.dims[1] = ell.num_cols; src.dims[2] = ell.num_nnzs;
  src.dims[3] = ell.max_row_width; src.dims[4] = (int64_t)ell.ld;
  src.arrays[0] = ell.col_index; src.arrays[1] = ell.values;
  const int64_t slots = (int64_t)ell.num_rows * (int64_t)ell.max_row_width;
  auto digest = [&](Digest &d, int mode) {
    d.array(ell.col_index, sizeof(I), slots, mode);
    d.array(ell.values, sizeof(V), slots, mode);
  };
  spops_matrix A = cached(ell.col_index, src, digest, [&](spops_matrix *h) {
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
  spops_matrix A = cached(s.reorder, sell_source<I, V>(s, 3), sell_digest<I, V>(s), [&](spops_matrix *h) {
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
  spops_matrix A = cached(s.reorder, sell_source<I, V>(s, 3), sell_digest<I, V>(s), [&](spops_matrix *h) {
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
  Source src;
  src.kind = 5 + (int)sizeof(I) * 10 + (int)sizeof(V) * 100;
  src.dims[0] = b.num_rows; src.dims[1] = b.num_cols; src.dims[2] = b.blockDim_r; src.dims[3] = b.blockDim_c;
  src.dims[4] = b.mb; src.dims[5] = b.nnzb;
  src.arrays[0] = b.row_ptr; src.arrays[1] = b.block_colindex; src.arrays[2] = b.block_data;
  auto digest = [&](Digest &d, int mode) {
    d.array(b.row_ptr, sizeof(I), (int64_t)b.mb + 1, mode);
    d.array(b.block_colindex, sizeof(I), b.nnzb, mode);
    d.array(b.block_data, sizeof(V), (int64_t)b.nnzb * b.blockDim_r * b.blockDim_c, mode);
  };
  spops_matrix A = cached(b.row_ptr, src, digest, [&](spops_matrix *h) {
    return spops_bsr_create_from_blocks(h, b.num_rows, b.num_cols, b.blockDim_r, b.blockDim_c, b.mb, b.nnzb,
                                        index_code<I>(), b.row_ptr, b.block_colindex, b.block_data, value_code<V>(),
                                        value_code<V>());
  });
  spmv_host<V>(A, b.num_rows, b.num_cols, alpha, x, beta, y);
}

#include "spgemm.h"
#include "spmv_testroutine.h"
#include "sparse_features.h"

#endif
