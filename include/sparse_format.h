/* Sparse matrix structs of the SpMV path, field for field the layout of the reference's
 * include/sparse_format.h (Matrix_Features :49-68, COO :77, CSR :96-104, BSR :113-131,
 * ELL :192-206, SELL_C_Sigma :245-272, SELL_C_R :275-298) so existing callers compile and
 * link unchanged; the delete_* helpers free what the conversions allocate
 * (:300-449).  CSR5, DIA and S-ELL are not part of this library. */
#ifndef SPOPS_DROPIN_SPARSE_FORMAT_H
#define SPOPS_DROPIN_SPARSE_FORMAT_H

#include "memopt.h"

typedef enum { RowMajor = 0, ColMajor = 1 } LeadingDimension;

template <typename IndexType>
struct Matrix_Features {
  typedef IndexType index_type;
  IndexType num_rows;
  IndexType num_cols;
  IndexType num_nnzs;
  double sparsity;
  IndexType *partition = nullptr;  // the reference's OpenMP row split; unused on the GPU
  double time, gflops, gbytes;
  int tag;
  int kernel_flag;                 // 0/1/2 select serial/OMP/balanced on the CPU; all run on the GPU here
};

template <typename IndexType, typename ValueType>
struct COO_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType *row_index;
  IndexType *col_index;
  ValueType *values;
};

template <typename IndexType, typename ValueType>
struct CSR_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType *row_offset;
  IndexType *col_index;
  ValueType *values;
};

template <typename IndexType, typename ValueType>
struct BSR_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType blockDim_r;
  IndexType blockDim_c;
  IndexType blockNNZ;
  IndexType mb;
  IndexType nb;
  IndexType nnzb;
  IndexType *row_ptr;
  IndexType *block_colindex;
  ValueType *block_data;  // row-major blocks
};

template <typename IndexType, typename ValueType>
struct ELL_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  LeadingDimension ld;
  IndexType max_row_width;
  IndexType min_row_width;
  IndexType *col_index;  // num_rows * max_row_width, -1 padding
  ValueType *values;
};

template <typename IndexType, typename ValueType>
struct SELL_C_Sigma_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType sliceWidth_Sigma;
  IndexType chunkWidth_C;
  IndexType sliceNum;
  IndexType chunkNum;
  IndexType validchunkNum;
  IndexType chunk_num_per_slice;
  IndexType alignment;
  IndexType *reorder;     // reorder[i] = original row at sorted position i
  IndexType *chunk_len;   // per chunk, max row length rounded up to alignment
  IndexType **col_index;  // per chunk, row-major (chunk_len * C), -1 padding
  ValueType **values;
};

template <typename IndexType, typename ValueType>
struct SELL_C_R_Matrix : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType chunkWidth_C;
  IndexType validchunkNum;
  IndexType alignment;
  IndexType *reorder;
  IndexType *chunk_len;
  IndexType **col_index;
  ValueType **values;
};

template <typename I, typename V>
void delete_coo_matrix(COO_Matrix<I, V> &m) {
  delete_array(m.row_index);
  delete_array(m.col_index);
  delete_array(m.values);
}
template <typename I, typename V>
void delete_csr_matrix(CSR_Matrix<I, V> &m) {
  delete_array(m.row_offset);
  delete_array(m.col_index);
  delete_array(m.values);
}
template <typename I, typename V>
void delete_bsr_matrix(BSR_Matrix<I, V> &m) {
  m.blockDim_c = 0;
  m.blockDim_r = 0;
  delete_array(m.row_ptr);
  delete_array(m.block_colindex);
  delete_array(m.block_data);
}
template <typename I, typename V>
void delete_ell_matrix(ELL_Matrix<I, V> &m) {
  m.max_row_width = 0;
  m.min_row_width = 0;
  delete_array(m.col_index);
  delete_array(m.values);
}
template <typename I, typename V>
void delete_s_ell_c_sigma_matrix(SELL_C_Sigma_Matrix<I, V> &m) {
  m.alignment = 0;
  m.chunk_num_per_slice = 0;
  delete_array(m.reorder);
  delete_array(m.chunk_len);
  for (I c = 0; c < m.validchunkNum; c++) {
    delete_array(m.col_index[c]);
    delete_array(m.values[c]);
  }
  delete[] m.col_index;
  delete[] m.values;
  m.sliceNum = 0;
  m.chunkNum = 0;
  m.validchunkNum = 0;
}
template <typename I, typename V>
void delete_s_ell_c_R_matrix(SELL_C_R_Matrix<I, V> &m) {
  m.alignment = 0;
  delete_array(m.reorder);
  delete_array(m.chunk_len);
  for (I c = 0; c < m.validchunkNum; c++) {
    delete_array(m.col_index[c]);
    delete_array(m.values[c]);
  }
  delete[] m.col_index;
  delete[] m.values;
  m.validchunkNum = 0;
}

template <typename I, typename V> void delete_host_matrix(COO_Matrix<I, V> &m) { delete_coo_matrix(m); }
template <typename I, typename V> void delete_host_matrix(CSR_Matrix<I, V> &m) { delete_csr_matrix(m); }
template <typename I, typename V> void delete_host_matrix(BSR_Matrix<I, V> &m) { delete_bsr_matrix(m); }
template <typename I, typename V> void delete_host_matrix(ELL_Matrix<I, V> &m) { delete_ell_matrix(m); }
template <typename I, typename V> void delete_host_matrix(SELL_C_Sigma_Matrix<I, V> &m) { delete_s_ell_c_sigma_matrix(m); }
template <typename I, typename V> void delete_host_matrix(SELL_C_R_Matrix<I, V> &m) { delete_s_ell_c_R_matrix(m); }

#endif
