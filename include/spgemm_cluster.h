/* Variable-length cluster matrix of the LeSpGEMM path, field for field the layout of the
 * reference's CSR_VlengthCluster (include/spgemm_cluster.h:108-143) and its deleter
 * (:152-169).  Cluster i holds cluster_sz[i] consecutive rows of the reordered matrix, its
 * distinct columns colids[rowptr[i] .. rowptr[i+1]) and, per column t, a dense block of
 * cluster_sz[i] values at values[rowptr_val[i] + (t - rowptr[i]) * cluster_sz[i] + row]. */
#ifndef SPOPS_DROPIN_SPGEMM_CLUSTER_H
#define SPOPS_DROPIN_SPGEMM_CLUSTER_H

#include "sparse_format.h"

template <typename IndexType, typename ValueType>
struct CSR_VlengthCluster : public Matrix_Features<IndexType> {
  typedef IndexType index_type;
  typedef ValueType value_type;
  IndexType csr_rows;
  IndexType rows;            // number of clusters
  IndexType cols;
  IndexType nnzc;            // distinct columns over all clusters
  IndexType nnzv;            // values stored (padded to cluster_sz per column)
  IndexType max_cluster_sz;
  IndexType *cluster_sz;
  IndexType *rowptr;
  IndexType *colids;
  IndexType *rowptr_val;
  ValueType *values;
  int dense_cluster_count;
  char *acc_flag = NULL;         // kernel 3: hash (0) / dense (1) classification per cluster
  IndexType *min_ccol = NULL;    // kernel 3: min / max C column per cluster
  IndexType *max_ccol = NULL;
  double mixed_acc_setup_ms = 0.0;
  double mixed_acc_malloc_ms = 0.0;
  double mixed_acc_symbolic_ms = 0.0;
  double mixed_acc_dispatcher_ms = 0.0;
  double mixed_acc_numeric_ms = 0.0;
};

template <typename IndexType, typename ValueType>
void delete_vlength_cluster_matrix(CSR_VlengthCluster<IndexType, ValueType> &c) {
  delete_array(c.cluster_sz);
  delete_array(c.rowptr);
  delete_array(c.colids);
  delete_array(c.rowptr_val);
  delete_array(c.values);
  c.cluster_sz = nullptr;
  c.rowptr = nullptr;
  c.colids = nullptr;
  c.rowptr_val = nullptr;
  c.values = nullptr;
  c.csr_rows = 0;
  c.rows = 0;
  c.cols = 0;
  c.nnzc = 0;
  c.nnzv = 0;
  c.max_cluster_sz = 0;
}

#endif
