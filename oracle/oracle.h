/* CPU restatement of the reference hot path (lsl036/SparseOps): TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker.  The product path (sparseops_amd, libsparseops_hip.so)
 * never calls it.
 *
 * All indices are int64, values float64 unless a suffix says otherwise.  Arrays returned
 * through T** out-parameters are malloc'ed here and released with orc_free().
 *
 * Parity status of each function (how it is pinned) is stated next to its definition in
 * oracle.c.  Every function is pinned; none is "parity unpinned".
 */
#ifndef SPARSEOPS_ORACLE_H
#define SPARSEOPS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_free(void *p);

/* ---- SpMV ---- */
void orc_spmv_csr(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                  double alpha, const double *x, double beta, double *y);
void orc_spmv_csr_f32(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const float *val,
                      float alpha, const float *x, float beta, float *y);
void orc_spmv_csr_ld(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                     double alpha, const double *x, double beta, double *y, double *absrow);

int64_t orc_sell_layout(int64_t n_rows, const int64_t *rowptr, int64_t sigma, int64_t C, int64_t align,
                        int64_t *reorder, int64_t *chunk_len, int64_t *chunk_off);
void orc_sell_fill(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                   int64_t C, const int64_t *reorder, const int64_t *chunk_len, const int64_t *chunk_off,
                   int64_t *scol, double *sval);
void orc_spmv_sell(int64_t n_rows, int64_t C, const int64_t *reorder, const int64_t *chunk_len,
                   const int64_t *chunk_off, const int64_t *scol, const double *sval,
                   double alpha, const double *x, double beta, double *y);

int64_t orc_ell_width(int64_t n_rows, const int64_t *rowptr);
void orc_ell_fill(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                  int64_t width, int col_major, int64_t *ecol, double *evals);
void orc_spmv_ell(int64_t n_rows, int64_t width, int col_major, const int64_t *ecol, const double *evals,
                  double alpha, const double *x, double beta, double *y);

int64_t orc_bsr_count(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col,
                      int64_t bR, int64_t bC, int64_t *brow_ptr);
void orc_bsr_fill(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col, const double *val,
                  int64_t bR, int64_t bC, const int64_t *brow_ptr, int64_t *bcol, double *bdata);
void orc_spmv_bsr(int64_t n_rows, int64_t n_cols, int64_t bR, int64_t bC, const int64_t *brow_ptr,
                  const int64_t *bcol, const double *bdata, double alpha, const double *x, double beta, double *y);

/* ---- LeSpGEMM preprocessing ---- */
void orc_minhash(int64_t n_rows, const int64_t *rowptr, const int64_t *col, int k, uint64_t seed,
                 uint64_t *sigs);
uint64_t orc_band_hash(const uint64_t *band_sig, int r);
int64_t orc_lsh_pairs(const uint64_t *sigs, int64_t n_rows, int k, int bands, int constrain,
                      int64_t **pi, int64_t **pj, double **score);
int64_t orc_hc(int version, int64_t n_rows, const int64_t *rowptr, const int64_t *col,
               int64_t npairs, const int64_t *pi, const int64_t *pj, const double *score,
               int score_f32, int64_t cluster_size, int64_t *perm, int64_t **offsets);
double orc_jaccard(const int64_t *rowptr, const int64_t *col, int64_t a, int64_t b, int sorted_unique);

/* ---- row-wise SpGEMM and the A * A^T top-k Jaccard front end ---- */
int64_t orc_spgemm_rows(int64_t n_rows, const int64_t *arp, const int64_t *acol, const double *aval, int64_t b_cols,
                        const int64_t *brp, const int64_t *bcol, const double *bval, int64_t *c_rowptr,
                        int64_t **c_col, double **c_val);
int64_t orc_cluster_matching(int64_t n, int64_t np, const int64_t *pi, const int64_t *pj, const double *score,
                             int64_t cs, int64_t max_rounds, int64_t *perm, int64_t **offsets, int64_t *rounds);
int64_t orc_topk_jaccard(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col, int64_t top_k,
                         int score_f32, int64_t *c_rowptr, int64_t **c_col, double **c_val);

/* ---- cluster format + cluster-wise SpGEMM ---- */
int64_t orc_cluster_build(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                          const int64_t *perm, int64_t nclusters, const int64_t *offsets,
                          int64_t *cl_rowptr, int64_t *cl_rowptr_val, int64_t *cl_sz,
                          int64_t **colids, double **values);
int64_t orc_spgemm_flop(int64_t a_rows, const int64_t *arowptr, const int64_t *acol,
                        int64_t b_rows, const int64_t *browptr);
void orc_cluster_symbolic(int64_t nclusters, const int64_t *cl_rowptr, const int64_t *colids,
                          const int64_t *browptr, const int64_t *bcol, int64_t b_cols,
                          int64_t *c_nz, int64_t *c_min, int64_t *c_max);
int64_t orc_cluster_classify(int64_t nclusters, const int64_t *cl_rowptr, const int64_t *cl_sz,
                             const int64_t *c_nz, const int64_t *c_min, const int64_t *c_max,
                             double b_rowdense, int64_t l2_budget, double density_threshold,
                             int64_t value_bytes, int8_t *acc_flag);
void orc_cluster_numeric(int64_t nclusters, const int64_t *cl_rowptr, const int64_t *cl_rowptr_val,
                         const int64_t *cl_sz, const int64_t *colids, const double *values,
                         const int64_t *browptr, const int64_t *bcol, const double *bval,
                         const int64_t *c_rowptr, const int64_t *c_rowptr_val,
                         int64_t *c_colids, double *c_values, int use_long_double);
int64_t orc_cluster_to_csr(int64_t n_rows, int64_t nclusters, const int64_t *cl_sz, const int64_t *perm,
                           const int64_t *c_rowptr, const int64_t *c_rowptr_val,
                           const int64_t *c_colids, const double *c_values,
                           int64_t *rowptr_out, int64_t **col_out, double **val_out);

/* ---- AI forms ---- */
void orc_spmm_csr(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                  int64_t k, const double *X, double *Y);
void orc_sddmm_csr(int64_t n_rows, const int64_t *rowptr, const int64_t *col, int64_t k,
                   const double *G, const double *X, double *dvals);
void orc_csr_transpose(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col,
                       const double *val, int64_t *t_rowptr, int64_t *t_col, double *t_val, int64_t *t_perm);

/* ---- matrix features (MTX<I,V>, src/sparse_features.cpp) ---- */
#define ORC_NUM_FEATURES 61
void orc_features(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col, const double *val,
                  double *out);

#ifdef __cplusplus
}
#endif
#endif
