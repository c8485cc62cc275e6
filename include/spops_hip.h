/*
 * spops_hip.h -- C ABI of libsparseops_hip.so, the MI355X (gfx950) implementation of the
 * SparseOps hot path: SpMV in CSR / ELL / SELL-C-sigma / SELL-C-R / BSR and the LeSpGEMM
 * pipeline (MinHash-LSH -> hierarchical clustering -> cluster-wise SpGEMM), plus batched
 * SpMM and the SpMM/SDDMM pair used by autograd.
 *
 * Plain C: pointers, sizes and int status codes; no HIP or torch types.  `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Device pointers are marked d_.
 * Every entry point returns SPOPS_SUCCESS (0) or an error code; spops_last_error()
 * gives the detail for the calling thread.
 *
 * Value storage types and the vector types that go with them:
 *   SPOPS_VALUE_F64  : values double, x/y double, FP64 accumulation
 *   SPOPS_VALUE_F32  : values float,  x/y float,  FP32 accumulation (compensated on long rows)
 *   SPOPS_VALUE_BF16 : values bf16,   x/y float,  FP32 accumulation (compensated on long rows)
 *
 * Reference interfaces replaced (lsl036/SparseOps):
 *   SpOps_csr          include/spmv_csr.h:20,        src/spmv_csr.cpp:165
 *   SpOps_ell          include/spmv_ell.h:53,        src/spmv_ell.cpp:225
 *   SpOps_sell_c_sigma include/spmv_sell_c_sigma.h:61, src/spmv_sell_c_sigma.cpp:190
 *   SpOps_sell_c_R     include/spmv_sell_c_R.h:61,   src/spmv_sell_c_R.cpp:190
 *   SpOps_bsr          include/spmv_bsr.h:59,        src/spmv_bsr.cpp:215
 *   (each entry below cites the one it backs)
 */
#ifndef SPOPS_HIP_H
#define SPOPS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define SPOPS_SUCCESS               0
#define SPOPS_ERROR_INVALID_VALUE   1   /* bad argument (shape, type code, null pointer)  */
#define SPOPS_ERROR_OUT_OF_MEMORY   2   /* device or host allocation failed               */
#define SPOPS_ERROR_HIP             3   /* a HIP runtime call failed                        */
#define SPOPS_ERROR_NOT_SUPPORTED   4   /* valid request the library does not implement    */
#define SPOPS_ERROR_LIMIT           5   /* exceeds a documented size limit                  */

#define SPOPS_INDEX_I32  0
#define SPOPS_INDEX_I64  1
#define SPOPS_VALUE_F64  0
#define SPOPS_VALUE_F32  1
#define SPOPS_VALUE_BF16 2

#define SPOPS_FORMAT_CSR   0
#define SPOPS_FORMAT_SELL  1   /* SELL-C-sigma (SELL-C-R is sigma = n_rows) */
#define SPOPS_FORMAT_ELL   2
#define SPOPS_FORMAT_BSR   3

const char *spops_status_string(int status);
const char *spops_last_error(void);
int spops_version(int *major, int *minor, int *patch);

/* ---- device memory and streams (so C++ callers need no HIP headers) --------------------- */
int spops_device_count(int *count);
int spops_set_device(int device);
int spops_malloc(void **d_ptr, size_t bytes);
int spops_free(void *d_ptr);
int spops_memcpy_htod(void *d_dst, const void *src, size_t bytes, void *stream);
int spops_memcpy_dtoh(void *dst, const void *d_src, size_t bytes, void *stream);
int spops_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream);
int spops_memset(void *d_ptr, int value, size_t bytes, void *stream);
int spops_stream_synchronize(void *stream);
int spops_stream_create(void **stream);
int spops_stream_destroy(void *stream);
/* events for kernel timing on a given stream */
int spops_event_create(void **event);
int spops_event_destroy(void *event);
int spops_event_record(void *event, void *stream);
int spops_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- SpMV matrices ------------------------------------------------------------------------
 * A spops_matrix is a device-resident, immutable matrix in one format with its launch plan
 * (row blocks, long-row lists, chunk descriptors).  Creation uploads and converts; it is
 * the place where the reference's csr_to_* conversions happen.  Every creator takes
 * HOST arrays; values are given as `in_value_type` (F64 or F32 host arrays) and stored as
 * `value_type` (F64, F32 or BF16, rounded to nearest even).  Row/column indices are `index_type` (int32 or int64 host arrays). */
typedef struct spops_matrix_s *spops_matrix;

/* CSR (include/sparse_format.h:96-104). Backs SpOps_csr. */
int spops_csr_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                     int index_type, const void *rowptr, const void *col,
                     const void *values, int in_value_type, int value_type);

/* The same CSR handle with its SpMV layout chosen explicitly instead of by the process-wide
 * SPOPS_TUNE_SPMV_LAYOUT: 0 chosen from the matrix (as spops_csr_create does by default),
 * 1 the row kernels only (no swept copy of the entries: for handles used by SpMM, SDDMM and
 * the transpose, which never run SpMV, so that spops_csr_set_values updates one copy),
 * 2 the swept layout beside the CSR arrays.  Backs SpOps_csr like spops_csr_create. */
int spops_csr_create_layout(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                            int index_type, const void *rowptr, const void *col,
                            const void *values, int in_value_type, int value_type, int spmv_layout);

/* SELL-C-sigma from CSR with the reference's ordering: each sigma-slice sorted by
 * (nnz desc, row desc), chunks of C sorted rows (include/sparse_conversion.h:345-487).
 * sigma = n_rows gives SELL-C-R (:491-606).  sigma % C != 0 -> SPOPS_ERROR_INVALID_VALUE.
 * Backs csr_to_sell_c_sigma + SpOps_sell_c_sigma and csr_to_sell_c_R + SpOps_sell_c_R. */
int spops_sell_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                      int index_type, const void *rowptr, const void *col,
                      const void *values, int in_value_type, int value_type, int64_t sigma, int64_t C);

/* SELL-C-sigma from the reference's own host layout (SELL_C_Sigma_Matrix,
 * include/sparse_format.h:245-272): reorder[n_rows], chunk_len[nchunks], and per chunk
 * arrays col[chunk][chunk_len*C], val[chunk][chunk_len*C] row-major inside the chunk with
 * col = -1 padding.  `in_value_type` (F64 or F32) is the host value type.  Indices int32
 * or int64 per index_type. */
int spops_sell_create_from_chunks(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                                  int64_t C, int64_t nchunks, int index_type,
                                  const void *reorder, const void *chunk_len,
                                  const void *const *chunk_col, const void *const *chunk_val,
                                  int in_value_type, int value_type);

/* ELL (include/sparse_format.h:192-206, csr_to_ell :195-256). Backs SpOps_ell. */
int spops_ell_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                     int index_type, const void *rowptr, const void *col,
                     const void *values, int in_value_type, int value_type);

/* BSR with bR x bC row-major blocks (include/sparse_format.h:113-131,
 * csr_to_bsr :777-934). Backs SpOps_bsr. */
int spops_bsr_create(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t nnz,
                     int index_type, const void *rowptr, const void *col,
                     const void *values, int in_value_type, int value_type, int64_t bR, int64_t bC);

/* BSR from the reference's own block layout (row_ptr[mb+1], block_colindex[nnzb], row-major
 * blocks block_data[nnzb*bR*bC]); backs SpOps_bsr on a BSR_Matrix built by the caller. */
int spops_bsr_create_from_blocks(spops_matrix *A, int64_t n_rows, int64_t n_cols, int64_t bR, int64_t bC,
                                 int64_t mb, int64_t nnzb, int index_type, const void *row_ptr,
                                 const void *block_colindex, const void *block_data, int in_value_type,
                                 int value_type);

/* y = alpha*A*x + beta*y on the device; beta == 0 does not read y (BLAS convention;
 * the reference's OMP kernels read y there, src/spmv_csr.cpp:119).  x/y element type
 * follows the value type (see top).  Asynchronous on `stream`. */
int spops_spmv(spops_matrix A, double alpha, const void *d_x, double beta, void *d_y, void *stream);

/* ---- matrix features ------------------------------------------------------------------------
 * The reference's feature extractor (MTX<I,V>, include/sparse_features.h:13-326;
 * src/sparse_features.cpp:170-480, 609-850) for a general real CSR matrix on the device:
 * out[SPOPS_NUM_FEATURES] in this order (the reference's names):
 *   0 num_rows  1 num_cols  2 real_nnz  3 nnz_ratio  4 num_of_diags  5 nnz_closed2diag_ratio
 *   6 nnz_lower  7 nnz_upper  8 nnz_diagonal
 *   9..16  rows: NoneZero_Row_Ratio, min, max, ave, var, standard_dev, P-ratio, Gini
 *   17..24 columns: the same
 *   25 diagonal_dominant_ratio  26 max_value_offdiag  27 max_value_diagonal
 *   28 row_variability  29 col_variability
 *   30 number of tiles per dimension (256)  31 rows per tile  32 columns per tile
 *   33..40 tiles: NE_ratio, min, max, ave, var, stand_dev, P-ratio, Gini
 *   41..48 row blocks (RB): the same   49..56 column blocks (CB): the same
 *   57 uniq_R  58 uniq_C  59 potReuseR  60 potReuseC
 * Synchronous on `stream` (the host finishes the statistics). */
#define SPOPS_NUM_FEATURES 61
int spops_matrix_features(const void *d_rowptr, const void *d_col, const void *d_values, int index_type,
                          int value_type, int64_t n_rows, int64_t n_cols, double *out, void *stream);

/* ---- batched SpMM ---------------------------------------------------------------------------
 * No reference counterpart as an entry point: many independent matrices, each with its own
 * number of right-hand sides k_i, multiplied in one persistent launch (Y_i = alpha A_i X_i +
 * beta Y_i; each product is SpOps_csr, src/spmv_csr.cpp:165-185, on k_i columns).  FP32
 * values, X and Y.  X_i (n_cols_i x ld_i) and Y_i (n_rows_i x ld_i) are row-major blocks of one
 * packed X and one packed Y buffer (16-byte aligned) at the offsets spops_batch_plan reports;
 * ld_i >= k_i is the row stride (columns past k_i are padding).  Per matrix an autotuner picks
 * the format (row chunks with {col, value} pairs, or CSR lane groups), the row ordering and
 * the lanes per row: tune = 0 uses the default rule, tune = 1 times the candidates per
 * feature class at creation. */
typedef struct spops_batch_s *spops_batch;
int spops_batch_create(spops_batch *out, int64_t nmat, const int64_t *n_rows, const int64_t *n_cols, const int64_t *k,
                       int index_type, const void *const *rowptr, const void *const *col, const void *const *values,
                       int in_value_type, int tune);
/* packed lengths (floats) of X and Y, 2*sum(nnz_i*k_i), algorithmic bytes, tasks */
int spops_batch_info(spops_batch B, int64_t *x_len, int64_t *y_len, int64_t *flop, int64_t *algo_bytes,
                     int64_t *ntasks);
/* per matrix: X / Y offsets (floats), row stride ld, format (0 CSR lane groups, 1 row
 * chunks), lanes per row, sigma (rows ordered by length inside windows of sigma rows; 1 = no
 * reordering), feature class; any pointer may be NULL */
int spops_batch_plan(spops_batch B, int64_t *x_off, int64_t *y_off, int32_t *ld, int32_t *fmt, int32_t *lanes,
                     int32_t *sigma, int32_t *cls);
int spops_batch_spmm(spops_batch B, float alpha, const float *d_X, float beta, float *d_Y, void *stream);
/* capture the launch on `stream` into a HIP graph kept in B; replays reuse X, Y, alpha, beta */
int spops_batch_capture(spops_batch B, float alpha, const float *d_X, float beta, float *d_Y, void *stream);
int spops_batch_replay(spops_batch B, void *stream);
int spops_batch_destroy(spops_batch B);

/* ---- multi-RHS and differentiable forms (CSR matrices) ------------------------------------
 * No reference counterpart as entry points: SpOps_csr (src/spmv_csr.cpp:165-185) applied to k
 * right-hand sides, its gradient with respect to the values (SDDMM on A's pattern) and the
 * transpose of csr_transpose (include/sparse_conversion.h:2067-2108) with the entry
 * permutation kept.  X row-major (n_cols x k), Y row-major (n_rows x k).  Accumulation is FP64
 * when the matrix or X is FP64, otherwise FP32.  Other formats return NOT_SUPPORTED. */
/* Y = alpha*A*X + beta*Y with X, Y of the matrix's x/y type (as for spops_spmv) */
int spops_spmm(spops_matrix A, int64_t k, double alpha, const void *d_X, double beta, void *d_Y, void *stream);
/* the same with explicit X and Y element types (SPOPS_VALUE_F64 / F32 / BF16) */
int spops_spmm_typed(spops_matrix A, int64_t k, double alpha, const void *d_X, int x_type, double beta, void *d_Y,
                     int y_type, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): spops_spmm_typed
 * with the same arguments and results, computed by the same kernels, whose only scratch (the
 * partial sums of rows longer than 256 entries, one row of k sums per 1024-entry segment) is
 * a static device array of 64 MiB instead of a per-handle buffer grown on demand.  It makes
 * no allocation.  A request needing more (segments x k x 8 bytes for FP64 values or X, x 4
 * otherwise) returns SPOPS_ERROR_LIMIT.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams, or host threads without ordering) race;
 * order them. */
int spops_spmm_typed_noalloc(spops_matrix A, int64_t k, double alpha, const void *d_X, int x_type, double beta,
                             void *d_Y, int y_type, void *stream);
/* No counterpart in the reference (for differentiable layers whose values change every
 * step): spops_spmm_typed with the matrix values read from d_values (nnz elements of A's
 * stored value type) instead of the stored ones, which are neither read nor changed.  As
 * with spops_csr_set_values, d_values of a handle made by spops_csr_transpose are in the
 * ORIGINAL matrix's entry order (the kernels read them through the kept permutation; such
 * handles of >= 2^31 entries return SPOPS_ERROR_LIMIT).  Saves the copy of
 * spops_csr_set_values before each product.  On a transposed handle the random value reads
 * land inside the gather-bound product: at R-MAT scale 20 (31 M entries, k = 64) that measured
 * slower than spops_csr_set_values followed by spops_spmm_typed. */
int spops_spmm_values(spops_matrix A, const void *d_values, int64_t k, double alpha, const void *d_X, int x_type,
                      double beta, void *d_Y, int y_type, void *stream);
/* out[e] = sum_c G[row_e, c] * X[col_e, c] for every entry e of A (A's values unused);
 * G row-major (n_rows x k), X row-major (n_cols x k), out[nnz] */
int spops_sddmm(spops_matrix A, int64_t k, const void *d_G, int g_type, const void *d_X, int x_type, void *d_out,
                int out_type, void *stream);
/* The backward of Y = A X in one pass over AT = A^T (a handle made by spops_csr_transpose, its
 * values set): dX = AT G (n_cols(A) x k, in X's type; alpha 1, beta 0) and dvals[e] =
 * sum_c G[row_e, c] * X[col_e, c] for every entry e of A, in A's entry order and AT's value
 * type.  Each row of G gathered for dX also serves the dot products, and X's rows are read once
 * per row of AT; the dot products are written in AT's entry order into a buffer the handle
 * keeps and moved into A's order by a two-pass permutation (coalesced passes through LDS instead
 * of a random gather).  The first call builds both permutations of the handle on the host (about
 * 16 bytes per entry of device memory, plus two nnz-value buffers; it allocates and copies
 * synchronously, so it cannot be captured into a graph, later calls can); afterwards
 * spops_csr_set_values on the handle uses the forward one too.  Calls on one AT handle must be
 * stream-ordered.  No reference counterpart (the differentiable form of SpOps_csr).  Returns
 * SPOPS_ERROR_NOT_SUPPORTED, having done nothing, unless k is a multiple of the elements of G in
 * 16 bytes and at most 64 times that, G is 16-byte aligned and X, dX are aligned to that many
 * elements of their type, and the handle's values take less than 2 GiB (the permutation's
 * staging); spops_spmm_typed on AT plus spops_sddmm on A compute the same. */
int spops_spmm_sddmm_transposed(spops_matrix AT, int64_t k, const void *d_G, int g_type, const void *d_X, int x_type,
                                void *d_dX, void *d_dvals, void *stream);
/* AT = A^T as a new CSR handle; AT's entry e is A's entry perm[e] (kept inside AT), rows of
 * AT hold A's entries in A's row order */
int spops_csr_transpose(spops_matrix A, spops_matrix *AT);
/* overwrite the stored values from a device array (cast to the stored type; BF16 storage
 * rounds each value once, to nearest even, from its FP64 / FP32 value).  For a handle made
 * by spops_csr_transpose the array is in the ORIGINAL matrix's entry order.  A handle with a
 * swept SpMV layout gets the new values there too, on the same stream, so every SpMV
 * ordered after this call (stream order, events, or a graph captured after it) sees them. */
int spops_csr_set_values(spops_matrix A, const void *d_values, int values_type, void *stream);

/* info[0]=format info[1]=n_rows info[2]=n_cols info[3]=nnz info[4]=value_type
 * info[5]=device bytes owned info[6]=algorithmic bytes of one SpMV (no padding)
 * info[7]=stored slots (incl. padding) info[8]=long rows info[9]=wave tasks (BSR: 1 when
 * the MFMA kernel runs it)
 * info[10]=C info[11]=sigma info[12]=bR (BSR) or column slices (SELL/ELL; 0 for the swept
 * layout, whose info[9] counts entry groups) info[13]=bC
 * (BSR) or max row width (ELL) or the entry groups of a CSR handle's swept SpMV layout (0
 * when its SpMV runs the row kernels) info[14]=nchunks (reference layout) info[15]=index_type */
int spops_matrix_info(spops_matrix A, int64_t info[16]);

/* Tuning (process-wide; no reference counterpart).  Returns the previous value in *old (may
 * be null).
 *   SPOPS_TUNE_SLICE_BYTES: bytes of x per column slice of the SELL/ELL layout of matrices
 *   created afterwards (default 2 MiB, one XCD's share of L2); the slice count is the
 *   smallest power of two <= 8 with n_cols * sizeof(x) / slices <= this.
 *   SPOPS_TUNE_DENSE_PRODUCTS: LeSpGEMM clusters with at least this many products (default
 *   2^22) accumulate in a dense block in device memory (many workgroups per cluster, atomic
 *   adds) instead of LDS chunks; applies to plans created afterwards.
 *   SPOPS_TUNE_SPMV_LAYOUT: device layout of SELL-C-sigma / SELL-C-R / ELL matrices created
 *   afterwards: 0 (default) chosen per matrix, 1 the column-sliced stream, 2 the swept layout
 *   (blocks of consecutive rows summed in FP64 in LDS, each block's entries sorted by column).
 *   The default picks the swept layout when one XCD's rows gather from more than the slice
 *   bytes of x and 16384 rows hold at least one entry per 128-byte line of x.  CSR matrices
 *   created afterwards follow the same choice for their SpMV: 1 keeps the row kernels, 2 and
 *   the default's swept choice add the swept layout beside the CSR arrays (which SpMM, SDDMM
 *   and the transpose keep using; spops_csr_set_values updates both copies).
 *   SPOPS_TUNE_SWEEP_WORKGROUPS: workgroups one swept-layout SpMV aims for (default 256, one
 *   per CU).  With B the fewest row blocks the row limit allows, a block is shared by
 *   H = this / B workgroups (1..16), and rows are cut into about this / H blocks of equal
 *   entry counts; the H partial sums of a row are added in a fixed order.
 *   SPOPS_TUNE_SWEEP_MAX_ROWS: rows of the largest row block of the swept layout, 1..20480
 *   (default 20480: 160 KiB of FP64 sums in LDS).  Both apply to matrices created afterwards.
 *   SPOPS_TUNE_SWEEP_SPLIT: how the swept layout forms its row blocks: 0 (default) consecutive
 *   rows of equal entry counts as above; T > 0 two classes, the rows of at most T entries and
 *   the longer ones, each class cut into blocks of the row limit in row order, and the blocks
 *   of long rows (many entries per 128-byte line of x) shared by as many workgroups as their
 *   entries ask for.  Applies to matrices created afterwards.
 *   SPOPS_TUNE_BSR_MFMA: which BSR matrices run on the matrix cores (v_mfma_f64_16x16x4f64,
 *   FP64 products and sums; dense blocks of at least 8 x 8 with bC % 4 == 0, bR <= 64,
 *   bR * bC <= 4096): 0 none, 1 (default) those whose blocks exceed the VALU kernel's 256
 *   entries, 2 all such blocks.  With one right-hand side the VALU kernel is the faster of
 *   the two where both apply.  Applies to matrices created afterwards.
 *   SPOPS_TUNE_TASK_PRODUCTS: a large LeSpGEMM cluster's numeric work is split over
 *   workgroups by ranges of its output-column chunks, about this many products per workgroup;
 *   0 (default) picks it per numeric call: 2^15, halved down to 2^13 while the call has fewer
 *   than 32768 such tasks.  Applies to numeric calls made afterwards.
 *   SPOPS_TUNE_WIDE_CURSORS: LeSpGEMM clusters of 1017..4096 A columns (B with < 2^31
 *   entries) in a variant that keeps one B-row cursor per column in LDS, at two workgroups
 *   per CU, instead of searching each B row once per chunk: 0 never, 1 always, 2 (default)
 *   when a numeric call has at least 8192 such tasks.
 *   Applies to numeric calls made afterwards.
 *   SPOPS_TUNE_HEAVY_PRODUCTS: LeSpGEMM clusters with at least this many products (default
 *   2^20) get their symbolic bitmap built by several workgroups (parts of about 2^18
 *   products) and kept by the plan for the emit pass, while the part bitmaps (1 GiB) and the
 *   kept ones (256 MiB) fit; applies to plans created afterwards. */
#define SPOPS_TUNE_SLICE_BYTES 0
#define SPOPS_TUNE_DENSE_PRODUCTS 1
#define SPOPS_TUNE_SPMV_LAYOUT 2
#define SPOPS_TUNE_SWEEP_WORKGROUPS 3
#define SPOPS_TUNE_SWEEP_MAX_ROWS 4
#define SPOPS_TUNE_BSR_MFMA 5
#define SPOPS_TUNE_TASK_PRODUCTS 6
#define SPOPS_TUNE_WIDE_CURSORS 7
#define SPOPS_TUNE_HEAVY_PRODUCTS 8
#define SPOPS_TUNE_SWEEP_SPLIT 9
int spops_set_tuning(int key, int64_t value, int64_t *old);

/* Device copies of the layout arrays a format conversion produced, for parity checks:
 * which=0: reorder (int32, n_rows; SELL only); which=1: chunk_len in the reference's
 * sense (int32, nchunks; SELL only, rounded to the reference alignment 256/8/sizeof(V));
 * which=2: the swept layout's parameters (int64[4]: row bits of a block-local row index,
 * workgroups per row block (the largest count), row blocks, entry groups; zeros for the
 * column-sliced layout); which=3: the swept layout's row classes (int64[4]: split length T,
 * 0 for one class; workgroups; partial-plane elements; tasks of the plane-adding pass).
 * `host_out` must hold the element count given by spops_matrix_info. */
int spops_matrix_layout(spops_matrix A, int which, void *host_out);

int spops_matrix_destroy(spops_matrix A);

/* ---- LeSpGEMM preprocessing: MinHash + LSH (device) --------------------------------------- */
typedef struct spops_pairs_s *spops_pairs;

/* d_sigs[r*k + d] = min over the columns c of row r of (a_d*c + b_d) mod 2^64, a_d/b_d from
 * the LCG of include/spgemm_MinHashLSH.h:40-48 seeded with `seed`; empty rows UINT64_MAX.
 * Backs minhash_signatures_flat (include/spgemm_MinHashLSH.h:119-180).  k <= 256. */
int spops_minhash_signatures(const void *d_rowptr, const void *d_col, int index_type, int64_t n_rows, int k,
                             uint64_t seed, uint64_t *d_sigs, void *stream);

/* Candidate pairs (i < j, sorted by (i,j), unique) of rows that share a band bucket, with
 * score = matching signature dimensions / k stored as `score_type` (F64 or F32).  Buckets
 * of more than 256 rows pair each row with its next 16 when `constrain` != 0.
 * Backs lsh_candidate_pairs_vector (include/spgemm_MinHashLSH.h:777-791). */
int spops_lsh_candidate_pairs(const void *d_rowptr, const void *d_col, int index_type, int64_t n_rows, int k,
                              int bands, uint64_t seed, int constrain, int score_type, spops_pairs *out,
                              void *stream);
/* Same, from precomputed device signatures (lsh_candidate_pairs_from_flat_vector, :427-605). */
int spops_lsh_pairs_from_signatures(const uint64_t *d_sigs, int64_t n_rows, int k, int bands, int constrain,
                                    int score_type, spops_pairs *out, void *stream);
int spops_pairs_count(spops_pairs P, int64_t *count);
int spops_pairs_device(spops_pairs P, const int64_t **d_i, const int64_t **d_j, const void **d_score);
int spops_pairs_download(spops_pairs P, int64_t *i, int64_t *j, void *score);
int spops_pairs_destroy(spops_pairs P);

/* ---- LeSpGEMM preprocessing: hierarchical clustering (host, sequential, bit-exact) --------
 * version 0: hierarchical_clustering_v0_fast (include/spgemm_utility.h:743-915)
 * version 1: hierarchical_clustering_v1      (include/spgemm_utility.h:609-733)
 * Host arrays.  pairs (pi, pj, score) as produced by LSH (sorted by (i,j)); score_type F64
 * or F32 selects the reference's ValueType.  perm[n_rows] receives the row permutation
 * (perm[new] = original row), offsets must hold n_rows + 1 entries; *nclusters is set. */
int spops_hierarchical_clustering(int version, int64_t n_rows, const int64_t *rowptr, const int64_t *col,
                                  int64_t npairs, const int64_t *pi, const int64_t *pj, const void *score,
                                  int score_type, int64_t cluster_size, int64_t *perm, int64_t *offsets,
                                  int64_t *nclusters);

/* hierarchical_clustering_v0 (include/spgemm_utility.h:517-606), the map-based original that
 * test/test_spgemm_hc.cpp runs on candidate pairs from a file (e.g. the A * A^T top-k pairs of
 * spops_spgemm_topk_jaccard).  The pairs (pi, pj, score) in the order of the caller's map,
 * keys as it holds them (non-negative; i > j allowed).  perm and offsets as above; roots[c]
 * receives the root row of cluster c (the reference returns map<root, rows>, clusters ordered
 * by root, rows ascending).  The caller's map is not extended with the pairs scored on the way
 * (the reference inserts them). */
int spops_hierarchical_clustering_v0(int64_t n_rows, const int64_t *rowptr, const int64_t *col, int64_t npairs,
                                     const int64_t *pi, const int64_t *pj, const void *score, int score_type,
                                     int64_t cluster_size, int64_t *perm, int64_t *offsets, int64_t *roots,
                                     int64_t *nclusters);

/* Approximate row clustering on the GPU, an alternative to the sequential hierarchical
 * clustering (include/spgemm_utility.h:743-915) behind a flag: no reference counterpart.  The
 * candidate pairs (device arrays, e.g. from spops_pairs_device; scores >= 0 of score_type F64 or
 * F32) are the edges of a graph over the rows; clusters merge in rounds in which every cluster
 * picks its best edge to another cluster whose size keeps the pair within cluster_size (highest
 * score, ties by the smaller splitmix64 hash of the (min id, max id) key) and clusters that picked
 * each other merge.
 * Rounds stop when none merges or after max_rounds (*rounds = rounds that merged).  perm[n_rows]
 * and offsets[n_rows + 1] (host) as spops_hierarchical_clustering: clusters ordered by their
 * smallest row, rows ascending inside; every cluster has at most cluster_size rows (the
 * reference's can exceed it: it stops growing a cluster once it reaches the size).  Deterministic.
 * Synchronous on `stream`. */
int spops_cluster_matching(int64_t n_rows, int64_t npairs, const int64_t *d_pi, const int64_t *d_pj,
                           const void *d_score, int score_type, int64_t cluster_size, int64_t max_rounds,
                           int64_t *perm, int64_t *offsets, int64_t *nclusters, int64_t *rounds, void *stream);

/* ---- LeSpGEMM: cluster format and cluster-wise SpGEMM (device) ----------------------------
 * A spops_cluster is a device CSR_VlengthCluster (include/spgemm_cluster.h:108-143):
 * cluster c = rows offsets[c]..offsets[c+1] of the permuted matrix, its distinct columns
 * (ascending) and a dense cluster_sz-wide value block per column. */
typedef struct spops_cluster_s *spops_cluster;

/* csr_reorder_rows + csr_to_vlength_cluster (include/sparse_conversion.h:1879, :1641).
 * A on the device (int64 rowptr/col, values F64 or F32); perm[n_rows] and
 * offsets[nclusters+1] on the host (as produced by spops_hierarchical_clustering). */
int spops_cluster_create(spops_cluster *out, int64_t n_rows, int64_t n_cols, int64_t nnz, const int64_t *d_rowptr,
                         const int64_t *d_col, const void *d_values, int value_type, const int64_t *perm,
                         int64_t nclusters, const int64_t *offsets, void *stream);
/* Device cluster matrix from a host CSR_VlengthCluster (the drop-in path of
 * LeSpGEMM_VLength, src/spgemm.cpp:662): arrays as in the struct, any column order. */
int spops_cluster_create_from_host(spops_cluster *out, int64_t csr_rows, int64_t cols, int64_t nclusters,
                                   const int64_t *cluster_sz, const int64_t *rowptr, const int64_t *rowptr_val,
                                   const int64_t *colids, const void *values, int value_type);
/* info[0]=clusters info[1]=csr rows info[2]=cols info[3]=nnzc info[4]=nnzv
 * info[5]=max cluster size info[6]=value type info[7]=device bytes */
int spops_cluster_info(spops_cluster A, int64_t info[8]);
/* host copies (int64 indices; values in the cluster's value type); any pointer may be NULL */
int spops_cluster_download(spops_cluster A, int64_t *cluster_sz, int64_t *rowptr, int64_t *rowptr_val,
                           int64_t *colids, void *values);
/* per-cluster symbolic results of a product (nz, min and max column), as kept by kernel 3 */
int spops_cluster_symbolic_download(spops_cluster C, int64_t *nz, int64_t *cmin, int64_t *cmax);
int spops_cluster_destroy(spops_cluster A);

/* C = A * B, cluster-wise with the hybrid accumulator (LeSpGEMM_VLength kernel 3,
 * src/spgemm.cpp:545-680).  B on the device: int64 rowptr/col, values of A's value type.
 * C's columns ascend inside each cluster.  B's rows need not be sorted (an unsorted B is
 * sorted into a temporary copy).  stats (may be NULL): [0]=1 if B was sorted into a copy
 * [1]=nnzc [2]=nnzv [3]=small clusters (one LDS block each) [4]=large clusters (chunked LDS
 * accumulator) [5..8]=setup/symbolic/malloc/numeric in us [9]=cluster-wise products
 * [10]=dense clusters [11]=numeric tasks of the large clusters [12]=those of them in the
 * wide-cursor variant. */
int spops_lespgemm(spops_cluster A, int64_t b_rows, int64_t b_cols, int64_t b_nnz, const int64_t *d_browptr,
                   const int64_t *d_bcol, const void *d_bval, spops_cluster *C, int64_t stats[16], void *stream);

/* The same product in panels (cluster ranges), for outputs larger than memory: no reference
 * counterpart as an entry point; each panel is what LeSpGEMM_VLength kernel 3
 * (src/spgemm.cpp:545-659) computes for clusters [c0, c1).  plan_create prepares B and runs
 * the symbolic phase for every cluster; A and the device arrays of B must outlive the plan.
 * info = {clusters, nnzc, nnzv, products, small clusters, large clusters, setup us,
 * symbolic us}.  plan_panels cuts [0, clusters) into ranges whose output (colids + values)
 * stays under budget_bytes (a range holds at least one cluster); bounds needs clusters + 2
 * entries.  plan_numeric computes one range into *C (created when *C is NULL, else its
 * buffers are reused); C then has c1 - c0 clusters.  stats as spops_lespgemm (setup and
 * symbolic count once, in the plan). */
typedef struct spops_spgemm_plan_s *spops_spgemm_plan;
int spops_lespgemm_plan_create(spops_cluster A, int64_t b_rows, int64_t b_cols, int64_t b_nnz,
                               const int64_t *d_browptr, const int64_t *d_bcol, const void *d_bval,
                               spops_spgemm_plan *plan, void *stream);
int spops_lespgemm_plan_info(spops_spgemm_plan plan, int64_t info[8]);
int spops_lespgemm_plan_panels(spops_spgemm_plan plan, int64_t budget_bytes, int64_t *npanels, int64_t *bounds);
int spops_lespgemm_plan_numeric(spops_spgemm_plan plan, int64_t c0, int64_t c1, spops_cluster *C, int64_t stats[16],
                                void *stream);
int spops_lespgemm_plan_destroy(spops_spgemm_plan plan);
/* checksums of a cluster matrix's values and column ids: out = {sum of values, sum of |values|,
 * sum of column ids, number of nonzero values} (for comparing panels with another implementation) */
int spops_cluster_checksum(spops_cluster C, double out[4], void *stream);

/* vlength_cluster2csr + inverse permutation (include/sparse_conversion.h:1778-1862,
 * test/test_spgemm_hc_lsh.cpp:296-303): C in the original row order, exact zeros dropped,
 * columns ascending.  The three device arrays are allocated here; free with spops_free. */
int spops_cluster_to_csr(spops_cluster C, spops_cluster A, int64_t *nnz, int64_t **d_rowptr, int64_t **d_col,
                         void **d_val, void *stream);

/* Row-wise C = A * B in CSR: the reference's LeSpGEMM (src/spgemm.cpp:684-705), whose kernel
 * flags 1 (hash, :45-104), 2 (array, :110-169) and 3 (SPA, :175-236) all give this C.  A and B
 * on the device (int64 rowptr/col, values of value_type F64 or F32; B's rows need not be
 * sorted).  C's columns ascend in every row (what sortOutput = true asks; with false the
 * reference leaves its hash order) and every column of the symbolic product is kept, zero sums
 * included.  The three output arrays (rowptr a_rows + 1, col and values c_nnz) are allocated
 * here; free with spops_free.  stats as spops_lespgemm (may be NULL).  Synchronous. */
int spops_spgemm_csr(int64_t a_rows, int64_t a_cols, int64_t a_nnz, const int64_t *d_arowptr, const int64_t *d_acol,
                     const void *d_aval, int64_t b_cols, int64_t b_nnz, const int64_t *d_browptr, const int64_t *d_bcol,
                     const void *d_bval, int value_type, int64_t *c_nnz, int64_t **d_crowptr, int64_t **d_ccol,
                     void **d_cval, int64_t stats[16], void *stream);

/* HashSpGEMMTopK (src/spgemm.cpp:707-842): the candidate pairs of the A * A^T front end of the
 * hierarchical clustering.  Row i of C holds the rows j != i of A that share a column with row
 * i, scored by s = |A_i & A_j| / (|A_i| + |A_j| - |A_i & A_j|) computed in the value type
 * (row lengths |A_i| count stored entries, duplicates included), at most top_k of them,
 * ordered by s descending then j ascending (src/spgemm_hash.cpp:99-108, 169-230).  A's pattern
 * on the device (int64 rowptr/col; its values are not read).  The intersection counts are
 * computed in row panels whose output stays under budget_bytes (<= 0: a quarter of the free
 * device memory).  Output arrays allocated here (free with spops_free).  Synchronous. */
int spops_spgemm_topk_jaccard(int64_t n_rows, int64_t n_cols, int64_t nnz, const int64_t *d_rowptr,
                              const int64_t *d_col, int value_type, int64_t top_k, int64_t budget_bytes,
                              int64_t *c_nnz, int64_t **d_crowptr, int64_t **d_ccol, void **d_cval, void *stream);

/* get_spgemm_flop (src/spgemm.cpp:24-42): 2 * sum over A entries (i,k), k < b_rows, of nnz(B_k). */
int spops_spgemm_flop(int64_t a_nnz, const int64_t *d_acol, int64_t b_rows, const int64_t *d_browptr, int64_t *flops,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SPOPS_HIP_H */
