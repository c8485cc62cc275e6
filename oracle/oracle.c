/* CPU restatement of the reference hot path (lsl036/SparseOps) -- TEST INFRASTRUCTURE ONLY.
 *
 * This is the oracle: tests/ and bench.py's cpu_baseline compare the HIP library
 * against it.  It is never linked into, called by or shipped with the product path.
 *
 * How each part is pinned (beyond the reference's own sources, which it follows line
 * by line as cited):
 *   - SpMV (CSR/ELL/SELL-C-sigma/SELL-C-R/BSR) and the format conversions: bit-identical
 *     layouts against the reference built from source (oracle/_ref/ref_harness) on the
 *     committed fixtures, y within FP tolerance of the reference; the 5-point Laplacian
 *     closed form (A*1 = boundary-degree vector) pins the CSR sum.
 *   - MinHash: exact equality with the reference's signatures (fixtures), plus the closed
 *     form of the LCG coefficients and a brute-force min over a 1-row matrix.
 *   - LSH pairs: exact pair-set equality with the reference (fixtures) and a brute-force
 *     band-collision enumeration on tiny inputs.
 *   - HC v0_fast / v1: byte-equal perm/offsets against the reference (fixtures); invariants
 *     (perm is a permutation, clusters contiguous, sizes <= 2*cs-2, roots ascending).
 *   - cluster SpGEMM: C structure equal to the reference's C (fixtures) and to a
 *     dense-matrix product on tiny inputs; values within 1e-12 relative.
 *   - SpMM/SDDMM/transpose: dense numpy products on tiny inputs.
 * Compiled with -ffp-contract=off (no FMA contraction) so sums follow the written order.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void *p) { free(p); }

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) abort();
  return p;
}

/* ------------------------------------------------------------------------------------ */
/* SpMV                                                                                   */
/* ------------------------------------------------------------------------------------ */

/* src/spmv_csr.cpp:65-90 (__spmv_csr_serial_simple): row sum in file order, then
 * y = sum when alpha==1 && beta==0, else y = alpha*sum + beta*y. */
void orc_spmv_csr(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                  double alpha, const double *x, double beta, double *y) {
  for (int64_t r = 0; r < n_rows; ++r) {
    double sum = 0.0;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) sum += val[j] * x[col[j]];
    if (alpha == 1.0 && beta == 0.0) y[r] = sum;
    else y[r] = alpha * sum + beta * y[r];
  }
}

/* Same as orc_spmv_csr with ValueType=float (src/spmv_csr.cpp:187 instantiation). */
void orc_spmv_csr_f32(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const float *val,
                      float alpha, const float *x, float beta, float *y) {
  for (int64_t r = 0; r < n_rows; ++r) {
    float sum = 0.0f;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) sum += val[j] * x[col[j]];
    if (alpha == 1.0f && beta == 0.0f) y[r] = sum;
    else y[r] = alpha * sum + beta * y[r];
  }
}

/* Extended-precision version of the same sum (the FP64 oracle for error reports);
 * absrow[r] = sum_j |a_rj * x_j| (the condition scale of row r), may be NULL. */
void orc_spmv_csr_ld(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                     double alpha, const double *x, double beta, double *y, double *absrow) {
  for (int64_t r = 0; r < n_rows; ++r) {
    long double sum = 0.0L, asum = 0.0L;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      long double p = (long double)val[j] * (long double)x[col[j]];
      sum += p;
      asum += fabsl(p);
    }
    long double out = (long double)alpha * sum;
    if (beta != 0.0) out += (long double)beta * (long double)y[r];
    y[r] = (double)out;
    if (absrow) absrow[r] = (double)asum;
  }
}

/* include/sparse_conversion.h:345-487 (csr_to_sell_c_sigma), steps 1-2.
 * reorder[g] = original row placed at sorted position g: each sigma-slice is sorted by
 * (nnz, row) descending (std::greater on pair, :391-408).  chunk_len[c] = max nnz of the
 * C sorted rows of chunk c rounded up to `align` (:413-441).  chunk_off[c] = start of
 * chunk c in the concatenated chunk arrays (each chunk holds chunk_len*C slots).
 * Returns the total slot count.  sigma % C != 0 returns -1 (the reference exits). */
typedef struct { int64_t nnz, row; } nnz_row_t;
static int cmp_nnz_row_desc(const void *a, const void *b) {
  const nnz_row_t *x = (const nnz_row_t *)a, *y = (const nnz_row_t *)b;
  if (x->nnz != y->nnz) return x->nnz > y->nnz ? -1 : 1;
  if (x->row != y->row) return x->row > y->row ? -1 : 1;
  return 0;
}

int64_t orc_sell_layout(int64_t n_rows, const int64_t *rowptr, int64_t sigma, int64_t C, int64_t align,
                        int64_t *reorder, int64_t *chunk_len, int64_t *chunk_off) {
  if (sigma % C) return -1;
  nnz_row_t *t = (nnz_row_t *)xmalloc(sizeof(nnz_row_t) * (size_t)(n_rows ? n_rows : 1));
  for (int64_t i = 0; i < n_rows; ++i) { t[i].nnz = rowptr[i + 1] - rowptr[i]; t[i].row = i; }
  for (int64_t s = 0; s * sigma < n_rows; ++s) {
    int64_t lo = s * sigma, hi = lo + sigma < n_rows ? lo + sigma : n_rows;
    qsort(t + lo, (size_t)(hi - lo), sizeof(nnz_row_t), cmp_nnz_row_desc);
  }
  for (int64_t i = 0; i < n_rows; ++i) reorder[i] = t[i].row;
  int64_t nchunks = (n_rows + C - 1) / C, total = 0;
  for (int64_t c = 0; c < nchunks; ++c) chunk_len[c] = 0;
  for (int64_t g = 0; g < n_rows; ++g) {
    int64_t r = reorder[g], len = rowptr[r + 1] - rowptr[r];
    if (len > chunk_len[g / C]) chunk_len[g / C] = len;
  }
  for (int64_t c = 0; c < nchunks; ++c) {
    chunk_len[c] = (chunk_len[c] + align - 1) / align * align;
    chunk_off[c] = total;
    total += chunk_len[c] * C;
  }
  free(t);
  return total;
}

/* include/sparse_conversion.h:446-484, step 3: row-major within a chunk,
 * pos = row_in_chunk * chunk_len + k, padding col -1 / value 0. */
void orc_sell_fill(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                   int64_t C, const int64_t *reorder, const int64_t *chunk_len, const int64_t *chunk_off,
                   int64_t *scol, double *sval) {
  int64_t nchunks = (n_rows + C - 1) / C;
  for (int64_t c = 0; c < nchunks; ++c)
    for (int64_t s = 0; s < chunk_len[c] * C; ++s) { scol[chunk_off[c] + s] = -1; sval[chunk_off[c] + s] = 0.0; }
  for (int64_t g = 0; g < n_rows; ++g) {
    int64_t r = reorder[g], c = g / C, l = g % C;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      int64_t pos = chunk_off[c] + l * chunk_len[c] + (j - rowptr[r]);
      scol[pos] = col[j];
      sval[pos] = val[j];
    }
  }
}

/* src/spmv_sell_c_sigma.cpp:18-57 (serial): per sorted row, sum over the chunk width
 * skipping padding, y[reorder[g]] = alpha*sum + beta*y.  Padding is masked here; the
 * reference reads x[(size_t)-1] for it (:46), which is undefined. */
void orc_spmv_sell(int64_t n_rows, int64_t C, const int64_t *reorder, const int64_t *chunk_len,
                   const int64_t *chunk_off, const int64_t *scol, const double *sval,
                   double alpha, const double *x, double beta, double *y) {
  int64_t nchunks = (n_rows + C - 1) / C;
  for (int64_t c = 0; c < nchunks; ++c) {
    for (int64_t l = 0; l < C; ++l) {
      int64_t g = c * C + l;
      if (g >= n_rows) break;
      double sum = 0.0;
      for (int64_t k = 0; k < chunk_len[c]; ++k) {
        int64_t pos = chunk_off[c] + l * chunk_len[c] + k;
        if (scol[pos] >= 0) sum += sval[pos] * x[scol[pos]];
      }
      y[reorder[g]] = alpha * sum + beta * y[reorder[g]];
    }
  }
}

/* include/sparse_conversion.h:195-256 (csr_to_ell): width = max row nnz, no rounding. */
int64_t orc_ell_width(int64_t n_rows, const int64_t *rowptr) {
  int64_t w = 0;
  for (int64_t i = 0; i < n_rows; ++i)
    if (rowptr[i + 1] - rowptr[i] > w) w = rowptr[i + 1] - rowptr[i];
  return w;
}

void orc_ell_fill(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                  int64_t width, int col_major, int64_t *ecol, double *evals) {
  for (int64_t s = 0; s < n_rows * width; ++s) { ecol[s] = -1; evals[s] = 0.0; }
  for (int64_t r = 0; r < n_rows; ++r)
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      int64_t t = j - rowptr[r];
      int64_t pos = col_major ? r + t * n_rows : r * width + t;
      ecol[pos] = col[j];
      evals[pos] = val[j];
    }
}

/* src/spmv_ell.cpp:17-68 (serial): skip slots with col < 0. */
void orc_spmv_ell(int64_t n_rows, int64_t width, int col_major, const int64_t *ecol, const double *evals,
                  double alpha, const double *x, double beta, double *y) {
  for (int64_t r = 0; r < n_rows; ++r) {
    double sum = 0.0;
    for (int64_t t = 0; t < width; ++t) {
      int64_t pos = col_major ? r + t * n_rows : r * width + t;
      if (ecol[pos] >= 0) sum += evals[pos] * x[ecol[pos]];
    }
    y[r] = alpha * sum + beta * y[r];
  }
}

/* include/sparse_conversion.h:777-894 (csr_to_bsr): block row i holds the distinct
 * block columns (col / bC) of rows [i*bR, (i+1)*bR), ascending.  Returns nnzb. */
int64_t orc_bsr_count(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col,
                      int64_t bR, int64_t bC, int64_t *brow_ptr) {
  int64_t mb = (n_rows + bR - 1) / bR, nb = (n_cols + bC - 1) / bC;
  char *mark = (char *)xmalloc((size_t)nb);
  brow_ptr[0] = 0;
  for (int64_t i = 0; i < mb; ++i) {
    memset(mark, 0, (size_t)nb);
    int64_t hi = (i + 1) * bR < n_rows ? (i + 1) * bR : n_rows, cnt = 0;
    for (int64_t j = rowptr[i * bR]; j < rowptr[hi]; ++j) mark[col[j] / bC] = 1;
    for (int64_t b = 0; b < nb; ++b) cnt += mark[b];
    brow_ptr[i + 1] = brow_ptr[i] + cnt;
  }
  free(mark);
  return brow_ptr[mb];
}

/* include/sparse_conversion.h:896-934: block data row-major inside each bR x bC block. */
void orc_bsr_fill(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col, const double *val,
                  int64_t bR, int64_t bC, const int64_t *brow_ptr, int64_t *bcol, double *bdata) {
  int64_t mb = (n_rows + bR - 1) / bR, nb = (n_cols + bC - 1) / bC;
  char *mark = (char *)xmalloc((size_t)nb);
  int64_t w = 0;
  for (int64_t i = 0; i < mb; ++i) {
    memset(mark, 0, (size_t)nb);
    int64_t hi = (i + 1) * bR < n_rows ? (i + 1) * bR : n_rows;
    for (int64_t j = rowptr[i * bR]; j < rowptr[hi]; ++j) mark[col[j] / bC] = 1;
    for (int64_t b = 0; b < nb; ++b) if (mark[b]) bcol[w++] = b;
  }
  free(mark);
  memset(bdata, 0, sizeof(double) * (size_t)(brow_ptr[mb] * bR * bC));
  for (int64_t r = 0; r < n_rows; ++r) {
    int64_t br = r / bR;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      int64_t bc = col[j] / bC, k;
      for (k = brow_ptr[br]; k < brow_ptr[br + 1]; ++k) if (bcol[k] == bc) break;
      bdata[k * bR * bC + (r % bR) * bC + col[j] % bC] = val[j];
    }
  }
}

/* src/spmv_bsr.cpp:18-63 (serial): tmp[br] accumulates blocks then in-block columns;
 * y[i*bR+br] = alpha*tmp + beta*y for rows < num_rows.  x beyond n_cols (padding
 * columns of the last block column) counts as 0; the reference reads past x there. */
void orc_spmv_bsr(int64_t n_rows, int64_t n_cols, int64_t bR, int64_t bC, const int64_t *brow_ptr,
                  const int64_t *bcol, const double *bdata, double alpha, const double *x, double beta, double *y) {
  int64_t mb = (n_rows + bR - 1) / bR;
  double *tmp = (double *)xmalloc(sizeof(double) * (size_t)bR);
  for (int64_t i = 0; i < mb; ++i) {
    for (int64_t r = 0; r < bR; ++r) tmp[r] = 0.0;
    for (int64_t j = brow_ptr[i]; j < brow_ptr[i + 1]; ++j)
      for (int64_t r = 0; r < bR; ++r)
        for (int64_t c = 0; c < bC; ++c) {
          int64_t xi = bcol[j] * bC + c;
          double xv = xi < n_cols ? x[xi] : 0.0;
          tmp[r] += bdata[j * bR * bC + r * bC + c] * xv;
        }
    for (int64_t r = 0; r < bR; ++r)
      if (i * bR + r < n_rows) y[i * bR + r] = alpha * tmp[r] + beta * y[i * bR + r];
  }
  free(tmp);
}

/* ------------------------------------------------------------------------------------ */
/* MinHash + LSH                                                                          */
/* ------------------------------------------------------------------------------------ */

/* include/spgemm_MinHashLSH.h:40-48 (lcg_next, hash_apply) and :119-180
 * (minhash_signatures_flat): for d in [0,k): state advances twice, a_d = (s1>>1)*2+1,
 * b_d = s2; sig[r*k+d] = min over the row's columns of a_d*col + b_d (mod 2^64);
 * empty rows keep UINT64_MAX. */
void orc_minhash(int64_t n_rows, const int64_t *rowptr, const int64_t *col, int k, uint64_t seed,
                 uint64_t *sigs) {
  uint64_t *a = (uint64_t *)xmalloc(sizeof(uint64_t) * (size_t)k);
  uint64_t *b = (uint64_t *)xmalloc(sizeof(uint64_t) * (size_t)k);
  uint64_t s = seed;
  for (int d = 0; d < k; ++d) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    a[d] = (s >> 1) * 2 + 1;
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    b[d] = s;
  }
  for (int64_t r = 0; r < n_rows; ++r) {
    uint64_t *sig = sigs + (size_t)r * (size_t)k;
    for (int d = 0; d < k; ++d) sig[d] = UINT64_MAX;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      uint64_t xv = (uint64_t)col[j];
      for (int d = 0; d < k; ++d) {
        uint64_t h = a[d] * xv + b[d];
        if (h < sig[d]) sig[d] = h;
      }
    }
  }
  free(a);
  free(b);
}

/* include/spgemm_MinHashLSH.h:277-283 (band_hash). */
uint64_t orc_band_hash(const uint64_t *band_sig, int r) {
  uint64_t h = 0;
  for (int i = 0; i < r; ++i) h = h * 0x9e3779b97f4a7c15ULL + band_sig[i];
  return h;
}

typedef struct { uint64_t bucket; int64_t row; } rec_t;
static int cmp_rec(const void *a, const void *b) {
  const rec_t *x = (const rec_t *)a, *y = (const rec_t *)b;
  if (x->bucket != y->bucket) return x->bucket < y->bucket ? -1 : 1;
  if (x->row != y->row) return x->row < y->row ? -1 : 1;
  return 0;
}
typedef struct { int64_t i, j; } pair_t;
static int cmp_pair(const void *a, const void *b) {
  const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  return 0;
}

typedef struct { pair_t *p; int64_t n, cap; } pairvec_t;
static void pv_push(pairvec_t *v, int64_t i, int64_t j) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 1024;
    v->p = (pair_t *)realloc(v->p, sizeof(pair_t) * (size_t)v->cap);
    if (!v->p) abort();
  }
  v->p[v->n].i = i;
  v->p[v->n].j = j;
  v->n++;
}

/* include/spgemm_MinHashLSH.h:427-605 (lsh_candidate_pairs_from_flat_vector) and
 * :235-270 (k_bucket_size_limit = 256, k_window_pairs = 16): per band, rows with equal
 * band hash form a bucket in ascending row order; a bucket of size <= 256 (or any size
 * when !constrain) emits all (a<b) pairs, a larger one pairs each row with its next <= 16
 * members.  The union over bands is sorted and made unique; score = matches / k.
 * Returns k % bands != 0 or empty input as 0 pairs (the reference returns empty). */
int64_t orc_lsh_pairs(const uint64_t *sigs, int64_t n_rows, int k, int bands, int constrain,
                      int64_t **pi, int64_t **pj, double **score) {
  pairvec_t pv = {0, 0, 0};
  int r = bands > 0 ? k / bands : 0;
  if (n_rows > 0 && bands > 0 && k > 0 && r > 0 && r * bands == k) {
    rec_t *rec = (rec_t *)xmalloc(sizeof(rec_t) * (size_t)n_rows);
    for (int band = 0; band < bands; ++band) {
      for (int64_t row = 0; row < n_rows; ++row) {
        rec[row].bucket = orc_band_hash(sigs + (size_t)row * (size_t)k + band * r, r);
        rec[row].row = row;
      }
      qsort(rec, (size_t)n_rows, sizeof(rec_t), cmp_rec);
      for (int64_t s = 0; s < n_rows;) {
        int64_t e = s + 1;
        while (e < n_rows && rec[e].bucket == rec[s].bucket) ++e;
        int64_t g = e - s;
        if (g > 1) {
          for (int64_t a = s; a < e; ++a) {
            int64_t bend = e;
            if (constrain && g > 256 && a + 1 + 16 < e) bend = a + 1 + 16;
            for (int64_t b = a + 1; b < bend; ++b) pv_push(&pv, rec[a].row, rec[b].row);
          }
        }
        s = e;
      }
    }
    free(rec);
  }
  if (pv.n) qsort(pv.p, (size_t)pv.n, sizeof(pair_t), cmp_pair);
  int64_t u = 0;
  for (int64_t t = 0; t < pv.n; ++t)
    if (u == 0 || pv.p[t].i != pv.p[u - 1].i || pv.p[t].j != pv.p[u - 1].j) pv.p[u++] = pv.p[t];
  *pi = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)u);
  *pj = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)u);
  *score = (double *)xmalloc(sizeof(double) * (size_t)u);
  for (int64_t t = 0; t < u; ++t) {
    const uint64_t *sa = sigs + (size_t)pv.p[t].i * (size_t)k, *sb = sigs + (size_t)pv.p[t].j * (size_t)k;
    int match = 0;
    for (int d = 0; d < k; ++d) match += sa[d] == sb[d];
    (*pi)[t] = pv.p[t].i;
    (*pj)[t] = pv.p[t].j;
    (*score)[t] = (double)match / (double)k;   /* :183-204 minhash_estimated_jaccard */
  }
  free(pv.p);
  return u;
}

/* ------------------------------------------------------------------------------------ */
/* Hierarchical clustering                                                                */
/* ------------------------------------------------------------------------------------ */

/* include/spgemm_utility.h:362-387 (jaccard_similarity, set-based; duplicates in row b
 * count with multiplicity) and :402-429 (jaccard_similarity_sorted_unique). */
double orc_jaccard(const int64_t *rowptr, const int64_t *col, int64_t a, int64_t b, int sorted_unique) {
  int64_t la = rowptr[a + 1] - rowptr[a], lb = rowptr[b + 1] - rowptr[b], c = 0;
  if (sorted_unique) {
    int64_t ia = rowptr[a], ib = rowptr[b];
    while (ia < rowptr[a + 1] && ib < rowptr[b + 1]) {
      if (col[ia] == col[ib]) { ++c; ++ia; ++ib; }
      else if (col[ia] < col[ib]) ++ia;
      else ++ib;
    }
  } else {
    for (int64_t j = rowptr[b]; j < rowptr[b + 1]; ++j) {
      int found = 0;
      for (int64_t i = rowptr[a]; i < rowptr[a + 1] && !found; ++i) found = col[i] == col[j];
      c += found;
    }
  }
  int64_t u = la + lb - c;
  return u == 0 ? 0.0 : (1.0 * (double)c / (double)u);
}

static int rows_sorted_unique(int64_t n, const int64_t *rowptr, const int64_t *col) {
  for (int64_t r = 0; r < n; ++r)
    for (int64_t j = rowptr[r] + 1; j < rowptr[r + 1]; ++j)
      if (col[j - 1] >= col[j]) return 0;
  return 1;
}

typedef struct { double s; int64_t i, j; } item_t;
/* comparator of spgemm_utility.h:774-776: a < b iff (score, (i,j)) is lexicographically less */
static int item_less(const item_t *a, const item_t *b) {
  if (a->s != b->s) return a->s < b->s;
  if (a->i != b->i) return a->i < b->i;
  return a->j < b->j;
}
typedef struct { item_t *h; int64_t n, cap; } heap_t;
static void heap_push(heap_t *q, item_t it) {
  if (q->n == q->cap) {
    q->cap = q->cap ? q->cap * 2 : 1024;
    q->h = (item_t *)realloc(q->h, sizeof(item_t) * (size_t)q->cap);
    if (!q->h) abort();
  }
  int64_t c = q->n++;
  q->h[c] = it;
  while (c > 0) {
    int64_t p = (c - 1) / 2;
    if (!item_less(&q->h[p], &q->h[c])) break;
    item_t t = q->h[p]; q->h[p] = q->h[c]; q->h[c] = t;
    c = p;
  }
}
static item_t heap_pop(heap_t *q) {
  item_t top = q->h[0];
  q->h[0] = q->h[--q->n];
  int64_t c = 0;
  for (;;) {
    int64_t l = 2 * c + 1, r = l + 1, m = c;
    if (l < q->n && item_less(&q->h[m], &q->h[l])) m = l;
    if (r < q->n && item_less(&q->h[m], &q->h[r])) m = r;
    if (m == c) break;
    item_t t = q->h[m]; q->h[m] = q->h[c]; q->h[c] = t;
    c = m;
  }
  return top;
}

/* open-addressing set of (i,j) pairs (membership only; replaces std::unordered_set /
 * std::set, whose iteration order the reference never uses) */
typedef struct { int64_t *k; int64_t cap, n; } pset_t;
static uint64_t mix64(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
  return x;
}
static void pset_init(pset_t *s, int64_t cap) {
  int64_t c = 1024;
  while (c < 2 * cap) c <<= 1;
  s->cap = c; s->n = 0;
  s->k = (int64_t *)xmalloc(sizeof(int64_t) * 2 * (size_t)c);
  for (int64_t t = 0; t < 2 * c; ++t) s->k[t] = -1;
}
static int pset_insert(pset_t *s, int64_t i, int64_t j);
static void pset_grow(pset_t *s) {
  pset_t t;
  pset_init(&t, s->cap);
  for (int64_t q = 0; q < s->cap; ++q)
    if (s->k[2 * q] >= 0) pset_insert(&t, s->k[2 * q], s->k[2 * q + 1]);
  free(s->k);
  *s = t;
}
/* returns 1 if newly inserted, 0 if present */
static int pset_insert(pset_t *s, int64_t i, int64_t j) {
  if (2 * (s->n + 1) > s->cap) pset_grow(s);
  uint64_t h = mix64((uint64_t)i * 0x9e3779b97f4a7c15ULL ^ mix64((uint64_t)j)) & (uint64_t)(s->cap - 1);
  for (;;) {
    if (s->k[2 * h] < 0) { s->k[2 * h] = i; s->k[2 * h + 1] = j; s->n++; return 1; }
    if (s->k[2 * h] == i && s->k[2 * h + 1] == j) return 0;
    h = (h + 1) & (uint64_t)(s->cap - 1);
  }
}
static int pset_has(const pset_t *s, int64_t i, int64_t j) {
  uint64_t h = mix64((uint64_t)i * 0x9e3779b97f4a7c15ULL ^ mix64((uint64_t)j)) & (uint64_t)(s->cap - 1);
  for (;;) {
    if (s->k[2 * h] < 0) return 0;
    if (s->k[2 * h] == i && s->k[2 * h + 1] == j) return 1;
    h = (h + 1) & (uint64_t)(s->cap - 1);
  }
}

static int64_t uf_find(int64_t *cl, int64_t x) {
  int64_t root = x;
  while (cl[root] != root) root = cl[root];
  while (cl[x] != root) { int64_t nx = cl[x]; cl[x] = root; x = nx; }
  return root;
}

static double as_value(double v, int f32) { return f32 ? (double)(float)v : v; }

/* std::lower_bound over (pi,pj) with the comparator of spgemm_utility.h:789-796 */
static int initial_has(int64_t np, const int64_t *pi, const int64_t *pj, int64_t i, int64_t j) {
  int64_t first = 0, len = np;
  while (len > 0) {
    int64_t half = len >> 1, mid = first + half;
    int less = pi[mid] != i ? pi[mid] < i : pj[mid] < j;
    if (less) { first = mid + 1; len = len - half - 1; }
    else len = half;
  }
  return first < np && pi[first] == i && pj[first] == j;
}

typedef struct { double s; int64_t pos; } ord_t;
static int cmp_ord(const void *a, const void *b) {
  const ord_t *x = (const ord_t *)a, *y = (const ord_t *)b;
  if (x->s != y->s) return x->s > y->s ? -1 : 1;       /* score descending */
  return x->pos > y->pos ? -1 : (x->pos < y->pos);     /* later input first (pop_back) */
}

/* Builds perm/offsets from the union-find forest: include/spgemm_utility.h:883-914 */
static int64_t hc_output(int64_t n, int64_t *cl, int64_t *perm, int64_t **offsets) {
  int64_t *root = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int64_t *count = (int64_t *)calloc((size_t)(n ? n : 1), sizeof(int64_t));
  int64_t *cid = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) { root[i] = uf_find(cl, i); count[root[i]]++; }
  int64_t nc = 0;
  for (int64_t r = 0; r < n; ++r) if (count[r] > 0) cid[r] = nc++;
  int64_t *off = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nc + 1));
  off[0] = 0;
  for (int64_t r = 0, c = 0; r < n; ++r) if (count[r] > 0) { off[c + 1] = off[c] + count[r]; ++c; }
  int64_t *cur = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nc ? nc : 1));
  for (int64_t c = 0; c < nc; ++c) cur[c] = off[c];
  for (int64_t i = 0; i < n; ++i) perm[cur[cid[root[i]]]++] = i;
  free(root); free(count); free(cid); free(cur);
  *offsets = off;
  return nc;
}

/* version 0: hierarchical_clustering_v0_fast, include/spgemm_utility.h:743-915
 * version 1: hierarchical_clustering_v1,      include/spgemm_utility.h:609-733
 * version 2: hierarchical_clustering_v0,      include/spgemm_utility.h:517-606 (the map-based
 *            original: pairs in the map's order, keys as the map holds them; a discovered root
 *            pair is scored unless its (min, max) key is a key already; set-semantics Jaccard)
 * score_f32 != 0 reproduces ValueType=float (scores and Jaccard stored as float).
 * Returns the number of clusters; *offsets is malloc'ed (nclusters+1 entries). */
int64_t orc_hc(int version, int64_t n, const int64_t *rowptr, const int64_t *col,
               int64_t np, const int64_t *pi, const int64_t *pj, const double *score,
               int score_f32, int64_t cs, int64_t *perm, int64_t **offsets) {
  int64_t *cl = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int64_t *sz = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  char *valid = (char *)xmalloc((size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) { cl[i] = i; sz[i] = 1; valid[i] = 1; }
  const int su = rows_sorted_unique(n, rowptr, col);
  heap_t dyn = {0, 0, 0};
  pset_t seen;
  pset_init(&seen, version >= 1 ? np : np / 8 + 1024);
  int nclusters = (int)n;

  if (version == 2) {
    for (int64_t t = 0; t < np; ++t) {
      pset_insert(&seen, pi[t], pj[t]);
      item_t it = {as_value(score[t], score_f32), pi[t] < pj[t] ? pi[t] : pj[t], pi[t] < pj[t] ? pj[t] : pi[t]};
      heap_push(&dyn, it);
    }
    int64_t left = n;
    while (dyn.n > 0 && left != 0) {
      item_t top = heap_pop(&dyn);
      int64_t i = top.i, j = top.j;
      if (i >= n || j >= n) continue;
      if (cl[i] == i && cl[j] == j) {
        if (!valid[i] || !valid[j]) continue;
        left--;
        if (sz[i] < sz[j]) { cl[i] = j; sz[j] += sz[i]; if (sz[j] >= cs) { valid[j] = 0; left--; } }
        else { cl[j] = i; sz[i] += sz[j]; if (sz[i] >= cs) { valid[i] = 0; left--; } }
      } else {
        int64_t ri = uf_find(cl, i), rj = uf_find(cl, j);
        if (!valid[ri] || !valid[rj]) continue;
        if (ri != rj) {
          int64_t a = ri < rj ? ri : rj, b = ri < rj ? rj : ri;
          if (pset_insert(&seen, a, b)) {
            item_t it = {as_value(orc_jaccard(rowptr, col, a, b, 0), score_f32), a, b};
            heap_push(&dyn, it);
          }
        }
      }
    }
  } else if (version == 1) {
    for (int64_t t = 0; t < np; ++t) {
      int64_t ii = pi[t], jj = pj[t];
      if (ii > jj) { int64_t s = ii; ii = jj; jj = s; }
      if (ii == jj || ii >= n || jj >= n) continue;
      if (!pset_insert(&seen, ii, jj)) continue;
      item_t it = {as_value(score[t], score_f32), ii, jj};
      heap_push(&dyn, it);
    }
    while (dyn.n > 0 && nclusters > 0) {
      item_t top = heap_pop(&dyn);
      int64_t i = top.i, j = top.j;
      if (i >= n || j >= n) continue;
      int64_t ri = uf_find(cl, i), rj = uf_find(cl, j);
      if (ri == rj) continue;
      if (!valid[ri] || !valid[rj]) continue;
      if (cl[i] == i && cl[j] == j) {
        nclusters--;
        if (sz[ri] < sz[rj]) { cl[ri] = rj; sz[rj] += sz[ri]; if (sz[rj] >= cs) { valid[rj] = 0; nclusters--; } }
        else { cl[rj] = ri; sz[ri] += sz[rj]; if (sz[ri] >= cs) { valid[ri] = 0; nclusters--; } }
      } else {
        int64_t a = ri < rj ? ri : rj, b = ri < rj ? rj : ri;
        if (pset_insert(&seen, a, b)) {
          item_t it = {as_value(orc_jaccard(rowptr, col, a, b, 0), score_f32), a, b};
          heap_push(&dyn, it);
        }
      }
    }
  } else {
    /* initial pairs in the order map<score,vector,greater> + pop_back yields */
    ord_t *ord = (ord_t *)xmalloc(sizeof(ord_t) * (size_t)(np ? np : 1));
    int64_t *oi = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(np ? np : 1));
    int64_t *oj = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(np ? np : 1));
    int64_t m = 0;
    for (int64_t t = 0; t < np; ++t) {
      int64_t ii = pi[t], jj = pj[t];
      if (ii > jj) { int64_t s = ii; ii = jj; jj = s; }
      if (ii == jj || ii >= n || jj >= n) continue;
      ord[m].s = as_value(score[t], score_f32);
      ord[m].pos = m;
      oi[m] = ii; oj[m] = jj;
      ++m;
    }
    qsort(ord, (size_t)m, sizeof(ord_t), cmp_ord);
    int64_t next = 0;
    while ((next < m || dyn.n > 0) && nclusters > 0) {
      item_t top;
      if (dyn.n == 0) {
        top.s = ord[next].s; top.i = oi[ord[next].pos]; top.j = oj[ord[next].pos]; ++next;
      } else if (next >= m) {
        top = heap_pop(&dyn);
      } else {
        item_t init = {ord[next].s, oi[ord[next].pos], oj[ord[next].pos]};
        if (item_less(&dyn.h[0], &init)) { top = init; ++next; }
        else top = heap_pop(&dyn);
      }
      int64_t i = top.i, j = top.j;
      if (i >= n || j >= n) continue;
      if (cl[i] == i && cl[j] == j) {
        if (!valid[i] || !valid[j]) continue;
        nclusters--;
        if (sz[i] < sz[j]) { cl[i] = j; sz[j] += sz[i]; if (sz[j] >= cs) { valid[j] = 0; nclusters--; } }
        else { cl[j] = i; sz[i] += sz[j]; if (sz[i] >= cs) { valid[i] = 0; nclusters--; } }
      } else {
        int64_t ri = uf_find(cl, i), rj = uf_find(cl, j);
        if (!valid[ri] || !valid[rj]) continue;
        if (ri != rj) {
          int64_t a = ri < rj ? ri : rj, b = ri < rj ? rj : ri;
          if (!initial_has(np, pi, pj, a, b) && pset_insert(&seen, a, b)) {
            item_t it = {as_value(orc_jaccard(rowptr, col, a, b, su), score_f32), a, b};
            heap_push(&dyn, it);
          }
        }
      }
    }
    free(ord); free(oi); free(oj);
  }
  (void)pset_has;
  int64_t nc = hc_output(n, cl, perm, offsets);
  free(dyn.h); free(seen.k); free(cl); free(sz); free(valid);
  return nc;
}

/* ------------------------------------------------------------------------------------ */
/* Cluster format and cluster-wise SpGEMM                                                 */
/* ------------------------------------------------------------------------------------ */

typedef struct { int64_t col, pos; } colpos_t;
static int cmp_colpos(const void *a, const void *b) {
  const colpos_t *x = (const colpos_t *)a, *y = (const colpos_t *)b;
  if (x->col != y->col) return x->col < y->col ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* csr_reorder_rows (include/sparse_conversion.h:1879-1976) followed by
 * csr_to_vlength_cluster (:1641-1765).  Cluster c holds reordered rows
 * [offsets[c], offsets[c+1]); colids are its distinct columns and
 * values[rowptr_val[c] + t*csz + l] = A'(offsets[c]+l, colids[t]) or 0 (a duplicate
 * entry overwrites an earlier one, as the reference's map assignment does).
 * Canonical column order: ascending.  The reference iterates a std::unordered_map here
 * (:1738-1753); that order changes only the FP64 summation order of the product, not
 * the structure of C.  Returns nnzc; *colids / *values malloc'ed. */
int64_t orc_cluster_build(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                          const int64_t *perm, int64_t nc, const int64_t *offsets,
                          int64_t *cl_rowptr, int64_t *cl_rowptr_val, int64_t *cl_sz,
                          int64_t **colids, double **values) {
  int64_t nnzc = 0, nnzv = 0, maxent = 0;
  (void)n_rows;
  for (int64_t c = 0; c < nc; ++c) {
    int64_t e = 0;
    for (int64_t t = offsets[c]; t < offsets[c + 1]; ++t) e += rowptr[perm[t] + 1] - rowptr[perm[t]];
    if (e > maxent) maxent = e;
  }
  colpos_t *buf = (colpos_t *)xmalloc(sizeof(colpos_t) * (size_t)(maxent ? maxent : 1));
  int64_t *ucount = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nc ? nc : 1));
  for (int64_t c = 0; c < nc; ++c) {
    int64_t e = 0;
    for (int64_t t = offsets[c]; t < offsets[c + 1]; ++t)
      for (int64_t j = rowptr[perm[t]]; j < rowptr[perm[t] + 1]; ++j) { buf[e].col = col[j]; buf[e].pos = e; ++e; }
    qsort(buf, (size_t)e, sizeof(colpos_t), cmp_colpos);
    int64_t u = 0;
    for (int64_t q = 0; q < e; ++q) if (q == 0 || buf[q].col != buf[q - 1].col) ++u;
    ucount[c] = u;
    cl_sz[c] = offsets[c + 1] - offsets[c];
    cl_rowptr[c] = nnzc;
    cl_rowptr_val[c] = nnzv;
    nnzc += u;
    nnzv += u * cl_sz[c];
  }
  cl_rowptr[nc] = nnzc;
  cl_rowptr_val[nc] = nnzv;
  *colids = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)nnzc);
  *values = (double *)calloc((size_t)(nnzv ? nnzv : 1), sizeof(double));
  for (int64_t c = 0; c < nc; ++c) {
    int64_t e = 0;
    for (int64_t t = offsets[c]; t < offsets[c + 1]; ++t)
      for (int64_t j = rowptr[perm[t]]; j < rowptr[perm[t] + 1]; ++j) { buf[e].col = col[j]; buf[e].pos = e; ++e; }
    qsort(buf, (size_t)e, sizeof(colpos_t), cmp_colpos);
    int64_t u = 0;
    for (int64_t q = 0; q < e; ++q) if (q == 0 || buf[q].col != buf[q - 1].col) (*colids)[cl_rowptr[c] + u++] = buf[q].col;
    /* values: walk rows in order so later duplicates overwrite */
    for (int64_t t = offsets[c]; t < offsets[c + 1]; ++t) {
      int64_t l = t - offsets[c];
      for (int64_t j = rowptr[perm[t]]; j < rowptr[perm[t] + 1]; ++j) {
        int64_t lo = cl_rowptr[c], hi = cl_rowptr[c + 1];
        while (lo < hi) { int64_t mid = (lo + hi) / 2; if ((*colids)[mid] < col[j]) lo = mid + 1; else hi = mid; }
        (*values)[cl_rowptr_val[c] + (lo - cl_rowptr[c]) * cl_sz[c] + l] = val[j];
      }
    }
  }
  free(buf); free(ucount);
  return nnzc;
}

/* src/spgemm.cpp:24-42 (get_spgemm_flop): 2 * sum over A entries (i,k), k < B.rows, of nnz(B_k). */
int64_t orc_spgemm_flop(int64_t a_rows, const int64_t *arowptr, const int64_t *acol,
                        int64_t b_rows, const int64_t *browptr) {
  int64_t f = 0;
  for (int64_t i = 0; i < a_rows; ++i)
    for (int64_t j = arowptr[i]; j < arowptr[i + 1]; ++j)
      if (acol[j] < b_rows) f += browptr[acol[j] + 1] - browptr[acol[j]];
  return 2 * f;
}

/* src/spgemm_Vlength_mixed.cpp:197-261 (spgemm_mixed_symbolic_omp_lb): per cluster, the
 * number of distinct B columns reached, with their min and max (0,0 when none). */
void orc_cluster_symbolic(int64_t nc, const int64_t *cl_rowptr, const int64_t *colids,
                          const int64_t *browptr, const int64_t *bcol, int64_t b_cols,
                          int64_t *c_nz, int64_t *c_min, int64_t *c_max) {
  int64_t *stamp = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(b_cols ? b_cols : 1));
  for (int64_t t = 0; t < b_cols; ++t) stamp[t] = -1;
  for (int64_t c = 0; c < nc; ++c) {
    int64_t nz = 0, mn = INT64_MAX, mx = 0;
    for (int64_t j = cl_rowptr[c]; j < cl_rowptr[c + 1]; ++j) {
      int64_t a = colids[j];
      for (int64_t q = browptr[a]; q < browptr[a + 1]; ++q) {
        int64_t key = bcol[q];
        if (stamp[key] != c) {
          stamp[key] = c;
          ++nz;
          if (key < mn) mn = key;
          if (key > mx) mx = key;
        }
      }
    }
    c_nz[c] = nz;
    c_min[c] = nz ? mn : 0;
    c_max[c] = nz ? mx : 0;
  }
  free(stamp);
}

/* src/spgemm_Vlength_mixed.cpp:268-319 (classify_clusters): dense (1) iff
 * acols*B_rowdense*(8+value_bytes) + col_range*csz*value_bytes <= L2 budget and
 * nz/col_range >= density threshold.  Returns the dense count. */
int64_t orc_cluster_classify(int64_t nc, const int64_t *cl_rowptr, const int64_t *cl_sz,
                             const int64_t *c_nz, const int64_t *c_min, const int64_t *c_max,
                             double b_rowdense, int64_t l2_budget, double thr,
                             int64_t vbytes, int8_t *acc_flag) {
  int64_t dense = 0;
  for (int64_t i = 0; i < nc; ++i) {
    acc_flag[i] = 0;
    if (c_nz[i] == 0) continue;
    int64_t range = c_max[i] - c_min[i] + 1;
    double density = (double)c_nz[i] / (double)range;
    size_t bbytes = (size_t)((double)(cl_rowptr[i + 1] - cl_rowptr[i]) * b_rowdense * (double)(8 + vbytes));
    size_t dbytes = (size_t)range * (size_t)cl_sz[i] * (size_t)vbytes;
    if (bbytes + dbytes <= (size_t)l2_budget && density >= thr) { acc_flag[i] = 1; ++dense; }
  }
  return dense;
}

/* src/spgemm_Vlength_mixed.cpp:326-434 (spgemm_mixed_numeric_omp_lb), dense-accumulator
 * semantics for every cluster: slot = 0, then for each A-cluster column t (colids order)
 * and each entry of B row colids[t] (B order): slot[bcol][l] += a[t][l] * b.  Output
 * columns ascending (sortOutput), values padded to csz per column.  use_long_double
 * accumulates in extended precision (error reference). */
void orc_cluster_numeric(int64_t nc, const int64_t *cl_rowptr, const int64_t *cl_rowptr_val,
                         const int64_t *cl_sz, const int64_t *colids, const double *values,
                         const int64_t *browptr, const int64_t *bcol, const double *bval,
                         const int64_t *c_rowptr, const int64_t *c_rowptr_val,
                         int64_t *c_colids, double *c_values, int use_ld) {
  for (int64_t c = 0; c < nc; ++c) {
    int64_t csz = cl_sz[c], nz = c_rowptr[c + 1] - c_rowptr[c];
    if (nz == 0) continue;
    /* gather distinct columns (sorted) */
    int64_t cap = 0;
    for (int64_t j = cl_rowptr[c]; j < cl_rowptr[c + 1]; ++j) cap += browptr[colids[j] + 1] - browptr[colids[j]];
    int64_t *cols = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)cap);
    int64_t m = 0;
    for (int64_t j = cl_rowptr[c]; j < cl_rowptr[c + 1]; ++j)
      for (int64_t q = browptr[colids[j]]; q < browptr[colids[j] + 1]; ++q) cols[m++] = bcol[q];
    colpos_t *cp = (colpos_t *)xmalloc(sizeof(colpos_t) * (size_t)(m ? m : 1));
    for (int64_t q = 0; q < m; ++q) { cp[q].col = cols[q]; cp[q].pos = q; }
    qsort(cp, (size_t)m, sizeof(colpos_t), cmp_colpos);
    int64_t u = 0;
    for (int64_t q = 0; q < m; ++q) if (q == 0 || cp[q].col != cp[q - 1].col) c_colids[c_rowptr[c] + u++] = cp[q].col;
    free(cp);
    free(cols);
    long double *acc = (long double *)calloc((size_t)(nz * csz), sizeof(long double));
    double *accd = (double *)calloc((size_t)(nz * csz), sizeof(double));
    const int64_t *ccol = c_colids + c_rowptr[c];
    for (int64_t j = cl_rowptr[c]; j < cl_rowptr[c + 1]; ++j) {
      const double *a = values + cl_rowptr_val[c] + (j - cl_rowptr[c]) * csz;
      int64_t bt = colids[j];
      for (int64_t q = browptr[bt]; q < browptr[bt + 1]; ++q) {
        int64_t lo = 0, hi = nz;
        while (lo < hi) { int64_t mid = (lo + hi) / 2; if (ccol[mid] < bcol[q]) lo = mid + 1; else hi = mid; }
        for (int64_t l = 0; l < csz; ++l) {
          if (use_ld) acc[lo * csz + l] += (long double)a[l] * (long double)bval[q];
          else accd[lo * csz + l] += a[l] * bval[q];
        }
      }
    }
    for (int64_t s = 0; s < nz * csz; ++s)
      c_values[c_rowptr_val[c] + s] = use_ld ? (double)acc[s] : accd[s];
    free(acc);
    free(accd);
  }
}

/* vlength_cluster2csr (include/sparse_conversion.h:1778-1862: drops |v| == 0) followed by
 * the inverse-permutation reorder of test/test_spgemm_hc_lsh.cpp:299-303.  Output rows in
 * original order, columns in cluster column order (ascending here).  Returns nnz. */
int64_t orc_cluster_to_csr(int64_t n_rows, int64_t nc, const int64_t *cl_sz, const int64_t *perm,
                           const int64_t *c_rowptr, const int64_t *c_rowptr_val,
                           const int64_t *c_colids, const double *c_values,
                           int64_t *rowptr_out, int64_t **col_out, double **val_out) {
  int64_t *cnt = (int64_t *)calloc((size_t)(n_rows + 1), sizeof(int64_t));
  int64_t row = 0;
  for (int64_t c = 0; c < nc; ++c) {
    for (int64_t j = c_rowptr[c]; j < c_rowptr[c + 1]; ++j)
      for (int64_t l = 0; l < cl_sz[c]; ++l)
        if (fabs(c_values[c_rowptr_val[c] + (j - c_rowptr[c]) * cl_sz[c] + l]) > 0.0) cnt[perm[row + l]]++;
    row += cl_sz[c];
  }
  rowptr_out[0] = 0;
  for (int64_t r = 0; r < n_rows; ++r) rowptr_out[r + 1] = rowptr_out[r] + cnt[r];
  int64_t nnz = rowptr_out[n_rows];
  *col_out = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)nnz);
  *val_out = (double *)xmalloc(sizeof(double) * (size_t)nnz);
  for (int64_t r = 0; r < n_rows; ++r) cnt[r] = rowptr_out[r];
  row = 0;
  for (int64_t c = 0; c < nc; ++c) {
    for (int64_t j = c_rowptr[c]; j < c_rowptr[c + 1]; ++j)
      for (int64_t l = 0; l < cl_sz[c]; ++l) {
        double v = c_values[c_rowptr_val[c] + (j - c_rowptr[c]) * cl_sz[c] + l];
        if (fabs(v) > 0.0) {
          int64_t o = perm[row + l], p = cnt[o]++;
          (*col_out)[p] = c_colids[j];
          (*val_out)[p] = v;
        }
      }
    row += cl_sz[c];
  }
  free(cnt);
  return nnz;
}

/* ------------------------------------------------------------------------------------ */
/* AI forms: SpMM, SDDMM, transpose                                                       */
/* ------------------------------------------------------------------------------------ */

/* Y[r,:] = sum_j a_rj X[col_j,:]  (per-matrix semantics of SpOps_csr applied column-wise,
 * alpha=1 beta=0); X is (n_cols x k) row-major, Y (n_rows x k). */
void orc_spmm_csr(int64_t n_rows, const int64_t *rowptr, const int64_t *col, const double *val,
                  int64_t k, const double *X, double *Y) {
  for (int64_t r = 0; r < n_rows; ++r) {
    for (int64_t q = 0; q < k; ++q) Y[r * k + q] = 0.0;
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
      for (int64_t q = 0; q < k; ++q) Y[r * k + q] += val[j] * X[col[j] * k + q];
  }
}

/* dvals[e] = <G[row_e,:], X[col_e,:]> on A's pattern. */
void orc_sddmm_csr(int64_t n_rows, const int64_t *rowptr, const int64_t *col, int64_t k,
                   const double *G, const double *X, double *dvals) {
  for (int64_t r = 0; r < n_rows; ++r)
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      double s = 0.0;
      for (int64_t q = 0; q < k; ++q) s += G[r * k + q] * X[col[j] * k + q];
      dvals[j] = s;
    }
}

/* Pattern-preserving transpose (a counting sort by column; within a transposed row the
 * original rows ascend).  t_perm[e'] = index of the original entry.  The reference's
 * csr_transpose (include/sparse_conversion.h:2067-2108) additionally sums duplicates and
 * drops exact zeros; inputs here have neither. */
void orc_csr_transpose(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col,
                       const double *val, int64_t *t_rowptr, int64_t *t_col, double *t_val, int64_t *t_perm) {
  for (int64_t c = 0; c <= n_cols; ++c) t_rowptr[c] = 0;
  for (int64_t j = 0; j < rowptr[n_rows]; ++j) t_rowptr[col[j] + 1]++;
  for (int64_t c = 0; c < n_cols; ++c) t_rowptr[c + 1] += t_rowptr[c];
  int64_t *cur = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n_cols ? n_cols : 1));
  for (int64_t c = 0; c < n_cols; ++c) cur[c] = t_rowptr[c];
  for (int64_t r = 0; r < n_rows; ++r)
    for (int64_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
      int64_t p = cur[col[j]]++;
      t_col[p] = r;
      t_val[p] = val[j];
      t_perm[p] = j;
    }
  free(cur);
}

/* ---- matrix features ------------------------------------------------------------------------
 * The reference's feature extractor for a general real matrix given as CSR, entries taken in
 * CSR order (the order MtxLoad reads a file written from that CSR):
 *   per entry     src/sparse_features.cpp:336-467 (general branch): row / column counts,
 *                 diagonal offsets, 256 x 256 tiles with the uneven split of :186-190 and
 *                 :357-358, first occurrence of a row in a tile column and of a column in a
 *                 tile row (Rows_flag / Cols_flag), Diag_Dom, lower / upper / diagonal counts,
 *                 largest |value| off and on the diagonal, log10 extremes per row / column
 *                 starting from 0.0;
 *   statistics    CalculateFeatures :643-760, CalculateTilesFeatures :763-850 (tile variance
 *                 and nonzero ratio start from -1.0), P_ratioAndGini :609-640,
 *                 trapezoidalRule :89-100.
 * The output order is the library's (spops_matrix_features).
 * Parity: pinned against the reference's own MTX::FeaturesPrint output on seeded matrices
 * (tests/golden/features_*.npz, printed to 6 significant digits) and, exactly, against the
 * device implementation on larger ones. */
static int cmp_i64(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

static void orc_p_ratio_gini(const int64_t *vec, int64_t length, int64_t nnz, double *p_ratio, double *gini) {
  int64_t *v = (int64_t *)malloc(sizeof(int64_t) * (size_t)length);
  memcpy(v, vec, sizeof(int64_t) * (size_t)length);
  qsort(v, (size_t)length, sizeof(int64_t), cmp_i64);
  int64_t quit = length, step = 0, count = 0;
  double p = 0.0, q = 0.0;
  do {
    --quit;
    ++count;
    step += v[quit];
    p = (double)count / (double)length;
    q = (double)step / (double)nnz;
  } while (p + q < 1.0);
  *p_ratio = p;
  /* cumulative sums, trapezoidal area with unit steps, area under the diagonal */
  double area = 0.0, x_step = (double)length / (double)length;
  int64_t cum = v[0], prev;
  area += (double)(cum + 0) / 2.0;
  for (int64_t i = 0; i + 1 < length; ++i) {
    prev = cum;
    cum += v[i + 1];
    area += (double)(prev + cum) / 2.0 * x_step;
  }
  const double total = (double)nnz * (double)length / 2.0;
  *gini = (total - area) / total;
  free(v);
}

static void orc_count_stats(const int64_t *n, int64_t len, int64_t nnz, double *o) {
  int64_t mn = 100000000, mx = 0, nz = 0;
  const double ave = (double)nnz / (double)len;
  double var = 0.0;
  for (int64_t i = 0; i < len; ++i) {
    if (n[i] < mn) mn = n[i];
    if (n[i] > mx) mx = n[i];
    if (n[i]) ++nz;
    const double d = (double)n[i] - ave;
    var += d * d;
  }
  var /= (double)len;
  o[0] = (double)nz / (double)len;
  o[1] = (double)mn;
  o[2] = (double)mx;
  o[3] = ave;
  o[4] = var;
  o[5] = sqrt(var);
  orc_p_ratio_gini(n, len, nnz, &o[6], &o[7]);
}

static void orc_tile_stats(const int64_t *n, int64_t len, int64_t nnz, double ave, double *o) {
  double nzr = -1.0, var = -1.0;
  int64_t mn = 100000000, mx = 0;
  for (int64_t i = 0; i < len; ++i) {
    if (n[i]) nzr += 1.0;
    const double d = (double)n[i] - ave;
    var += d * d;
    if (n[i] < mn) mn = n[i];
    if (n[i] > mx) mx = n[i];
  }
  nzr /= (double)len;
  var /= (double)len;
  o[0] = nzr;
  o[1] = (double)mn;
  o[2] = (double)mx;
  o[3] = ave;
  o[4] = var;
  o[5] = sqrt(var > 0.0 ? var : 0.0);
  orc_p_ratio_gini(n, len, nnz, &o[6], &o[7]);
}

void orc_features(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col, const double *val,
                  double *out) {
  enum { T = 256 };
  const int64_t nnz = rowptr[n_rows] - rowptr[0];
  const int64_t ndiag = n_rows + n_cols - 1, close_thr = ndiag / 20;
  const int64_t rb_num = n_rows / T, rb_mod = n_rows % T, rb_thr = rb_mod * (rb_num + 1);
  const int64_t cb_num = n_cols / T, cb_mod = n_cols % T, cb_thr = cb_mod * (cb_num + 1);
  int64_t *rowlen = (int64_t *)calloc((size_t)n_rows, 8), *collen = (int64_t *)calloc((size_t)n_cols, 8);
  int64_t *offcnt = (int64_t *)calloc((size_t)ndiag, 8);
  int64_t *tiles = (int64_t *)calloc(T * T, 8), *rbc = (int64_t *)calloc(T, 8), *cbc = (int64_t *)calloc(T, 8);
  int64_t *urb = (int64_t *)calloc(T * T, 8), *ucb = (int64_t *)calloc(T * T, 8);
  unsigned char *rflag = (unsigned char *)calloc((size_t)n_rows * T, 1), *cflag = (unsigned char *)calloc((size_t)n_cols * T, 1);
  double *dom = (double *)calloc((size_t)n_rows, 8);
  double *rmax = (double *)calloc((size_t)n_rows, 8), *rmin = (double *)calloc((size_t)n_rows, 8);
  double *cmax = (double *)calloc((size_t)n_cols, 8), *cmin = (double *)calloc((size_t)n_cols, 8);
  int64_t lower = 0, upper = 0, diag = 0;
  double maxoff = -1.7976931348623157e308, maxdiag = -1.7976931348623157e308;
  for (int64_t r = 0; r < n_rows; ++r)
    for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
      const int64_t c = col[p];
      const double a = fabs(val[p]);
      rowlen[r]++;
      collen[c]++;
      offcnt[c - r + n_rows - 1]++;
      const int64_t tr = r < rb_thr ? r / (rb_num + 1) : rb_mod + (r - rb_thr) / rb_num;
      const int64_t tc = c < cb_thr ? c / (cb_num + 1) : cb_mod + (c - cb_thr) / cb_num;
      tiles[tr * T + tc]++;
      rbc[tr]++;
      cbc[tc]++;
      if (!rflag[r * T + tc]) { urb[tr * T + tc]++; rflag[r * T + tc] = 1; }
      if (!cflag[c * T + tr]) { ucb[tr * T + tc]++; cflag[c * T + tr] = 1; }
      if (r == c) {
        diag++;
        maxdiag = maxdiag > a ? maxdiag : a;
        dom[r] += a;
      } else {
        if (r > c) lower++; else upper++;
        maxoff = maxoff > a ? maxoff : a;
        dom[r] -= a;
      }
      const double lg = log10(a);
      if (rmax[r] < lg) rmax[r] = lg;
      if (a > 0.0 && rmin[r] > lg) rmin[r] = lg;
      if (cmax[c] < lg) cmax[c] = lg;
      if (a > 0.0 && cmin[c] > lg) cmin[c] = lg;
    }
  int64_t ndiags = 0, close = 0;
  for (int64_t d = 0; d < ndiag; ++d)
    if (offcnt[d]) {
      ndiags++;
      const int64_t off = d - (n_rows - 1);
      if ((off < 0 ? -off : off) <= close_thr) close += offcnt[d];
    }
  out[0] = (double)n_rows;
  out[1] = (double)n_cols;
  out[2] = (double)nnz;
  out[3] = (double)nnz / ((double)n_rows * (double)n_cols);
  out[4] = (double)ndiags;
  out[5] = (double)close / (double)nnz;
  out[6] = (double)lower;
  out[7] = (double)upper;
  out[8] = (double)diag;
  orc_count_stats(rowlen, n_rows, nnz, out + 9);
  orc_count_stats(collen, n_cols, nnz, out + 17);
  int64_t dominance = 0;
  double rv = -1.0, cv = -1.0;
  for (int64_t r = 0; r < n_rows; ++r) {
    if (dom[r] > 0) dominance++;
    if (rmin[r] != 0.0) { const double t = rmax[r] - rmin[r]; rv = rv > t ? rv : t; }
  }
  for (int64_t c = 0; c < n_cols; ++c)
    if (cmin[c] != 0.0) { const double t = cmax[c] - cmin[c]; cv = cv > t ? cv : t; }
  out[25] = (double)dominance / (double)(n_rows < n_cols ? n_rows : n_cols);
  out[26] = maxoff;
  out[27] = maxdiag;
  out[28] = rv;
  out[29] = cv;
  out[30] = (double)T;
  out[31] = (double)rb_num;
  out[32] = (double)cb_num;
  orc_tile_stats(tiles, T * T, nnz, (double)nnz / ((double)T * (double)T), out + 33);
  orc_tile_stats(rbc, T, nnz, (double)nnz / (double)T, out + 41);
  orc_tile_stats(cbc, T, nnz, (double)nnz / (double)T, out + 49);
  int64_t su = 0, sv = 0;
  for (int i = 0; i < T * T; ++i) { su += urb[i]; sv += ucb[i]; }
  out[57] = (double)su / (double)nnz;
  out[58] = (double)sv / (double)nnz;
  out[59] = (double)su / (double)n_rows;
  out[60] = (double)sv / (double)n_cols;
  free(rowlen); free(collen); free(offcnt); free(tiles); free(rbc); free(cbc); free(urb); free(ucb);
  free(rflag); free(cflag); free(dom); free(rmax); free(rmin); free(cmax); free(cmin);
}

/* ------------------------------------------------------------------------------------ */
/* Row-wise SpGEMM and the A * A^T top-k Jaccard front end                                */
/* ------------------------------------------------------------------------------------ */

/* The row-wise LeSpGEMM (src/spgemm.cpp:684-705) as its SPA kernel computes it
 * (src/spgemm.cpp:175-236, spgemm_spa_numeric): per row, every product a_ik * b_kj added into
 * C(i, j) in the order of A's row entries, then B's; C's columns are the symbolic union, ascending,
 * zero sums kept.  *c_col / *c_val malloc'ed; c_rowptr has n_rows + 1 entries.
 * Pinned: scipy's product (structure exact, values 1e-12) and the reference's own output for
 * kernel flags 1, 2 and 3 (tests/golden/rowwise_*). */
int64_t orc_spgemm_rows(int64_t n_rows, const int64_t *arp, const int64_t *acol, const double *aval, int64_t b_cols,
                        const int64_t *brp, const int64_t *bcol, const double *bval, int64_t *c_rowptr,
                        int64_t **c_col, double **c_val) {
  int64_t *mark = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(b_cols ? b_cols : 1));
  double *acc = (double *)xmalloc(sizeof(double) * (size_t)(b_cols ? b_cols : 1));
  for (int64_t j = 0; j < b_cols; ++j) mark[j] = -1;
  c_rowptr[0] = 0;
  for (int64_t i = 0; i < n_rows; ++i) {           /* symbolic: distinct columns per row */
    int64_t nz = 0;
    for (int64_t e = arp[i]; e < arp[i + 1]; ++e)
      for (int64_t f = brp[acol[e]]; f < brp[acol[e] + 1]; ++f)
        if (mark[bcol[f]] != i) { mark[bcol[f]] = i; ++nz; }
    c_rowptr[i + 1] = c_rowptr[i] + nz;
  }
  const int64_t nnz = c_rowptr[n_rows];
  *c_col = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  *c_val = (double *)xmalloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  for (int64_t j = 0; j < b_cols; ++j) mark[j] = -1;
  for (int64_t i = 0; i < n_rows; ++i) {
    int64_t *cc = *c_col + c_rowptr[i];
    int64_t nz = 0;
    for (int64_t e = arp[i]; e < arp[i + 1]; ++e)
      for (int64_t f = brp[acol[e]]; f < brp[acol[e] + 1]; ++f) {
        const int64_t j = bcol[f];
        if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cc[nz++] = j; }
        acc[j] += aval[e] * bval[f];
      }
    qsort(cc, (size_t)nz, sizeof(int64_t), cmp_i64);
    for (int64_t t = 0; t < nz; ++t) (*c_val)[c_rowptr[i] + t] = acc[cc[t]];
  }
  free(mark);
  free(acc);
  return nnz;
}

typedef struct { double s; int64_t c; } cand_t;
static int cmp_cand(const void *x, const void *y) {   /* score descending, column ascending */
  const cand_t *a = (const cand_t *)x, *b = (const cand_t *)y;
  if (a->s != b->s) return a->s > b->s ? -1 : 1;
  return (a->c > b->c) - (a->c < b->c);
}

/* HashSpGEMMTopK (src/spgemm.cpp:707-842; selection src/spgemm_hash.cpp:99-108, 169-230): for
 * row i the rows j != i sharing a column, intersection |A_i & A_j| counted with multiplicity
 * (the product of the 0/1 patterns), union |A_i| + |A_j| - inter (stored entries), score
 * 1 * inter / union in the value type (float with score_f32), the top_k by score descending then
 * j ascending.  *c_col / *c_val malloc'ed; c_rowptr has n_rows + 1 entries.
 * Pinned: the reference's own HashSpGEMMTopK output (tests/golden/topk_*) and a brute force over
 * Python sets on small matrices. */
int64_t orc_topk_jaccard(int64_t n_rows, int64_t n_cols, const int64_t *rowptr, const int64_t *col, int64_t top_k,
                         int score_f32, int64_t *c_rowptr, int64_t **c_col, double **c_val) {
  const int64_t nnz = rowptr[n_rows];
  int64_t *tp = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n_cols + 1));
  int64_t *tr = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  for (int64_t c = 0; c <= n_cols; ++c) tp[c] = 0;
  for (int64_t e = 0; e < nnz; ++e) tp[col[e] + 1]++;
  for (int64_t c = 0; c < n_cols; ++c) tp[c + 1] += tp[c];
  int64_t *cur = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n_cols ? n_cols : 1));
  for (int64_t c = 0; c < n_cols; ++c) cur[c] = tp[c];
  for (int64_t i = 0; i < n_rows; ++i)
    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) tr[cur[col[e]]++] = i;
  double *cnt = (double *)xmalloc(sizeof(double) * (size_t)(n_rows ? n_rows : 1));
  int64_t *seen = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n_rows ? n_rows : 1));
  int64_t *list = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n_rows ? n_rows : 1));
  cand_t *cand = (cand_t *)xmalloc(sizeof(cand_t) * (size_t)(n_rows ? n_rows : 1));
  for (int64_t j = 0; j < n_rows; ++j) seen[j] = -1;
  int64_t cap = n_rows * (top_k < n_rows ? top_k : n_rows) + 1;
  *c_col = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)cap);
  *c_val = (double *)xmalloc(sizeof(double) * (size_t)cap);
  c_rowptr[0] = 0;
  for (int64_t i = 0; i < n_rows; ++i) {
    int64_t m = 0;
    for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
      for (int64_t f = tp[col[e]]; f < tp[col[e] + 1]; ++f) {
        const int64_t j = tr[f];
        if (seen[j] != i) { seen[j] = i; cnt[j] = 0.0; list[m++] = j; }
        cnt[j] += 1.0;
      }
    int64_t q = 0;
    const int64_t li = rowptr[i + 1] - rowptr[i];
    for (int64_t t = 0; t < m; ++t) {
      const int64_t j = list[t];
      if (j == i) continue;
      const int64_t u = li + (rowptr[j + 1] - rowptr[j]) - (int64_t)cnt[j];
      double sc;
      if (score_f32) sc = u == 0 ? 0.0 : (double)((float)1.0f * (float)cnt[j] / (float)u);
      else sc = u == 0 ? 0.0 : 1.0 * cnt[j] / (double)u;
      cand[q].s = sc;
      cand[q].c = j;
      ++q;
    }
    qsort(cand, (size_t)q, sizeof(cand_t), cmp_cand);
    const int64_t keep = q < top_k ? q : top_k;
    for (int64_t t = 0; t < keep; ++t) {
      (*c_col)[c_rowptr[i] + t] = cand[t].c;
      (*c_val)[c_rowptr[i] + t] = cand[t].s;
    }
    c_rowptr[i + 1] = c_rowptr[i] + keep;
  }
  free(tp); free(tr); free(cur); free(cnt); free(seen); free(list); free(cand);
  return c_rowptr[n_rows];
}

/* ------------------------------------------------------------------------------------ */
/* Approximate clustering by rounds of mutual best edges (no reference counterpart)      */
/* ------------------------------------------------------------------------------------ */

/* A bijective 64-bit mix of an edge's (min id, max id) key (the splitmix64 finalizer): the
 * tie-break among equal scores, a pseudo-random strict order that keeps chains of equal-score
 * pairs (the windowed LSH buckets of identical rows) from merging one pair per round. */
static uint64_t edge_rank_hash(int64_t lo, int64_t hi) {
  uint64_t x = (uint64_t)lo << 32 | (uint64_t)hi;
  x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27; x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

/* The GPU clustering of spops_cluster_matching, round by round: every cluster picks its best
 * valid edge (the two clusters of a pair fit cluster_size together; highest score, ties by the
 * smaller edge_rank_hash of its (min id, max id) key), clusters that picked each other merge into the smaller id, rows
 * take their cluster's new id; until no merge or max_rounds.  Clusters ordered by id (their
 * smallest row), rows ascending.  Returns the number of clusters; *offsets malloc'ed; *rounds =
 * rounds that merged.  Pinned: a pure-Python restatement on small graphs and the invariants
 * (permutation, sizes, every cluster connected by candidate pairs) in tests/test_cluster_matching_cpu.py. */
int64_t orc_cluster_matching(int64_t n, int64_t np, const int64_t *pi, const int64_t *pj, const double *score,
                             int64_t cs, int64_t max_rounds, int64_t *perm, int64_t **offsets, int64_t *rounds) {
  int64_t *label = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int64_t *size = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int64_t *parent = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  double *bs = (double *)xmalloc(sizeof(double) * (size_t)(n ? n : 1));
  int64_t *be_lo = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int64_t *be_hi = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  char *has = (char *)xmalloc((size_t)(n ? n : 1));
  uint64_t *bh = (uint64_t *)xmalloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) { label[i] = i; size[i] = 1; }
  *rounds = 0;
  for (int64_t r = 0; r < max_rounds; ++r) {
    for (int64_t c = 0; c < n; ++c) { has[c] = 0; parent[c] = c; }
    for (int64_t t = 0; t < np; ++t) {
      if (pi[t] < 0 || pi[t] >= n || pj[t] < 0 || pj[t] >= n) continue;
      const int64_t a = label[pi[t]], b = label[pj[t]];
      if (a == b || size[a] + size[b] > cs) continue;
      const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
      const uint64_t h = edge_rank_hash(lo, hi);
      const int64_t ends[2] = {a, b};
      for (int e = 0; e < 2; ++e) {
        const int64_t c = ends[e];
        if (!has[c] || score[t] > bs[c] || (score[t] == bs[c] && h < bh[c])) {
          has[c] = 1; bs[c] = score[t]; bh[c] = h; be_lo[c] = lo; be_hi[c] = hi;
        }
      }
    }
    int64_t merged = 0;
    for (int64_t c = 0; c < n; ++c) {
      if (!has[c] || c != be_hi[c]) continue;
      const int64_t lo = be_lo[c];
      if (!has[lo] || be_lo[lo] != lo || be_hi[lo] != c) continue;
      parent[c] = lo;
      ++merged;
    }
    for (int64_t c = 0; c < n; ++c)
      if (parent[c] != c) size[parent[c]] += size[c];
    for (int64_t i = 0; i < n; ++i) label[i] = parent[label[i]];
    if (merged == 0) break;
    ++*rounds;
  }
  /* clusters by id, rows ascending: a counting pass over the ids */
  int64_t *count = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n + 1));
  for (int64_t i = 0; i <= n; ++i) count[i] = 0;
  for (int64_t i = 0; i < n; ++i) count[label[i] + 1]++;
  int64_t nc = 0;
  for (int64_t i = 0; i < n; ++i) nc += count[i + 1] > 0;
  *offsets = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(nc + 1));
  int64_t *start = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int64_t pos = 0, k = 0;
  for (int64_t id = 0; id < n; ++id) {
    start[id] = pos;
    if (count[id + 1] > 0) { (*offsets)[k++] = pos; pos += count[id + 1]; }
  }
  (*offsets)[nc] = n;
  for (int64_t i = 0; i < n; ++i) perm[start[label[i]]++] = i;
  free(label); free(size); free(parent); free(bs); free(be_lo); free(be_hi); free(has); free(count); free(start);
  free(bh);
  return nc;
}
