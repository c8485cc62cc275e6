/* Host format conversions of the SpMV and LeSpGEMM paths with the reference's names,
 * signatures and output layouts (include/sparse_conversion.h):
 *   coo_to_csr :38, csr_to_ell :195, csr_to_sell_c_sigma :345, csr_to_sell_c_R :491,
 *   csr_to_bsr :777, csr_to_vlength_cluster :1641, vlength_cluster2csr :1778,
 *   csr_reorder_rows :1879/:1989, csr_transpose :2067.
 * csr_to_vlength_cluster runs on the GPU (spops_cluster_create) and returns the host
 * struct; the others are host code producing exactly the reference's arrays (the drop-in
 * SpOps_* functions then upload them once, see spops_dropin.h).  Error behaviour follows the
 * reference: sigma % C != 0 prints and exits(-1); bad offsets throw std::invalid_argument;
 * allocation failure throws std::bad_alloc. */
#ifndef SPOPS_DROPIN_SPARSE_CONVERSION_H
#define SPOPS_DROPIN_SPARSE_CONVERSION_H

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <new>
#include <numeric>
#include <stdexcept>
#include <utility>
#include <vector>

#include "sparse_format.h"
#include "spgemm_cluster.h"
#include "spops_dropin.h"

#include "general_config.h"
#include "plat_config.h"

template <typename T>
inline T *spops_checked(T *p) {
  if (!p) throw std::bad_alloc();
  return p;
}

/* The reference's allocation check (include/sparse_conversion.h:25-35): a null pointer
 * prints where it happened and throws std::bad_alloc. */
inline void checkAlloc(const void *ptr, const char *func, const char *file, int line) {
  if (ptr != nullptr) return;
  std::cerr << "Memory allocation failed in " << func << " (" << file << ":" << line << ")" << std::endl;
  throw std::bad_alloc();
}
#ifndef CHECK_ALLOC
#define CHECK_ALLOC(ptr) checkAlloc((ptr), __FUNCTION__, __FILE__, __LINE__)
#endif

/* Sum duplicate entries row by row and drop exact-zero sums; the columns of a row come out
 * in reverse order of their first appearance (the order the reference's linked list yields,
 * include/sparse_operation.h:21-73). */
template <class I, class V>
void sum_csr_duplicates(const I num_rows, const I num_cols, I *row_ptr, I *col_index, V *values) {
  std::vector<I> next(num_cols, (I)-1);
  std::vector<V> sums(num_cols, V(0));
  I nnz = 0, start = 0;
  for (I i = 0; i < num_rows; i++) {
    const I end = row_ptr[i + 1];
    I head = (I)-2;
    for (I jj = start; jj < end; jj++) {
      const I j = col_index[jj];
      sums[j] += values[jj];
      if (next[j] == (I)-1) { next[j] = head; head = j; }
    }
    while (head != (I)-2) {
      const I cur = head;
      head = next[cur];
      if (sums[cur] != V(0)) { col_index[nnz] = cur; values[nnz] = sums[cur]; nnz++; }
      next[cur] = (I)-1;
      sums[cur] = V(0);
    }
    start = end;
    row_ptr[i + 1] = nnz;
  }
}

template <class I, class V>
CSR_Matrix<I, V> coo_to_csr(const COO_Matrix<I, V> &coo, bool compact = false) {
  CSR_Matrix<I, V> csr;
  csr.num_rows = coo.num_rows;
  csr.num_cols = coo.num_cols;
  csr.num_nnzs = coo.num_nnzs;
  csr.tag = 0;
  csr.kernel_flag = 0;
  csr.partition = nullptr;
  csr.row_offset = spops_checked(new_array<I>(csr.num_rows + 1));
  csr.col_index = spops_checked(new_array<I>(csr.num_nnzs));
  csr.values = spops_checked(new_array<V>(csr.num_nnzs));
  std::fill_n(csr.row_offset, csr.num_rows + 1, (I)0);
  for (I i = 0; i < coo.num_nnzs; i++) csr.row_offset[coo.row_index[i] + 1]++;
  for (I i = 0; i < csr.num_rows; i++) csr.row_offset[i + 1] += csr.row_offset[i];
  std::vector<I> fill(csr.row_offset, csr.row_offset + csr.num_rows);
  for (I i = 0; i < coo.num_nnzs; i++) {      // row-stable: file order kept inside a row
    const I d = fill[coo.row_index[i]]++;
    csr.col_index[d] = coo.col_index[i];
    csr.values[d] = coo.values[i];
  }
  if (compact) {
    sum_csr_duplicates(csr.num_rows, csr.num_cols, csr.row_offset, csr.col_index, csr.values);
    csr.num_nnzs = csr.row_offset[csr.num_rows];
  }
  return csr;
}

template <class I, class V>
ELL_Matrix<I, V> csr_to_ell(const CSR_Matrix<I, V> &csr, const LeadingDimension ld = RowMajor) {
  ELL_Matrix<I, V> ell;
  ell.num_rows = csr.num_rows;
  ell.num_cols = csr.num_cols;
  ell.num_nnzs = csr.num_nnzs;
  ell.tag = 0;
  ell.kernel_flag = 0;
  ell.partition = nullptr;
  ell.ld = ld;
  I w = 0, wmin = csr.num_rows ? csr.num_cols : 0;
  for (I i = 0; i < csr.num_rows; i++) {
    w = std::max(w, csr.row_offset[i + 1] - csr.row_offset[i]);
    wmin = std::min(wmin, csr.row_offset[i + 1] - csr.row_offset[i]);
  }
  ell.max_row_width = w;
  ell.min_row_width = wmin;
  const size_t slots = (size_t)ell.num_rows * (size_t)w;
  ell.col_index = spops_checked(new_array<I>(slots));
  ell.values = spops_checked(new_array<V>(slots));
  std::fill_n(ell.col_index, slots, (I)-1);
  std::fill_n(ell.values, slots, V(0));
  for (I r = 0; r < csr.num_rows; r++)
    for (I j = csr.row_offset[r]; j < csr.row_offset[r + 1]; j++) {
      const size_t t = (size_t)(j - csr.row_offset[r]);
      const size_t pos = ld == ColMajor ? (size_t)r + t * (size_t)ell.num_rows : (size_t)r * (size_t)w + t;
      ell.col_index[pos] = csr.col_index[j];
      ell.values[pos] = csr.values[j];
    }
  return ell;
}

namespace spops_dropin {
// rows of each sigma-slice sorted by (nnz, row) descending
template <class I>
std::vector<I> sell_order(const I *rp, I n, I sigma) {
  std::vector<I> perm(n);
  std::iota(perm.begin(), perm.end(), (I)0);
  for (I lo = 0; lo < n; lo += sigma) {
    const I hi = std::min<I>(n, lo + sigma);
    std::sort(perm.begin() + lo, perm.begin() + hi, [&](I a, I b) {
      const I la = rp[a + 1] - rp[a], lb = rp[b + 1] - rp[b];
      return la != lb ? la > lb : a > b;
    });
  }
  return perm;
}

template <class I, class V>
void sell_chunks(const CSR_Matrix<I, V> &csr, const std::vector<I> &perm, I C, I align, I nchunks, I *&chunk_len,
                 I **&col_index, V **&values) {
  chunk_len = spops_checked(new_array<I>(nchunks));
  std::fill_n(chunk_len, nchunks, (I)0);
  for (I g = 0; g < csr.num_rows; g++)
    chunk_len[g / C] = std::max(chunk_len[g / C], csr.row_offset[perm[g] + 1] - csr.row_offset[perm[g]]);
  for (I c = 0; c < nchunks; c++) chunk_len[c] = (chunk_len[c] + align - 1) / align * align;
  col_index = new I *[nchunks];
  values = new V *[nchunks];
  for (I c = 0; c < nchunks; c++) {
    const size_t e = (size_t)chunk_len[c] * (size_t)C;
    col_index[c] = spops_checked(new_array<I>(e));
    values[c] = spops_checked(new_array<V>(e));
    std::fill_n(col_index[c], e, (I)-1);
    std::fill_n(values[c], e, V(0));
  }
  for (I g = 0; g < csr.num_rows; g++) {
    const I r = perm[g], c = g / C, l = g % C;
    for (I j = csr.row_offset[r]; j < csr.row_offset[r + 1]; j++) {
      const size_t pos = (size_t)l * (size_t)chunk_len[c] + (size_t)(j - csr.row_offset[r]);
      col_index[c][pos] = csr.col_index[j];
      values[c][pos] = csr.values[j];
    }
  }
}
}  // namespace spops_dropin

template <class I, class V>
SELL_C_Sigma_Matrix<I, V> csr_to_sell_c_sigma(const CSR_Matrix<I, V> &csr, FILE *fp_feature, const int slicewidth = SELL_SIGMA,
                                              const int chunkwidth = CHUNK_SIZE,
                                              const I alignment = (SIMD_WIDTH / 8 / sizeof(V))) {
  (void)fp_feature;
  SELL_C_Sigma_Matrix<I, V> s;
  s.num_rows = csr.num_rows;
  s.num_cols = csr.num_cols;
  s.num_nnzs = csr.num_nnzs;
  s.tag = 0;
  s.kernel_flag = 0;
  s.partition = nullptr;
  s.sliceWidth_Sigma = slicewidth;
  s.chunkWidth_C = chunkwidth;
  s.alignment = alignment;
  s.sliceNum = (s.num_rows + slicewidth - 1) / slicewidth;
  s.validchunkNum = (s.num_rows + chunkwidth - 1) / chunkwidth;
  if (slicewidth % chunkwidth) {
    std::printf(" Sigma must be multiple divided by c\n");
    std::exit(-1);
  }
  s.chunk_num_per_slice = slicewidth / chunkwidth;
  s.chunkNum = s.sliceNum * s.chunk_num_per_slice;
  std::vector<I> perm = spops_dropin::sell_order<I>(csr.row_offset, csr.num_rows, (I)slicewidth);
  s.reorder = spops_checked(new_array<I>(s.num_rows));
  std::copy(perm.begin(), perm.end(), s.reorder);
  spops_dropin::sell_chunks(csr, perm, (I)chunkwidth, alignment, s.validchunkNum, s.chunk_len, s.col_index, s.values);
  return s;
}

template <class I, class V>
SELL_C_R_Matrix<I, V> csr_to_sell_c_R(const CSR_Matrix<I, V> &csr, FILE *fp_feature, const int chunkwidth = CHUNK_SIZE,
                                      const I alignment = (SIMD_WIDTH / 8 / sizeof(V))) {
  (void)fp_feature;
  SELL_C_R_Matrix<I, V> s;
  s.num_rows = csr.num_rows;
  s.num_cols = csr.num_cols;
  s.num_nnzs = csr.num_nnzs;
  s.tag = 0;
  s.kernel_flag = 0;
  s.partition = nullptr;
  s.chunkWidth_C = chunkwidth;
  s.alignment = alignment;
  s.validchunkNum = (s.num_rows + chunkwidth - 1) / chunkwidth;
  std::vector<I> perm = spops_dropin::sell_order<I>(csr.row_offset, csr.num_rows, std::max<I>(csr.num_rows, 1));
  s.reorder = spops_checked(new_array<I>(s.num_rows));
  std::copy(perm.begin(), perm.end(), s.reorder);
  spops_dropin::sell_chunks(csr, perm, (I)chunkwidth, alignment, s.validchunkNum, s.chunk_len, s.col_index, s.values);
  return s;
}

template <class I, class V>
BSR_Matrix<I, V> csr_to_bsr(const CSR_Matrix<I, V> &csr, const I blockDimRow = BSR_BlockDimRow,
                            I blockDimCol = (SIMD_WIDTH / 8 / sizeof(V))) {
  BSR_Matrix<I, V> b;
  b.num_rows = csr.num_rows;
  b.num_cols = csr.num_cols;
  b.num_nnzs = csr.num_nnzs;
  b.tag = 0;
  b.kernel_flag = 0;
  b.partition = nullptr;
  b.blockDim_r = blockDimRow;
  b.blockDim_c = blockDimCol;
  b.blockNNZ = blockDimRow * blockDimCol;
  b.mb = (b.num_rows + blockDimRow - 1) / blockDimRow;
  b.nb = (b.num_cols + blockDimCol - 1) / blockDimCol;
  b.row_ptr = spops_checked(new_array<I>(b.mb + 1));
  std::vector<I> cols;
  std::vector<I> mark(b.nb, (I)-1);
  b.row_ptr[0] = 0;
  for (I i = 0; i < b.mb; i++) {
    const I hi = std::min<I>(b.num_rows, (i + 1) * blockDimRow);
    std::vector<I> mine;
    for (I j = csr.row_offset[i * blockDimRow]; j < csr.row_offset[hi]; j++) {
      const I bc = csr.col_index[j] / blockDimCol;
      if (mark[bc] != i) { mark[bc] = i; mine.push_back(bc); }
    }
    std::sort(mine.begin(), mine.end());       // block columns ascending per block row
    cols.insert(cols.end(), mine.begin(), mine.end());
    b.row_ptr[i + 1] = (I)cols.size();
  }
  b.nnzb = (I)cols.size();
  b.block_colindex = spops_checked(new_array<I>(b.nnzb));
  std::copy(cols.begin(), cols.end(), b.block_colindex);
  const size_t slots = (size_t)b.nnzb * (size_t)b.blockNNZ;
  b.block_data = spops_checked(new_array<V>(slots));
  std::fill_n(b.block_data, slots, V(0));
  for (I r = 0; r < b.num_rows; r++) {
    const I br = r / blockDimRow;
    for (I j = csr.row_offset[r]; j < csr.row_offset[r + 1]; j++) {
      const I bc = csr.col_index[j] / blockDimCol;
      const I k = (I)(std::lower_bound(b.block_colindex + b.row_ptr[br], b.block_colindex + b.row_ptr[br + 1], bc) -
                      b.block_colindex);
      b.block_data[(size_t)k * b.blockNNZ + (size_t)(r % blockDimRow) * blockDimCol + csr.col_index[j] % blockDimCol] =
          csr.values[j];
    }
  }
  return b;
}

/* new row i = original row row_permutation[i] */
template <typename I, typename V>
CSR_Matrix<I, V> csr_reorder_rows(const CSR_Matrix<I, V> &csr, const I *row_permutation) {
  CSR_Matrix<I, V> r;
  r.num_rows = csr.num_rows;
  r.num_cols = csr.num_cols;
  r.tag = 0;
  r.kernel_flag = 0;
  r.partition = nullptr;
  r.row_offset = spops_checked(new_array<I>(r.num_rows + 1));
  r.row_offset[0] = 0;
  for (I i = 0; i < r.num_rows; i++) {
    const I o = row_permutation[i];
    const I len = (o >= 0 && o < csr.num_rows) ? csr.row_offset[o + 1] - csr.row_offset[o] : 0;
    r.row_offset[i + 1] = r.row_offset[i] + len;
  }
  r.num_nnzs = r.row_offset[r.num_rows];
  r.col_index = r.num_nnzs ? spops_checked(new_array<I>(r.num_nnzs)) : nullptr;
  r.values = r.num_nnzs ? spops_checked(new_array<V>(r.num_nnzs)) : nullptr;
  for (I i = 0; i < r.num_rows; i++) {
    const I o = row_permutation[i];
    if (o < 0 || o >= csr.num_rows) continue;
    std::copy(csr.col_index + csr.row_offset[o], csr.col_index + csr.row_offset[o + 1], r.col_index + r.row_offset[i]);
    std::copy(csr.values + csr.row_offset[o], csr.values + csr.row_offset[o + 1], r.values + r.row_offset[i]);
  }
  return r;
}

template <typename I, typename V>
CSR_Matrix<I, V> csr_reorder_rows(const CSR_Matrix<I, V> &csr, const std::vector<I> &row_permutation) {
  if ((I)row_permutation.size() < csr.num_rows) throw std::invalid_argument("permutation shorter than num_rows");
  return csr_reorder_rows(csr, row_permutation.data());
}

/* A row permutation file (include/sparse_conversion.h:2012-2052): one 0-based row per line,
 * or a Matrix Market file whose header and size line are skipped.  With num_rows > 0 the
 * list is padded with (size, size+1, ...) or cut to num_rows, as the reference does. */
template <typename I>
std::vector<I> read_row_permutation(const char *filename, I num_rows = 0) {
  std::vector<I> rows;
  std::ifstream file(filename);
  if (!file.is_open()) {
    std::cerr << "Error: Cannot open permutation file: " << filename << std::endl;
    throw std::runtime_error("Cannot open permutation file");
  }
  std::string first;
  std::getline(file, first);
  I v;
  if (first.find("%%MatrixMarket") != std::string::npos) {
    std::string line;
    while (std::getline(file, line))
      if (line.empty() || line[0] != '%') break;
  } else {
    file.clear();
    file.seekg(0, std::ios::beg);
  }
  while (file >> v) rows.push_back(v);
  if (num_rows > 0) {
    if ((I)rows.size() < num_rows)
      for (I i = (I)rows.size(); i < num_rows; ++i) rows.push_back(i);
    else if ((I)rows.size() > num_rows)
      rows.resize(num_rows);
  }
  return rows;
}

/* Cluster format of an already reordered CSR (offsets delimit the clusters).  Built on the
 * GPU; the columns of a cluster come out ascending (the reference iterates an
 * unordered_map, which changes only the summation order of a later product). */
template <class I, class V>
CSR_VlengthCluster<I, V> csr_to_vlength_cluster(const CSR_Matrix<I, V> &csr, const std::vector<I> &offsets) {
  using namespace spops_dropin;
  if (offsets.size() < 2) {
    std::cerr << "Error: offsets must have at least 2 elements (start and end)" << std::endl;
    throw std::invalid_argument("offsets size must be >= 2");
  }
  if (offsets[0] != 0 || offsets.back() != csr.num_rows) {
    std::cerr << "Error: offsets must start at 0 and end at csr.num_rows" << std::endl;
    throw std::invalid_argument("Invalid offsets: must be [0, ..., csr.num_rows]");
  }
  std::lock_guard<std::mutex> g(state().mu);
  const int64_t n = csr.num_rows, nnz = csr.num_nnzs, nc = (int64_t)offsets.size() - 1;
  const int64_t *drp = upload_index64(0, csr.row_offset, n + 1);
  const int64_t *dcol = upload_index64(1, csr.col_index, nnz);
  void *dval = state().scratch[2].get((size_t)(nnz ? nnz : 1) * sizeof(V));
  check(spops_memcpy_htod(dval, csr.values, (size_t)nnz * sizeof(V), nullptr), "upload values");
  std::vector<int64_t> perm(n), off(offsets.begin(), offsets.end());
  std::iota(perm.begin(), perm.end(), (int64_t)0);
  spops_cluster h = nullptr;
  check(spops_cluster_create(&h, n, csr.num_cols, nnz, drp, dcol, dval, value_code<V>(), perm.data(), nc, off.data(),
                             nullptr),
        "spops_cluster_create");
  int64_t info[8];
  spops_cluster_info(h, info);
  CSR_VlengthCluster<I, V> c;
  c.csr_rows = csr.num_rows;
  c.rows = (I)nc;
  c.cols = csr.num_cols;
  c.nnzc = (I)info[3];
  c.nnzv = (I)info[4];
  c.max_cluster_sz = std::max<I>(256, (I)info[5]);   // the reference starts the maximum at 256
  c.acc_flag = nullptr;
  std::vector<int64_t> csz(nc), rp(nc + 1), rpv(nc + 1), col(c.nnzc);
  c.values = spops_checked(new_array<V>(c.nnzv));
  check(spops_cluster_download(h, csz.data(), rp.data(), rpv.data(), col.data(), c.values), "download cluster");
  spops_cluster_destroy(h);
  c.cluster_sz = spops_checked(new_array<I>(nc));
  c.rowptr = spops_checked(new_array<I>(nc + 1));
  c.rowptr_val = spops_checked(new_array<I>(nc + 1));
  c.colids = spops_checked(new_array<I>(c.nnzc));
  std::copy(csz.begin(), csz.end(), c.cluster_sz);
  std::copy(rp.begin(), rp.end(), c.rowptr);
  std::copy(rpv.begin(), rpv.end(), c.rowptr_val);
  std::copy(col.begin(), col.end(), c.colids);
  c.num_rows = c.csr_rows;
  c.num_cols = c.cols;
  c.num_nnzs = c.nnzv;
  return c;
}

/* Back to CSR in cluster row order, dropping exact zeros (include/sparse_conversion.h:1778). */
template <class I, class V>
CSR_Matrix<I, V> vlength_cluster2csr(const CSR_VlengthCluster<I, V> &cl) {
  CSR_Matrix<I, V> csr;
  csr.num_rows = cl.csr_rows;
  csr.num_cols = cl.cols;
  csr.kernel_flag = 0;
  csr.tag = 0;
  csr.partition = nullptr;
  csr.row_offset = spops_checked(new_array<I>(csr.num_rows + 1));
  std::fill_n(csr.row_offset, csr.num_rows + 1, (I)0);
  I row = 0;
  for (I c = 0; c < cl.rows; c++) {
    for (I j = cl.rowptr[c]; j < cl.rowptr[c + 1]; j++)
      for (I l = 0; l < cl.cluster_sz[c]; l++)
        if (cl.values[cl.rowptr_val[c] + (j - cl.rowptr[c]) * cl.cluster_sz[c] + l] != V(0)) csr.row_offset[row + l + 1]++;
    row += cl.cluster_sz[c];
  }
  for (I r = 0; r < csr.num_rows; r++) csr.row_offset[r + 1] += csr.row_offset[r];
  csr.num_nnzs = csr.row_offset[csr.num_rows];
  csr.col_index = csr.num_nnzs ? spops_checked(new_array<I>(csr.num_nnzs)) : nullptr;
  csr.values = csr.num_nnzs ? spops_checked(new_array<V>(csr.num_nnzs)) : nullptr;
  std::vector<I> fill(csr.row_offset, csr.row_offset + csr.num_rows);
  row = 0;
  for (I c = 0; c < cl.rows; c++) {
    for (I j = cl.rowptr[c]; j < cl.rowptr[c + 1]; j++)
      for (I l = 0; l < cl.cluster_sz[c]; l++) {
        const V v = cl.values[cl.rowptr_val[c] + (j - cl.rowptr[c]) * cl.cluster_sz[c] + l];
        if (v != V(0)) {
          const I p = fill[row + l]++;
          csr.col_index[p] = cl.colids[j];
          csr.values[p] = v;
        }
      }
    row += cl.cluster_sz[c];
  }
  return csr;
}

template <class I, class V>
COO_Matrix<I, V> csr_to_coo(const CSR_Matrix<I, V> &csr) {
  COO_Matrix<I, V> coo;
  coo.num_rows = csr.num_rows;
  coo.num_cols = csr.num_cols;
  coo.num_nnzs = csr.num_nnzs;
  coo.tag = 0;
  coo.kernel_flag = 0;
  coo.partition = nullptr;
  coo.row_index = spops_checked(new_array<I>(coo.num_nnzs));
  coo.col_index = spops_checked(new_array<I>(coo.num_nnzs));
  coo.values = spops_checked(new_array<V>(coo.num_nnzs));
  for (I r = 0; r < csr.num_rows; r++)
    for (I j = csr.row_offset[r]; j < csr.row_offset[r + 1]; j++) {
      coo.row_index[j] = r;
      coo.col_index[j] = csr.col_index[j];
      coo.values[j] = csr.values[j];
    }
  return coo;
}

/* Transpose through COO with duplicates summed and exact zeros dropped (the reference's
 * csr_transpose, include/sparse_conversion.h:2067-2108). */
template <typename I, typename V>
CSR_Matrix<I, V> csr_transpose(const CSR_Matrix<I, V> &csr) {
  COO_Matrix<I, V> coo = csr_to_coo(csr);
  std::swap(coo.row_index, coo.col_index);
  std::swap(coo.num_rows, coo.num_cols);
  CSR_Matrix<I, V> t = coo_to_csr(coo, true);
  delete_host_matrix(coo);
  return t;
}

/* csr_to_binary_pattern (include/sparse_conversion.h:2121-2158): the same structure with every
 * value 1; an empty matrix gets a zero row pointer and null column / value arrays. */
template <class I, class V>
CSR_Matrix<I, V> csr_to_binary_pattern(const CSR_Matrix<I, V> &csr) {
  CSR_Matrix<I, V> b;
  b.num_rows = csr.num_rows;
  b.num_cols = csr.num_cols;
  b.num_nnzs = csr.num_nnzs;
  b.tag = 0;
  b.kernel_flag = csr.kernel_flag;
  b.partition = nullptr;
  b.row_offset = spops_checked(new_array<I>((size_t)b.num_rows + 1));
  if (csr.num_nnzs == 0) {
    std::fill_n(b.row_offset, (size_t)b.num_rows + 1, (I)0);
    b.col_index = nullptr;
    b.values = nullptr;
    return b;
  }
  b.col_index = spops_checked(new_array<I>((size_t)b.num_nnzs));
  b.values = spops_checked(new_array<V>((size_t)b.num_nnzs));
  std::copy(csr.row_offset, csr.row_offset + b.num_rows + 1, b.row_offset);
  std::copy(csr.col_index, csr.col_index + b.num_nnzs, b.col_index);
  std::fill_n(b.values, (size_t)b.num_nnzs, (V)1);
  return b;
}

#endif
