// Test harness around the reference implementation (lsl036/SparseOps), built from
// the reference's own sources by oracle/Makefile into oracle/_ref/.
//
// This file is test infrastructure only: it drives the reference's functions on
// inputs written by the test suite and dumps their outputs as raw little-endian
// arrays, so the CPU restatement (oracle/oracle.c) and the HIP library can be
// checked against the reference itself.  It is also the "reference" CPU
// baseline timed by bench.py.
//
// Input CSR file ("SPOCSR01"): int64 n_rows, n_cols, nnz; int64 rowptr[n+1];
// int64 col[nnz]; float64 val[nnz].
//
// Modes:
//   spmv   <A.csr> <x.f64> <outdir> <sigma> <C> <bR> <bC> <prec 32|64>
//   lespgemm <A.csr> <B.csr> <outdir> <k> <bands> <seed> <hc_v> <cluster_size> <kernel> <prec>
//   bench_spmv <A.csr> <fmt csr|sell> <prec> <threads> <min_seconds> [sigma C]
//   bench_lespgemm <A.csr> <B.csr> <threads> <iterations> <k> <bands> <seed> <cluster_size> <kernel>
//   bench_lespgemm_panels <A.csr> <B.csr> <threads> <iterations> <k> <bands> <seed> <cluster_size> <kernel>
//                         <c0:c1,c0:c1,...>   LeSpGEMM_VLength on cluster ranges of the clustered A
//   features <A.mtx>      the reference's feature extractor (MTX<int,double>: MtxLoad,
//                         CalculateFeatures, FeaturesPrint) on a Matrix Market file
//   bench_spmm <A.csr> <k> <threads> <min_seconds> <prec> <transpose 0|1>
//                         k right-hand sides as k calls of the reference's OpenMP SpOps_csr
//                         (kernel_flag 2, nnz-balanced rows) on A (or on csr_transpose(A),
//                         built before the timing); the reference has no SpMM entry point
//   bench_spmm_batch <list> <threads> <prec>
//                         the same over a batch: <list> holds lines "<A.csr> <k>"; every
//                         matrix is timed on one k-column product after a warm-up
//   rowwise <A.csr> <B.csr> <outdir> <kernel 1|2|3> <prec>
//                         the row-wise LeSpGEMM<sortOutput = true> (src/spgemm.cpp:684-705):
//                         C's rowptr, col and values
//   features_full <A.mtx> <matID>
//                         MtxLoad, CalculateFeatures, CalculateTilesExtraFeatures, FeaturesPrint,
//                         ExtraFeaturesPrint: the whole printout of test/Get_features.cpp
//   read_mtx <A.mtx> <outdir> <compact 0|1>
//                         the reference's read_csr_matrix<int64_t, double> (utils/sparse_io.cpp:
//                         26-214): rowptr, col and values, dimensions in "dims"
//   bench_rowwise <A.csr> <B.csr> <threads> <iterations> <kernel 1|2|3>
//                         the row-wise LeSpGEMM<sortOutput = true> in FP64: one warm-up, then the mean
//                         wall time of the iterations and C's nonzero count
//   bench_topk <A.csr> <threads> <iterations> <top_k>
//                         HashSpGEMMTopK in FP64 (A A^T top-k Jaccard candidates): the same timing
//   topk_hc <A.csr> <outdir> <top_k> <cluster_size> <prec>
//                         HashSpGEMMTopK (C's rowptr, col, values), then the flow of
//                         test/test_spgemm_hc.cpp on those pairs: C written as "%20.16g" text
//                         and read back into map<pair, V> (test/test_spgemm_hc.cpp:37-85),
//                         hierarchical_clustering_v0 and reordered_dict_to_permutation_and_offsets
//                         (perm, offsets and the map's root keys)
#include "SpOps.h"
#include "spgemm.h"
#include "spgemm_MinHashLSH.h"
#include "spgemm_utility.h"
#include "sparse_conversion.h"
#include "sparse_features.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#include <omp.h>

namespace {

struct HostCsr {
  int64_t n_rows = 0, n_cols = 0, nnz = 0;
  std::vector<int64_t> rowptr, col;
  std::vector<double> val;
};

HostCsr read_csr_file(const char *path) {
  HostCsr m;
  FILE *f = std::fopen(path, "rb");
  if (!f) { std::fprintf(stderr, "cannot open %s\n", path); std::exit(2); }
  char magic[8];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "SPOCSR01", 8) != 0) {
    std::fprintf(stderr, "bad magic in %s\n", path); std::exit(2);
  }
  int64_t hdr[3];
  if (std::fread(hdr, sizeof(int64_t), 3, f) != 3) std::exit(2);
  m.n_rows = hdr[0]; m.n_cols = hdr[1]; m.nnz = hdr[2];
  m.rowptr.resize(m.n_rows + 1); m.col.resize(m.nnz); m.val.resize(m.nnz);
  if (std::fread(m.rowptr.data(), 8, m.n_rows + 1, f) != (size_t)(m.n_rows + 1)) std::exit(2);
  if (std::fread(m.col.data(), 8, m.nnz, f) != (size_t)m.nnz) std::exit(2);
  if (std::fread(m.val.data(), 8, m.nnz, f) != (size_t)m.nnz) std::exit(2);
  std::fclose(f);
  return m;
}

template <typename T>
void dump(const std::string &dir, const char *name, const T *p, size_t n) {
  std::string path = dir + "/" + name;
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { std::fprintf(stderr, "cannot write %s\n", path.c_str()); std::exit(2); }
  if (n) std::fwrite(p, sizeof(T), n, f);
  std::fclose(f);
}

template <typename I, typename V>
CSR_Matrix<I, V> to_ref_csr(const HostCsr &h) {
  CSR_Matrix<I, V> c;
  c.num_rows = (I)h.n_rows; c.num_cols = (I)h.n_cols; c.num_nnzs = (I)h.nnz;
  c.row_offset = new_array<I>(h.n_rows + 1);
  c.col_index = new_array<I>(h.nnz > 0 ? h.nnz : 1);
  c.values = new_array<V>(h.nnz > 0 ? h.nnz : 1);
  for (int64_t i = 0; i <= h.n_rows; ++i) c.row_offset[i] = (I)h.rowptr[i];
  for (int64_t i = 0; i < h.nnz; ++i) { c.col_index[i] = (I)h.col[i]; c.values[i] = (V)h.val[i]; }
  c.partition = nullptr; c.kernel_flag = 0; c.tag = 0;
  return c;
}

// The reference's SELL / SELL-C-R kernels read x[(size_t)-1] for padding slots
// (src/spmv_sell_c_sigma.cpp:46).  Give x a zero guard element in front so the
// harness result is defined (padding values are 0).
template <typename V>
V *guarded_x(const std::vector<double> &x) {
  V *buf = new_array<V>(x.size() + 1);
  buf[0] = 0;
  for (size_t i = 0; i < x.size(); ++i) buf[i + 1] = (V)x[i];
  return buf + 1;
}

template <typename V>
int run_spmv(const HostCsr &h, const std::vector<double> &xd, const std::string &out,
             int sigma, int C, int bR, int bC) {
  typedef int I;
  CSR_Matrix<I, V> A = to_ref_csr<I, V>(h);
  V *x = guarded_x<V>(xd);
  std::vector<V> y(h.n_rows);
  const V alpha = (V)0.8, beta = (V)0.7;
  std::vector<V> y0(h.n_rows);
  for (int64_t i = 0; i < h.n_rows; ++i) y0[i] = (V)(0.25 * (double)((i * 7919) % 13) - 1.0);

  // CSR serial, alpha=1 beta=0 (src/spmv_csr.cpp:65)
  A.kernel_flag = 0;
  SpOps_csr<I, V>((V)1, A, x, (V)0, y.data());
  dump(out, "y_csr_a1b0", y.data(), y.size());
  y = y0;
  SpOps_csr<I, V>(alpha, A, x, beta, y.data());
  dump(out, "y_csr_ab", y.data(), y.size());
  // nnz-balanced OMP variant (kernel_flag 2, src/spmv_csr.cpp:138)
  A.kernel_flag = 2;
  y = y0;
  SpOps_csr<I, V>(alpha, A, x, beta, y.data());
  dump(out, "y_csr_lb_ab", y.data(), y.size());

  // SELL-C-sigma (include/sparse_conversion.h:345)
  {
    const I align = (I)(SIMD_WIDTH / 8 / sizeof(V));
    SELL_C_Sigma_Matrix<I, V> S = csr_to_sell_c_sigma(A, nullptr, sigma, C, align);
    dump(out, "sell_reorder", S.reorder, (size_t)S.num_rows);
    dump(out, "sell_chunk_len", S.chunk_len, (size_t)S.validchunkNum);
    std::vector<I> cols; std::vector<V> vals;
    for (I c = 0; c < S.validchunkNum; ++c) {
      size_t e = (size_t)S.chunk_len[c] * S.chunkWidth_C;
      cols.insert(cols.end(), S.col_index[c], S.col_index[c] + e);
      vals.insert(vals.end(), S.values[c], S.values[c] + e);
    }
    dump(out, "sell_col", cols.data(), cols.size());
    dump(out, "sell_val", vals.data(), vals.size());
    int64_t meta[6] = {S.sliceWidth_Sigma, S.chunkWidth_C, S.sliceNum, S.chunkNum, S.validchunkNum, S.alignment};
    dump(out, "sell_meta", meta, 6);
    S.kernel_flag = 0;
    y = y0;
    SpOps_sell_c_sigma<I, V>(alpha, S, x, beta, y.data());
    dump(out, "y_sell_ab", y.data(), y.size());
    delete_host_matrix(S);
  }
  // SELL-C-R (include/sparse_conversion.h:491)
  {
    const I align = (I)(SIMD_WIDTH / 8 / sizeof(V));
    SELL_C_R_Matrix<I, V> R = csr_to_sell_c_R(A, nullptr, C, align);
    dump(out, "sellr_reorder", R.reorder, (size_t)R.num_rows);
    dump(out, "sellr_chunk_len", R.chunk_len, (size_t)R.validchunkNum);
    R.kernel_flag = 0;
    y = y0;
    SpOps_sell_c_R<I, V>(alpha, R, x, beta, y.data());
    dump(out, "y_sellr_ab", y.data(), y.size());
    delete_host_matrix(R);
  }
  // ELL, row-major (include/sparse_conversion.h:195)
  {
    ELL_Matrix<I, V> E = csr_to_ell(A, RowMajor);
    int64_t w = E.max_row_width;
    dump(out, "ell_width", &w, 1);
    E.kernel_flag = 1;
    y = y0;
    SpOps_ell<I, V>(alpha, E, x, beta, y.data());
    dump(out, "y_ell_ab", y.data(), y.size());
    delete_host_matrix(E);
  }
  // BSR (include/sparse_conversion.h:777); the reference reads x past num_cols
  // when num_cols % bC != 0 (src/spmv_bsr.cpp:46), so only call it when aligned.
  if (h.n_cols % bC == 0) {
    BSR_Matrix<I, V> B = csr_to_bsr(A, (I)bR, (I)bC);
    dump(out, "bsr_row_ptr", B.row_ptr, (size_t)B.mb + 1);
    dump(out, "bsr_colidx", B.block_colindex, (size_t)B.nnzb);
    dump(out, "bsr_data", B.block_data, (size_t)B.nnzb * bR * bC);
    B.kernel_flag = 0;
    y = y0;
    SpOps_bsr<I, V>(alpha, B, x, beta, y.data());
    dump(out, "y_bsr_ab", y.data(), y.size());
    delete_host_matrix(B);
  }
  dump(out, "y0", y0.data(), y0.size());
  delete_array(x - 1);
  delete_host_matrix(A);
  return 0;
}

template <typename V>
int run_lespgemm(const HostCsr &ha, const HostCsr &hb, const std::string &out, int k, int bands,
                 uint64_t seed, int hc_v, int64_t cluster_size, int kernel, double l2_fraction) {
  typedef int64_t I;
  CSR_Matrix<I, V> A = to_ref_csr<I, V>(ha);
  CSR_Matrix<I, V> B = to_ref_csr<I, V>(hb);

  std::vector<uint64_t> sigs = minhash_signatures_flat<I>(A.row_offset, A.col_index, A.num_rows, k, seed);
  dump(out, "signatures", sigs.data(), sigs.size());

  const bool constrain = (hc_v == 0);
  std::vector<CandidatePair<I, V>> pairs = lsh_candidate_pairs_vector<I, V>(
      A.row_offset, A.col_index, A.num_rows, k, bands, seed, constrain);
  std::vector<int64_t> pi(pairs.size()), pj(pairs.size());
  std::vector<double> ps(pairs.size());
  for (size_t t = 0; t < pairs.size(); ++t) { pi[t] = pairs[t].i; pj[t] = pairs[t].j; ps[t] = (double)pairs[t].score; }
  dump(out, "pairs_i", pi.data(), pi.size());
  dump(out, "pairs_j", pj.data(), pj.size());
  dump(out, "pairs_score", ps.data(), ps.size());

  std::vector<I> perm, offsets;
  if (hc_v == 1)
    hierarchical_clustering_v1<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, (I)cluster_size, perm, offsets);
  else
    hierarchical_clustering_v0_fast<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, (I)cluster_size, perm, offsets);
  dump(out, "perm", perm.data(), perm.size());
  dump(out, "offsets", offsets.data(), offsets.size());

  CSR_Matrix<I, V> Ar = csr_reorder_rows<I, V>(A, perm);
  CSR_VlengthCluster<I, V> Acl = csr_to_vlength_cluster<I, V>(Ar, offsets);
  int64_t acl_meta[3] = {Acl.rows, Acl.nnzc, Acl.nnzv};
  dump(out, "acl_meta", acl_meta, 3);

  long long flops = get_spgemm_flop(A, B);
  dump(out, "flops", &flops, 1);

  CSR_VlengthCluster<I, V> Ccl;
  Ccl.acc_flag = new_array<char>(Acl.rows);
  Ccl.min_ccol = new_array<I>(Acl.rows);
  Ccl.max_ccol = new_array<I>(Acl.rows);
  LeSpGEMM_VLength<false, I, V>(Acl, B, Ccl, kernel, l2_fraction);
  int64_t ccl_meta[3] = {Ccl.rows, Ccl.nnzc, Ccl.nnzv};
  dump(out, "ccl_meta", ccl_meta, 3);
  if (kernel == 3) {
    int64_t dense = Ccl.dense_cluster_count;
    dump(out, "dense_cluster_count", &dense, 1);
  }
  dump(out, "ccl_rowptr", Ccl.rowptr, (size_t)Ccl.rows + 1);
  if (kernel == 3) {
    std::vector<int64_t> mn(Ccl.min_ccol, Ccl.min_ccol + Ccl.rows), mx(Ccl.max_ccol, Ccl.max_ccol + Ccl.rows);
    dump(out, "ccl_min", mn.data(), mn.size());
    dump(out, "ccl_max", mx.data(), mx.size());
    std::vector<int8_t> fl(Ccl.acc_flag, Ccl.acc_flag + Ccl.rows);
    dump(out, "ccl_acc_flag", fl.data(), fl.size());
  }
  CSR_Matrix<I, V> Cc = vlength_cluster2csr<I, V>(Ccl);
  std::vector<I> inv(A.num_rows);
  for (I t = 0; t < A.num_rows; ++t) inv[perm[t]] = t;
  CSR_Matrix<I, V> C = csr_reorder_rows<I, V>(Cc, inv);
  dump(out, "C_rowptr", C.row_offset, (size_t)C.num_rows + 1);
  dump(out, "C_col", C.col_index, (size_t)C.num_nnzs);
  std::vector<double> cv(C.values, C.values + C.num_nnzs);
  dump(out, "C_val", cv.data(), cv.size());

  delete_host_matrix(C); delete_host_matrix(Cc);
  delete_array(Ccl.acc_flag); delete_array(Ccl.min_ccol); delete_array(Ccl.max_ccol);
  delete_vlength_cluster_matrix(Ccl); delete_vlength_cluster_matrix(Acl);
  delete_host_matrix(Ar); delete_host_matrix(A); delete_host_matrix(B);
  return 0;
}

template <typename V>
int bench_spmv(const HostCsr &h, const std::string &fmt, double min_seconds, int sigma, int C) {
  typedef int I;
  CSR_Matrix<I, V> A = to_ref_csr<I, V>(h);
  std::vector<double> xd(h.n_cols);
  for (int64_t i = 0; i < h.n_cols; ++i) xd[i] = (double)((i * 2654435761ULL) % 1000) / 1000.0 - 0.5;
  V *x = guarded_x<V>(xd);
  std::vector<V> y(h.n_rows, 0);
  const int threads = Le_get_thread_num();
  double t_iter = 0; int iters = 0;
  if (fmt == "csr") {
    A.kernel_flag = 2;
    A.partition = new_array<I>(threads + 1);
    balanced_partition_row_by_nnz(A.row_offset, A.num_rows, threads, A.partition);
    SpOps_csr<I, V>((V)1, A, x, (V)0, y.data());          // warm-up
    double t0 = omp_get_wtime();
    do { SpOps_csr<I, V>((V)1, A, x, (V)0, y.data()); ++iters; } while (iters < 20 || omp_get_wtime() - t0 < min_seconds);
    t_iter = (omp_get_wtime() - t0) / iters;
    delete_array(A.partition); A.partition = nullptr;
  } else {
    const I align = (I)(SIMD_WIDTH / 8 / sizeof(V));
    SELL_C_Sigma_Matrix<I, V> S = csr_to_sell_c_sigma(A, nullptr, sigma, C, align);
    S.kernel_flag = 2;
    S.partition = new_array<I>(threads + 1);
    balanced_partition_row_by_nnz_sell(S.col_index, S.num_nnzs, S.chunkWidth_C, S.validchunkNum, S.chunk_len, threads, S.partition);
    SpOps_sell_c_sigma<I, V>((V)1, S, x, (V)0, y.data());
    double t0 = omp_get_wtime();
    do { SpOps_sell_c_sigma<I, V>((V)1, S, x, (V)0, y.data()); ++iters; } while (iters < 20 || omp_get_wtime() - t0 < min_seconds);
    t_iter = (omp_get_wtime() - t0) / iters;
    delete_array(S.partition); S.partition = nullptr;
    delete_host_matrix(S);
  }
  double checksum = 0; for (double v : y) checksum += v;
  std::printf("{\"format\": \"%s\", \"threads\": %d, \"iterations\": %d, \"seconds_per_iter\": %.9e, \"checksum\": %.17g}\n",
              fmt.c_str(), threads, iters, t_iter, checksum);
  delete_array(x - 1);
  delete_host_matrix(A);
  return 0;
}

// One k-column product: k calls of SpOps_csr (kernel_flag 2) on columns of X (column-major,
// n_cols x k) into Y (n_rows x k).  X holds a deterministic pattern (timing does not depend on
// the values).
template <typename V>
struct SpmmCase {
  CSR_Matrix<int, V> A;
  int k = 0;
  std::vector<V> X, Y;
  void prepare(int threads) {
    X.resize((size_t)A.num_cols * k);
    Y.assign((size_t)A.num_rows * k, (V)0);
    for (size_t i = 0; i < X.size(); ++i) X[i] = (V)((double)((i * 2654435761ULL) % 1000) / 1000.0 - 0.5);
    A.kernel_flag = 2;
    A.partition = new_array<int>(threads + 1);
    balanced_partition_row_by_nnz(A.row_offset, A.num_rows, threads, A.partition);
  }
  void run() {
    for (int j = 0; j < k; ++j)
      SpOps_csr<int, V>((V)1, A, X.data() + (size_t)j * A.num_cols, (V)0, Y.data() + (size_t)j * A.num_rows);
  }
  void release() {
    delete_array(A.partition);
    A.partition = nullptr;
    delete_host_matrix(A);
  }
};

template <typename V>
int bench_spmm(const HostCsr &h, int k, double min_seconds, bool transpose) {
  SpmmCase<V> c;
  CSR_Matrix<int, V> A = to_ref_csr<int, V>(h);
  if (transpose) {
    c.A = csr_transpose<int, V>(A);
    delete_host_matrix(A);
  } else {
    c.A = A;
  }
  c.k = k;
  const int threads = Le_get_thread_num();
  c.prepare(threads);
  c.run();                                              // warm-up
  int iters = 0;
  double t0 = omp_get_wtime();
  do { c.run(); ++iters; } while (iters < 3 || omp_get_wtime() - t0 < min_seconds);
  const double t = (omp_get_wtime() - t0) / iters;
  std::printf("{\"threads\": %d, \"k\": %d, \"nnz\": %d, \"iterations\": %d, \"seconds_per_spmm\": %.9e}\n",
              threads, k, c.A.num_nnzs, iters, t);
  c.release();
  return 0;
}

template <typename V>
int bench_spmm_batch(const char *list) {
  FILE *f = std::fopen(list, "r");
  if (!f) { std::fprintf(stderr, "cannot open %s\n", list); return 2; }
  char path[4096];
  int k = 0;
  const int threads = Le_get_thread_num();
  double total = 0.0, flop = 0.0;
  int count = 0;
  while (std::fscanf(f, "%4095s %d", path, &k) == 2) {
    SpmmCase<V> c;
    c.A = to_ref_csr<int, V>(read_csr_file(path));
    c.k = k;
    c.prepare(threads);
    c.run();                                            // warm-up
    const double t0 = omp_get_wtime();
    c.run();
    total += omp_get_wtime() - t0;
    flop += 2.0 * (double)c.A.num_nnzs * k;
    ++count;
    c.release();
  }
  std::fclose(f);
  std::printf("{\"threads\": %d, \"matrices\": %d, \"flop\": %.17g, \"seconds\": %.9e}\n", threads, count, flop,
              total);
  return 0;
}

int bench_lespgemm(const HostCsr &ha, const HostCsr &hb, int iterations, int k, int bands, uint64_t seed,
                   int64_t cluster_size, int kernel) {
  typedef int64_t I; typedef double V;
  CSR_Matrix<I, V> A = to_ref_csr<I, V>(ha);
  CSR_Matrix<I, V> B = to_ref_csr<I, V>(hb);
  double t0 = omp_get_wtime();
  std::vector<CandidatePair<I, V>> pairs = lsh_candidate_pairs_vector<I, V>(
      A.row_offset, A.col_index, A.num_rows, k, bands, seed, true);
  double t_pairs = omp_get_wtime() - t0;
  std::vector<I> perm, offsets;
  t0 = omp_get_wtime();
  hierarchical_clustering_v0_fast<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, (I)cluster_size, perm, offsets);
  double t_hc = omp_get_wtime() - t0;
  CSR_Matrix<I, V> Ar = csr_reorder_rows<I, V>(A, perm);
  t0 = omp_get_wtime();
  CSR_VlengthCluster<I, V> Acl = csr_to_vlength_cluster<I, V>(Ar, offsets);
  double t_conv = omp_get_wtime() - t0;
  long long flops = get_spgemm_flop(A, B);
  CSR_VlengthCluster<I, V> Ccl;
  Ccl.acc_flag = new_array<char>(Acl.rows);
  Ccl.min_ccol = new_array<I>(Acl.rows);
  Ccl.max_ccol = new_array<I>(Acl.rows);
  LeSpGEMM_VLength<false, I, V>(Acl, B, Ccl, kernel, -1.0);     // warm-up
  int64_t nnzc = Ccl.nnzc, nnzv = Ccl.nnzv;
  delete_vlength_cluster_matrix(Ccl);
  double total = 0;
  for (int it = 0; it < iterations; ++it) {
    t0 = omp_get_wtime();
    LeSpGEMM_VLength<false, I, V>(Acl, B, Ccl, kernel, -1.0);
    total += omp_get_wtime() - t0;
    delete_vlength_cluster_matrix(Ccl);
  }
  double avg = total / iterations;
  std::printf("{\"threads\": %d, \"pairs\": %zu, \"clusters\": %zu, \"genpairs_ms\": %.3f, \"hc_ms\": %.3f, "
              "\"conversion_ms\": %.3f, \"lespgemm_ms\": %.3f, \"flops\": %lld, \"gflops\": %.4f, \"nnzc\": %lld, \"nnzv\": %lld}\n",
              Le_get_thread_num(), pairs.size(), offsets.size() - 1, t_pairs * 1e3, t_hc * 1e3, t_conv * 1e3, avg * 1e3,
              flops, (double)flops / 1e9 / avg, (long long)nnzc, (long long)nnzv);
  delete_array(Ccl.acc_flag); delete_array(Ccl.min_ccol); delete_array(Ccl.max_ccol);
  delete_vlength_cluster_matrix(Acl); delete_host_matrix(Ar); delete_host_matrix(A); delete_host_matrix(B);
  return 0;
}

// Clusters [c0, c1) of a clustered matrix as a matrix of its own (arrays copied, pointers
// rebased), so that LeSpGEMM_VLength computes that panel of the product.
template <typename I, typename V>
CSR_VlengthCluster<I, V> cluster_slice(const CSR_VlengthCluster<I, V> &A, I c0, I c1) {
  CSR_VlengthCluster<I, V> S{};
  S.rows = c1 - c0;
  S.cols = A.cols;
  S.cluster_sz = new_array<I>(S.rows > 0 ? S.rows : 1);
  S.rowptr = new_array<I>(S.rows + 1);
  S.rowptr_val = new_array<I>(S.rows + 1);
  S.csr_rows = 0;
  S.max_cluster_sz = 0;
  for (I c = 0; c < S.rows; ++c) {
    S.cluster_sz[c] = A.cluster_sz[c0 + c];
    S.csr_rows += S.cluster_sz[c];
    S.max_cluster_sz = std::max(S.max_cluster_sz, S.cluster_sz[c]);
  }
  for (I c = 0; c <= S.rows; ++c) {
    S.rowptr[c] = A.rowptr[c0 + c] - A.rowptr[c0];
    S.rowptr_val[c] = A.rowptr_val[c0 + c] - A.rowptr_val[c0];
  }
  S.nnzc = S.rowptr[S.rows];
  S.nnzv = S.rowptr_val[S.rows];
  S.colids = new_array<I>(S.nnzc > 0 ? S.nnzc : 1);
  S.values = new_array<V>(S.nnzv > 0 ? S.nnzv : 1);
  std::copy(A.colids + A.rowptr[c0], A.colids + A.rowptr[c1], S.colids);
  std::copy(A.values + A.rowptr_val[c0], A.values + A.rowptr_val[c1], S.values);
  return S;
}

int bench_lespgemm_panels(const HostCsr &ha, const HostCsr &hb, int iterations, int k, int bands, uint64_t seed,
                          int64_t cluster_size, int kernel, const char *ranges, const char *dump_dir) {
  typedef int64_t I; typedef double V;
  CSR_Matrix<I, V> A = to_ref_csr<I, V>(ha);
  CSR_Matrix<I, V> B = to_ref_csr<I, V>(hb);
  double t0 = omp_get_wtime();
  std::vector<CandidatePair<I, V>> pairs = lsh_candidate_pairs_vector<I, V>(
      A.row_offset, A.col_index, A.num_rows, k, bands, seed, true);
  double t_pairs = omp_get_wtime() - t0;
  std::vector<I> perm, offsets;
  t0 = omp_get_wtime();
  hierarchical_clustering_v0_fast<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, (I)cluster_size, perm, offsets);
  double t_hc = omp_get_wtime() - t0;
  if (dump_dir != nullptr) {
    dump(dump_dir, "perm", perm.data(), perm.size());
    dump(dump_dir, "offsets", offsets.data(), offsets.size());
  }
  CSR_Matrix<I, V> Ar = csr_reorder_rows<I, V>(A, perm);
  t0 = omp_get_wtime();
  CSR_VlengthCluster<I, V> Acl = csr_to_vlength_cluster<I, V>(Ar, offsets);
  double t_conv = omp_get_wtime() - t0;
  // fingerprints of the clustering: sum of perm[i] * (i + 1) and offsets[i] * (i + 1), mod 2^64
  uint64_t fp_perm = 0, fp_off = 0;
  for (size_t i = 0; i < perm.size(); ++i) fp_perm += (uint64_t)perm[i] * (uint64_t)(i + 1);
  for (size_t i = 0; i < offsets.size(); ++i) fp_off += (uint64_t)offsets[i] * (uint64_t)(i + 1);
  std::printf("{\"threads\": %d, \"pairs\": %zu, \"clusters\": %zu, \"genpairs_ms\": %.3f, \"hc_ms\": %.3f, "
              "\"conversion_ms\": %.3f, \"perm_fingerprint\": %llu, \"offsets_fingerprint\": %llu}\n",
              Le_get_thread_num(), pairs.size(), offsets.size() - 1, t_pairs * 1e3, t_hc * 1e3, t_conv * 1e3,
              (unsigned long long)fp_perm, (unsigned long long)fp_off);
  std::fflush(stdout);
  const char *p = ranges;
  while (*p) {
    char *end = nullptr;
    const I c0 = std::strtoll(p, &end, 10);
    if (*end != ':') { std::fprintf(stderr, "bad range list\n"); return 2; }
    const I c1 = std::strtoll(end + 1, &end, 10);
    p = *end == ',' ? end + 1 : end;
    if (c0 < 0 || c1 < c0 || c1 > Acl.rows) { std::fprintf(stderr, "bad range %lld:%lld\n", (long long)c0, (long long)c1); return 2; }
    CSR_VlengthCluster<I, V> S = cluster_slice(Acl, c0, c1);
    long long products = 0;
    for (I c = 0; c < S.rows; ++c)
      for (I t = S.rowptr[c]; t < S.rowptr[c + 1]; ++t) products += B.row_offset[S.colids[t] + 1] - B.row_offset[S.colids[t]];
    CSR_VlengthCluster<I, V> Ccl{};
    Ccl.acc_flag = new_array<char>(S.rows > 0 ? S.rows : 1);
    Ccl.min_ccol = new_array<I>(S.rows > 0 ? S.rows : 1);
    Ccl.max_ccol = new_array<I>(S.rows > 0 ? S.rows : 1);
    LeSpGEMM_VLength<false, I, V>(S, B, Ccl, kernel, -1.0);     // warm-up
    double total = 0;
    for (int it = 0; it < iterations; ++it) {
      delete_vlength_cluster_matrix(Ccl);
      t0 = omp_get_wtime();
      LeSpGEMM_VLength<false, I, V>(S, B, Ccl, kernel, -1.0);
      total += omp_get_wtime() - t0;
    }
    double sum = 0, sum_abs = 0, colsum = 0, nonzeros = 0;
    for (I i = 0; i < Ccl.nnzv; ++i) { sum += Ccl.values[i]; sum_abs += std::fabs(Ccl.values[i]); nonzeros += Ccl.values[i] != 0.0; }
    for (I i = 0; i < Ccl.nnzc; ++i) colsum += (double)Ccl.colids[i];
    std::printf("{\"c0\": %lld, \"c1\": %lld, \"lespgemm_ms\": %.3f, \"products\": %lld, \"nnzc\": %lld, \"nnzv\": %lld, "
                "\"sum\": %.17g, \"sum_abs\": %.17g, \"colid_sum\": %.17g, \"nonzeros\": %.17g}\n",
                (long long)c0, (long long)c1, iterations ? total / iterations * 1e3 : 0.0, products,
                (long long)Ccl.nnzc, (long long)Ccl.nnzv, sum, sum_abs, colsum, nonzeros);
    std::fflush(stdout);
    delete_vlength_cluster_matrix(Ccl);
    delete_array(Ccl.acc_flag); delete_array(Ccl.min_ccol); delete_array(Ccl.max_ccol);
    delete_vlength_cluster_matrix(S);
  }
  delete_vlength_cluster_matrix(Acl); delete_host_matrix(Ar); delete_host_matrix(A); delete_host_matrix(B);
  return 0;
}

std::vector<double> read_f64(const char *path, size_t n) {
  std::vector<double> v(n);
  FILE *f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), 8, n, f) != n) { std::fprintf(stderr, "cannot read %s\n", path); std::exit(2); }
  std::fclose(f);
  return v;
}

template <typename V>
void dump_csr(const std::string &out, const char *tag, const CSR_Matrix<int64_t, V> &C) {
  std::vector<double> v(C.values, C.values + C.num_nnzs);
  dump(out, (std::string(tag) + "_rowptr").c_str(), C.row_offset, (size_t)C.num_rows + 1);
  dump(out, (std::string(tag) + "_col").c_str(), C.col_index, (size_t)C.num_nnzs);
  dump(out, (std::string(tag) + "_val").c_str(), v.data(), v.size());
}

template <typename V>
int run_rowwise(const HostCsr &ha, const HostCsr &hb, const std::string &out, int kernel) {
  CSR_Matrix<int64_t, V> A = to_ref_csr<int64_t, V>(ha), B = to_ref_csr<int64_t, V>(hb), C;
  LeSpGEMM<true, int64_t, V>(A, B, C, kernel);
  dump_csr(out, "C", C);
  delete_host_matrix(A);
  delete_host_matrix(B);
  delete_host_matrix(C);
  return 0;
}

int bench_rowwise(const HostCsr &ha, const HostCsr &hb, int iterations, int kernel) {
  CSR_Matrix<int64_t, double> A = to_ref_csr<int64_t, double>(ha), B = to_ref_csr<int64_t, double>(hb);
  int64_t nnz = 0;
  double total = 0.0;
  for (int it = 0; it <= iterations; ++it) {
    CSR_Matrix<int64_t, double> C;
    const double t0 = omp_get_wtime();
    LeSpGEMM<true, int64_t, double>(A, B, C, kernel);
    const double t = omp_get_wtime() - t0;
    if (it > 0) total += t;
    nnz = C.num_nnzs;
    delete_host_matrix(C);
  }
  std::printf("{\"threads\": %d, \"iterations\": %d, \"kernel\": %d, \"seconds\": %.9e, \"nnz_C\": %lld}\n",
              Le_get_thread_num(), iterations, kernel, total / std::max(iterations, 1), (long long)nnz);
  delete_host_matrix(A);
  delete_host_matrix(B);
  return 0;
}

int bench_topk(const HostCsr &ha, int iterations, int64_t top_k) {
  CSR_Matrix<int64_t, double> A = to_ref_csr<int64_t, double>(ha);
  int64_t nnz = 0;
  double total = 0.0;
  for (int it = 0; it <= iterations; ++it) {
    CSR_Matrix<int64_t, double> C;
    const double t0 = omp_get_wtime();
    HashSpGEMMTopK<int64_t, double>(A, C, top_k);
    const double t = omp_get_wtime() - t0;
    if (it > 0) total += t;
    nnz = C.num_nnzs;
    delete_host_matrix(C);
  }
  std::printf("{\"threads\": %d, \"iterations\": %d, \"top_k\": %lld, \"seconds\": %.9e, \"nnz_C\": %lld}\n",
              Le_get_thread_num(), iterations, (long long)top_k, total / std::max(iterations, 1), (long long)nnz);
  delete_host_matrix(A);
  return 0;
}

template <typename V>
int run_topk_hc(const HostCsr &ha, const std::string &out, int64_t top_k, int64_t cluster_size) {
  CSR_Matrix<int64_t, V> A = to_ref_csr<int64_t, V>(ha), C;
  HashSpGEMMTopK<int64_t, V>(A, C, top_k);
  dump_csr(out, "topk", C);
  // the candidate file round trip of test/generate_candidate_pairs.cpp + test/test_spgemm_hc.cpp
  std::map<std::pair<int64_t, int64_t>, V> close_pairs;
  for (int64_t i = 0; i < C.num_rows; ++i)
    for (int64_t e = C.row_offset[i]; e < C.row_offset[i + 1]; ++e) {
      char buf[64];
      std::snprintf(buf, sizeof buf, "%20.16g", (double)C.values[e]);
      std::istringstream iss(buf);
      V sim;
      iss >> sim;
      close_pairs[std::make_pair(i, (int64_t)C.col_index[e])] = sim;
    }
  std::vector<double> file_scores;
  std::vector<int64_t> keys;
  for (const auto &kv : close_pairs) {
    keys.push_back(kv.first.first);
    keys.push_back(kv.first.second);
    file_scores.push_back((double)kv.second);
  }
  dump(out, "pair_keys", keys.data(), keys.size());
  dump(out, "pair_scores", file_scores.data(), file_scores.size());
  std::map<int64_t, std::vector<int64_t>> dict =
      hierarchical_clustering_v0<int64_t, V>(A.row_offset, A.col_index, A.num_rows, close_pairs, cluster_size);
  std::vector<int64_t> perm, offsets, roots;
  reordered_dict_to_permutation_and_offsets<int64_t>(dict, A.num_rows, perm, offsets);
  for (const auto &kv : dict) roots.push_back(kv.first);
  dump(out, "perm", perm.data(), perm.size());
  dump(out, "offsets", offsets.data(), offsets.size());
  dump(out, "roots", roots.data(), roots.size());
  delete_host_matrix(A);
  delete_host_matrix(C);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: ref_harness <mode> ...\n"); return 2; }
  std::string mode = argv[1];
  const char *thr = std::getenv("OMP_NUM_THREADS");
  Le_set_thread_num(thr ? std::atoi(thr) : omp_get_max_threads());
  if (mode == "spmv" && argc >= 10) {
    HostCsr A = read_csr_file(argv[2]);
    std::vector<double> x = read_f64(argv[3], A.n_cols);
    int sigma = std::atoi(argv[5]), C = std::atoi(argv[6]), bR = std::atoi(argv[7]), bC = std::atoi(argv[8]);
    int prec = std::atoi(argv[9]);
    return prec == 32 ? run_spmv<float>(A, x, argv[4], sigma, C, bR, bC)
                      : run_spmv<double>(A, x, argv[4], sigma, C, bR, bC);
  }
  if (mode == "lespgemm" && argc >= 12) {
    HostCsr A = read_csr_file(argv[2]);
    HostCsr B = read_csr_file(argv[3]);
    int k = std::atoi(argv[5]), bands = std::atoi(argv[6]);
    uint64_t seed = std::strtoull(argv[7], nullptr, 10);
    int hc_v = std::atoi(argv[8]);
    int64_t cs = std::atoll(argv[9]);
    int kernel = std::atoi(argv[10]);
    int prec = std::atoi(argv[11]);
    double l2f = argc > 12 ? std::atof(argv[12]) : -1.0;   // kernel 3's L2 fraction (-1: its default)
    return prec == 32 ? run_lespgemm<float>(A, B, argv[4], k, bands, seed, hc_v, cs, kernel, l2f)
                      : run_lespgemm<double>(A, B, argv[4], k, bands, seed, hc_v, cs, kernel, l2f);
  }
  if (mode == "bench_spmv" && argc >= 7) {
    HostCsr A = read_csr_file(argv[2]);
    std::string fmt = argv[3];
    int prec = std::atoi(argv[4]);
    Le_set_thread_num(std::atoi(argv[5]));
    double secs = std::atof(argv[6]);
    int sigma = argc > 7 ? std::atoi(argv[7]) : 1024, C = argc > 8 ? std::atoi(argv[8]) : 32;
    return prec == 32 ? bench_spmv<float>(A, fmt, secs, sigma, C) : bench_spmv<double>(A, fmt, secs, sigma, C);
  }
  if (mode == "bench_lespgemm" && argc >= 11) {
    HostCsr A = read_csr_file(argv[2]);
    HostCsr B = read_csr_file(argv[3]);
    Le_set_thread_num(std::atoi(argv[4]));
    return bench_lespgemm(A, B, std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7]),
                          std::strtoull(argv[8], nullptr, 10), std::atoll(argv[9]), std::atoi(argv[10]));
  }
  if (mode == "bench_spmm" && argc >= 8) {
    HostCsr A = read_csr_file(argv[2]);
    Le_set_thread_num(std::atoi(argv[4]));
    const int k = std::atoi(argv[3]), prec = std::atoi(argv[6]);
    const double secs = std::atof(argv[5]);
    const bool tr = std::atoi(argv[7]) != 0;
    return prec == 32 ? bench_spmm<float>(A, k, secs, tr) : bench_spmm<double>(A, k, secs, tr);
  }
  if (mode == "bench_spmm_batch" && argc >= 5) {
    Le_set_thread_num(std::atoi(argv[3]));
    return std::atoi(argv[4]) == 32 ? bench_spmm_batch<float>(argv[2]) : bench_spmm_batch<double>(argv[2]);
  }
  if (mode == "read_mtx" && argc >= 5) {
    CSR_Matrix<int64_t, double> A = read_csr_matrix<int64_t, double>(argv[2], std::atoi(argv[4]) != 0);
    const int64_t dims[3] = {(int64_t)A.num_rows, (int64_t)A.num_cols, (int64_t)A.num_nnzs};
    dump(argv[3], "dims", dims, 3);
    if (A.num_rows > 0) dump_csr(argv[3], "A", A);
    return 0;
  }
  if (mode == "rowwise" && argc >= 7) {
    HostCsr A = read_csr_file(argv[2]);
    HostCsr B = read_csr_file(argv[3]);
    return std::atoi(argv[6]) == 32 ? run_rowwise<float>(A, B, argv[4], std::atoi(argv[5]))
                                    : run_rowwise<double>(A, B, argv[4], std::atoi(argv[5]));
  }
  if (mode == "bench_rowwise" && argc >= 7) {
    HostCsr A = read_csr_file(argv[2]);
    HostCsr B = read_csr_file(argv[3]);
    Le_set_thread_num(std::atoi(argv[4]));
    return bench_rowwise(A, B, std::atoi(argv[5]), std::atoi(argv[6]));
  }
  if (mode == "bench_topk" && argc >= 6) {
    HostCsr A = read_csr_file(argv[2]);
    Le_set_thread_num(std::atoi(argv[3]));
    return bench_topk(A, std::atoi(argv[4]), std::atoll(argv[5]));
  }
  if (mode == "topk_hc" && argc >= 7) {
    HostCsr A = read_csr_file(argv[2]);
    return std::atoi(argv[6]) == 32 ? run_topk_hc<float>(A, argv[3], std::atoll(argv[4]), std::atoll(argv[5]))
                                    : run_topk_hc<double>(A, argv[3], std::atoll(argv[4]), std::atoll(argv[5]));
  }
  if (mode == "features_full" && argc >= 4) {
    MTX<int, double> m(std::atoi(argv[3]));
    if (!m.MtxLoad(argv[2])) return 2;
    m.CalculateFeatures();
    m.CalculateTilesExtraFeatures(argv[2]);
    m.FeaturesPrint();
    m.ExtraFeaturesPrint();
    return 0;
  }
  if (mode == "features" && argc >= 3) {
    MTX<int, double> m(0);
    if (!m.MtxLoad(argv[2])) return 2;
    m.CalculateFeatures();
    m.FeaturesPrint();
    m.ExtraFeaturesPrint();
    return 0;
  }
  if (mode == "bench_lespgemm_panels" && argc >= 12) {
    HostCsr A = read_csr_file(argv[2]);
    HostCsr B = read_csr_file(argv[3]);
    Le_set_thread_num(std::atoi(argv[4]));
    return bench_lespgemm_panels(A, B, std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7]),
                                 std::strtoull(argv[8], nullptr, 10), std::atoll(argv[9]), std::atoi(argv[10]),
                                 argv[11], argc > 12 ? argv[12] : nullptr);
  }
  std::fprintf(stderr, "bad arguments for mode %s\n", mode.c_str());
  return 2;
}
