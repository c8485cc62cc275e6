// test_spgemm_hc_lsh: the reference's LeSpGEMM driver (test/test_spgemm_hc_lsh.cpp) on the
// drop-in headers, i.e. on the GPU.  Same options and the same stdout lines the reference's
// scripts parse (genPairs time, HC time, Format Conversion time, Average time, Average GFLOPS,
// [mixed_acc] dense clusters, phase breakdown).
//   test_spgemm_hc_lsh A.mtx B.mtx [--precision=32|64] [--iterations=N] [--kernel=1|2|3]
//       [--l2_fraction=f] [--cluster_size=8] [--sort=0|1] [--k=64] [--bands=16] [--seed=7]
//       [--hc_v=0|1|2] [--test_type=correctness|performance] [--print_bd=0|1] [--threads=N]
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "SpOps.h"
#include "cmdline.h"
#include "sparse_io.h"

using namespace std;

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void usage(char **argv) {
  cout << "Usage:\n\t" << argv[0] << " <A.mtx> <B.mtx> [options]\n"
       << "\t  --precision=32|64 --threads=N --iterations=10 --kernel=1|2|3 --l2_fraction=f --cluster_size=8\n"
       << "\t  --sort=0|1 --k=64 --bands=16 --seed=7 --hc_v=0|1|2 --test_type=correctness|performance --print_bd=0|1\n";
}

template <typename I, typename V>
static void pairs_and_clusters(const CSR_Matrix<I, V> &A, int k, int bands, uint64_t seed, int hc_v, I cs,
                               vector<I> &perm, vector<I> &offsets, double &t_lsh, double &t_hc, size_t &npairs) {
  double t0 = now_ms();
  vector<CandidatePair<I, V>> pairs = lsh_candidate_pairs_vector<I, V>(A.row_offset, A.col_index, A.num_rows, k, bands,
                                                                       seed, hc_v == 0);
  t_lsh = now_ms() - t0;
  npairs = pairs.size();
  cout << "LSH candidate pairs: " << pairs.size() << " (genPairs time: " << t_lsh << " ms)" << endl;
  cout << "Hierarchical clustering (cluster_size=" << cs << ") [" << (hc_v == 1 ? "v1" : "fast v0") << "]..." << endl;
  t0 = now_ms();
  if (hc_v == 1)
    hierarchical_clustering_v1<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, cs, perm, offsets);
  else
    hierarchical_clustering_v0_fast<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, cs, perm, offsets);
  t_hc = now_ms() - t0;
}

template <typename I, typename V>
static void run(const char *pa, const char *pb, bool correctness, int iterations, int kernel, bool sortOutput, I cs,
                int k, int bands, uint64_t seed, int hc_v, double l2f, bool print_bd) {
  cout << "========================================" << endl;
  cout << "MinHash+LSH -> Vlength-Cluster SpGEMM " << (correctness ? "Correctness" : "Performance") << endl;
  cout << "========================================" << endl;
  cout << "Reading A: " << pa << endl;
  CSR_Matrix<I, V> A = read_csr_matrix<I, V>(pa);
  cout << "A: " << A.num_rows << " x " << A.num_cols << ", nnz: " << A.num_nnzs << endl;
  cout << "Reading B: " << pb << endl;
  CSR_Matrix<I, V> B = read_csr_matrix<I, V>(pb);
  cout << "B: " << B.num_rows << " x " << B.num_cols << ", nnz: " << B.num_nnzs << endl;
  if (A.num_cols != B.num_rows) { cerr << "Error: A.cols != B.rows" << endl; return; }
  if (k <= 0 || bands <= 0 || k % bands != 0) {
    cerr << "Error: k must be positive and k % bands == 0. k=" << k << " bands=" << bands << endl;
    return;
  }
  cout << "Generating candidate pairs by MinHash+LSH (k=" << k << ", bands=" << bands << ", seed=" << seed
       << ") [hc_v=" << hc_v << "]..." << endl;
  vector<I> perm, offsets;
  double t_lsh = 0, t_hc = 0;
  size_t npairs = 0;
  pairs_and_clusters<I, V>(A, k, bands, seed, hc_v, cs, perm, offsets, t_lsh, t_hc, npairs);
  cout << "Clusters: " << (offsets.empty() ? 0 : offsets.size() - 1) << " (HC time: " << t_hc << " ms)" << endl;
  cout << "Reordering A by cluster..." << endl;
  CSR_Matrix<I, V> Ar = csr_reorder_rows<I, V>(A, perm);
  cout << "Converting A to CSR_VlengthCluster..." << endl;
  double t0 = now_ms();
  CSR_VlengthCluster<I, V> Acl = csr_to_vlength_cluster<I, V>(Ar, offsets);
  cout << "Format Conversion time: " << now_ms() - t0 << " ms" << endl;
  delete_host_matrix(Ar);
  cout << "A_cluster: " << Acl.rows << " clusters, nnzc=" << Acl.nnzc << ", nnzv=" << Acl.nnzv << endl;
  long long flops = get_spgemm_flop(A, B);
  cout << "FLOPs: " << flops << endl;
  cout << "Kernel: " << kernel << endl;
  CSR_VlengthCluster<I, V> C;
  if (kernel == 3) {
    C.acc_flag = new_array<char>(Acl.rows);
    C.min_ccol = new_array<I>(Acl.rows);
    C.max_ccol = new_array<I>(Acl.rows);
  }
  auto spgemm = [&]() {
    if (sortOutput) LeSpGEMM_VLength<true, I, V>(Acl, B, C, kernel, l2f);
    else LeSpGEMM_VLength<false, I, V>(Acl, B, C, kernel, l2f);
  };
  if (correctness) {
    t0 = now_ms();
    spgemm();
    double ms = now_ms() - t0;
    cout << "C_cluster: " << C.rows << " clusters, nnzc=" << C.nnzc << endl;
    cout << "Time: " << ms << " ms" << endl;
    cout << "Performance: " << (flops / 1e9) / (ms / 1000.0) << " GFLOPS" << endl;
    CSR_Matrix<I, V> Cc = vlength_cluster2csr<I, V>(C);
    vector<I> inv(A.num_rows);
    for (I t = 0; t < A.num_rows; t++) inv[perm[t]] = t;
    CSR_Matrix<I, V> Co = csr_reorder_rows<I, V>(Cc, inv);
    string out = extractFileNameWithoutExtension(pa) + "_SpOps_" +
                 (kernel == 1 ? "hashvlengthcluster_hc_lsh" : kernel == 2 ? "arrayvlengthcluster_hc_lsh" : "mixedvlengthcluster_hc_lsh") +
                 ".mtx";
    write_mtx_crd(out.c_str(), Co);
    cout << "C_csr: " << Cc.num_rows << " x " << Cc.num_cols << ", nnz: " << Cc.num_nnzs << endl;
    delete_host_matrix(Co);
    delete_host_matrix(Cc);
    delete_vlength_cluster_matrix(C);
    cout << "Correctness test done." << endl;
  } else {
    spgemm();    // warm-up
    if (kernel == 3)
      cout << "[mixed_acc] dense clusters: " << C.dense_cluster_count << " / " << C.rows << " ("
           << (100.0 * C.dense_cluster_count / (C.rows ? C.rows : 1)) << "%)" << endl;
    delete_vlength_cluster_matrix(C);
    double total = 0, s_setup = 0, s_malloc = 0, s_sym = 0, s_disp = 0, s_num = 0;
    for (int it = 0; it < iterations; it++) {
      t0 = now_ms();
      spgemm();
      total += now_ms() - t0;
      s_setup += C.mixed_acc_setup_ms; s_malloc += C.mixed_acc_malloc_ms; s_sym += C.mixed_acc_symbolic_ms;
      s_disp += C.mixed_acc_dispatcher_ms; s_num += C.mixed_acc_numeric_ms;
      delete_vlength_cluster_matrix(C);
    }
    const double avg = total / (iterations ? iterations : 1);
    cout << "Average time (LeSpGEMM_VLength): " << avg << " ms" << endl;
    cout << "Average GFLOPS: " << (flops / 1e9) / (avg / 1000.0) << endl;
    if (print_bd && kernel == 3 && iterations > 0) {
      const double inv = 1.0 / iterations;
      cout << "[mixed_acc] phase breakdown avg over " << iterations << " runs (ms): setup=" << s_setup * inv
           << " malloc=" << s_malloc * inv << " symbolic=" << s_sym * inv << " dispatcher=" << s_disp * inv
           << " numeric=" << s_num * inv << " | sum=" << (s_setup + s_malloc + s_sym + s_disp + s_num) * inv << endl;
    }
    cout << "Performance test done." << endl;
  }
  if (kernel == 3) {
    delete_array(C.acc_flag);
    delete_array(C.min_ccol);
    delete_array(C.max_ccol);
  }
  delete_vlength_cluster_matrix(Acl);
  delete_host_matrix(A);
  delete_host_matrix(B);
}

int main(int argc, char **argv) {
  if (get_arg(argc, argv, "help") != nullptr || argc < 3) { usage(argv); return EXIT_SUCCESS; }
  const char *pa = nullptr, *pb = nullptr;
  for (int i = 1; i < argc; i++)
    if (argv[i][0] != '-') { if (!pa) pa = argv[i]; else if (!pb) { pb = argv[i]; break; } }
  if (!pa || !pb) { printf("Error: need A.mtx B.mtx\n"); usage(argv); return EXIT_SUCCESS; }
  char *p;
  int precision = (p = get_argval(argc, argv, "precision")) ? atoi(p) : 64;
  int threads = (p = get_argval(argc, argv, "threads")) ? atoi(p) : 1;
  int iterations = (p = get_argval(argc, argv, "iterations")) ? atoi(p) : 10;
  int kernel = (p = get_argval(argc, argv, "kernel")) ? atoi(p) : 1;
  if (kernel != 1 && kernel != 2 && kernel != 3) kernel = 1;
  long long cs = (p = get_argval(argc, argv, "cluster_size")) ? atoll(p) : 8;
  bool sortOutput = (p = get_argval(argc, argv, "sort")) ? atoi(p) != 0 : false;
  int k = (p = get_argval(argc, argv, "k")) ? atoi(p) : 64;
  int bands = (p = get_argval(argc, argv, "bands")) ? atoi(p) : 16;
  uint64_t seed = (p = get_argval(argc, argv, "seed")) ? (uint64_t)atoll(p) : 7;
  int hc_v = (p = get_argval(argc, argv, "hc_v")) ? atoi(p) : 0;
  if (hc_v != 0 && hc_v != 1 && hc_v != 2) hc_v = 0;
  double l2f = (p = get_argval(argc, argv, "l2_fraction")) ? atof(p) : -1.0;
  const char *tt = (p = get_argval(argc, argv, "test_type")) ? p : "performance";
  bool print_bd = (p = get_argval(argc, argv, "print_bd")) ? atoi(p) != 0 : false;
  printf("Precision: %d-bit, threads: %d\n\n", precision, threads);
  cout << "A: " << pa << ", B: " << pb << endl;
  cout << "test_type: " << tt << ", LSH: k=" << k << ", bands=" << bands << ", seed=" << seed << ", hc_v: " << hc_v << endl;
  cout << "kernel: " << kernel << ", cluster_size: " << cs << endl;
  const bool corr = strcmp(tt, "correctness") == 0;
  if (precision == 32) run<int64_t, float>(pa, pb, corr, iterations, kernel, sortOutput, cs, k, bands, seed, hc_v, l2f, print_bd);
  else if (precision == 64) run<int64_t, double>(pa, pb, corr, iterations, kernel, sortOutput, cs, k, bands, seed, hc_v, l2f, print_bd);
  else { usage(argv); return EXIT_FAILURE; }
  return EXIT_SUCCESS;
}
