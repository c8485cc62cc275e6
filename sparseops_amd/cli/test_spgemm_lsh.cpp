// test_spgemm_lsh: the reference's LeSpGEMM driver on saved clusterings
// (test/test_spgemm_lsh.cpp): reads A, B and <lsh-order-dir>/<A name>.perm / .offsets (as
// record_permutation_lsh writes them), reorders A, builds the cluster format and times
// LeSpGEMM_VLength kernel 3 only, on the drop-in headers (GPU).  Same options and stdout lines.
//   test_spgemm_lsh A.mtx B.mtx [--lsh-order-dir=DIR] [--precision=32|64] [--threads=N]
//       [--iterations=10] [--l2_fraction=f] [--sort=0|1]
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "SpOps.h"
#include "cmdline.h"
#include "sparse_io.h"

using namespace std;

#ifndef LSH_ORDER_OUT_DIR
#define LSH_ORDER_OUT_DIR "./lsh_order"
#endif

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void usage(char **argv) {
  cout << "Usage:\n\t" << argv[0] << " <A.mtx> <B.mtx> [options]\n"
       << "\t  Reads permutation and offsets from lsh_order/<mtx_name>.perm and .offsets (mtx_name from A path).\n"
       << "\t  Runs LeSpGEMM_VLength kernel=3 only.\n"
       << "\tOptions:\n"
       << "\t  --lsh-order-dir = directory for .perm/.offsets (default: " LSH_ORDER_OUT_DIR ")\n"
       << "\t  --precision     = 32|64 (default 64)\n"
       << "\t  --threads      = OMP threads (accepted, unused)\n"
       << "\t  --iterations   = iterations for timing (default 10)\n"
       << "\t  --l2_fraction  = L2 fraction for kernel=3 (classification report only)\n"
       << "\t  --sort         = 0|1 (sort output columns, default 0)\n";
}

// offsets file: one integer per line (test/test_spgemm_lsh.cpp:31-44)
template <typename I>
static vector<I> read_offsets_file(const char *filename) {
  vector<I> out;
  ifstream f(filename);
  if (!f.is_open()) {
    cerr << "Error: cannot open offsets file: " << filename << endl;
    throw runtime_error("Cannot open offsets file");
  }
  I x;
  while (f >> x) out.push_back(x);
  return out;
}

template <typename I, typename V>
static int run(const char *pa, const char *pb, const string &dir, int iterations, bool sortOutput, double l2f) {
  const int kernel = 3;
  cout << "========================================" << endl;
  cout << "LSH-order SpGEMM (LeSpGEMM_VLength kernel=3)" << endl;
  cout << "========================================" << endl;
  cout << "Reading A: " << pa << endl;
  CSR_Matrix<I, V> A = read_csr_matrix<I, V>(pa);
  cout << "A: " << A.num_rows << " x " << A.num_cols << ", nnz: " << A.num_nnzs << endl;
  cout << "Reading B: " << pb << endl;
  CSR_Matrix<I, V> B = read_csr_matrix<I, V>(pb);
  cout << "B: " << B.num_rows << " x " << B.num_cols << ", nnz: " << B.num_nnzs << endl;
  if (A.num_cols != B.num_rows) {
    cerr << "Error: A.cols != B.rows" << endl;
    delete_host_matrix(A);
    delete_host_matrix(B);
    return EXIT_FAILURE;
  }
  const string name = extractFileNameWithoutExtension(string(pa));
  const string perm_path = dir + "/" + name + ".perm", off_path = dir + "/" + name + ".offsets";
  cout << "Reading permutation: " << perm_path << endl;
  vector<I> perm = read_row_permutation<I>(perm_path.c_str(), A.num_rows);
  if ((I)perm.size() != A.num_rows) {
    cerr << "Error: permutation size (" << perm.size() << ") != A.num_rows (" << A.num_rows << ")" << endl;
    delete_host_matrix(A);
    delete_host_matrix(B);
    return EXIT_FAILURE;
  }
  cout << "Reading offsets: " << off_path << endl;
  vector<I> offsets = read_offsets_file<I>(off_path.c_str());
  if (offsets.empty() || offsets.front() != 0 || offsets.back() != A.num_rows) {
    cerr << "Error: offsets must be non-empty, start with 0 and end with A.num_rows. Got " << offsets.size()
         << " entries, first=" << (offsets.empty() ? -1 : (int64_t)offsets.front())
         << " last=" << (offsets.empty() ? -1 : (int64_t)offsets.back()) << endl;
    delete_host_matrix(A);
    delete_host_matrix(B);
    return EXIT_FAILURE;
  }
  long long flops = get_spgemm_flop(A, B);
  cout << "FLOPs: " << flops << endl;
  cout << "Converting A to CSR_VlengthCluster..." << endl;
  double t0 = now_ms();
  CSR_Matrix<I, V> Ar = csr_reorder_rows<I, V>(A, perm);
  CSR_VlengthCluster<I, V> Acl = csr_to_vlength_cluster<I, V>(Ar, offsets);
  cout << "Format Conversion time: " << now_ms() - t0 << " ms" << endl;
  cout << "Clusters: " << (offsets.size() - 1) << endl;
  delete_host_matrix(Ar);
  delete_host_matrix(A);
  cout << "A_cluster: " << Acl.rows << " clusters, nnzc=" << Acl.nnzc << ", nnzv=" << Acl.nnzv << endl;
  CSR_VlengthCluster<I, V> C;
  C.acc_flag = new_array<char>(Acl.rows);
  C.min_ccol = new_array<I>(Acl.rows);
  C.max_ccol = new_array<I>(Acl.rows);
  auto spgemm = [&]() {
    if (sortOutput) LeSpGEMM_VLength<true, I, V>(Acl, B, C, kernel, l2f);
    else LeSpGEMM_VLength<false, I, V>(Acl, B, C, kernel, l2f);
  };
  spgemm();
  cout << "[mixed_acc] dense clusters: " << C.dense_cluster_count << " / " << C.rows << " ("
       << (100.0 * C.dense_cluster_count / (C.rows ? C.rows : 1)) << "%)" << endl;
  delete_vlength_cluster_matrix(C);
  double total = 0;
  for (int i = 0; i < iterations; i++) {
    t0 = now_ms();
    spgemm();
    total += now_ms() - t0;
    if (i < iterations - 1) delete_vlength_cluster_matrix(C);
  }
  const double avg = total / (iterations ? iterations : 1);
  cout << "Average time (LeSpGEMM_VLength kernel=3): " << avg << " ms" << endl;
  cout << "Average GFLOPS: " << (flops / 1e9) / (avg / 1000.0) << endl;
  if (iterations > 0) delete_vlength_cluster_matrix(C);
  delete_array(C.acc_flag);
  delete_array(C.min_ccol);
  delete_array(C.max_ccol);
  delete_vlength_cluster_matrix(Acl);
  delete_host_matrix(B);
  cout << "Done." << endl;
  return EXIT_SUCCESS;
}

int main(int argc, char **argv) {
  if (get_arg(argc, argv, "help") != nullptr || argc < 3) { usage(argv); return EXIT_SUCCESS; }
  const char *pa = nullptr, *pb = nullptr;
  for (int i = 1; i < argc; i++)
    if (argv[i][0] != '-') { if (!pa) pa = argv[i]; else if (!pb) { pb = argv[i]; break; } }
  if (!pa || !pb) { cerr << "Error: need A.mtx B.mtx\n"; usage(argv); return EXIT_SUCCESS; }
  char *p;
  int threads = (p = get_argval(argc, argv, "threads")) ? atoi(p) : 1;
  string dir = (p = get_argval(argc, argv, "lsh-order-dir")) ? p : LSH_ORDER_OUT_DIR;
  int iterations = (p = get_argval(argc, argv, "iterations")) ? atoi(p) : 10;
  bool sortOutput = (p = get_argval(argc, argv, "sort")) ? atoi(p) != 0 : false;
  double l2f = (p = get_argval(argc, argv, "l2_fraction")) ? atof(p) : -1.0;
  int precision = (p = get_argval(argc, argv, "precision")) ? atoi(p) : 64;
  cout << "Precision: " << precision << "-bit, threads: " << threads << endl;
  if (precision == 32) return run<int64_t, float>(pa, pb, dir, iterations, sortOutput, l2f);
  if (precision == 64) return run<int64_t, double>(pa, pb, dir, iterations, sortOutput, l2f);
  usage(argv);
  return EXIT_FAILURE;
}
