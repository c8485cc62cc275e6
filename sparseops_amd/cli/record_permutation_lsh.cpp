// record_permutation_lsh: the reference's clustering recorder (test/record_permutation_lsh.cpp)
// on the drop-in headers, i.e. MinHash + LSH on the GPU and the bit-exact host HC.  Writes the
// row permutation and the cluster offsets of A, one integer per line, to
// <out-dir>/<mtx name>.perm and .offsets (:160-181), for test_spgemm_lsh to load.
//   record_permutation_lsh A.mtx [--out-dir=DIR] [--k=64] [--bands=16] [--seed=7] [--hc_v=0|1|2]
//       [--cluster_size=8] [--save=0|1] [--threads=N] [--precision=32|64]
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <sys/stat.h>
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
  cout << "Usage:\n\t" << argv[0] << " <A.mtx> [options]\n"
       << "\t  Runs LSH + HC, writes permutation and offsets to out_dir/mtx_name.perm and .offsets\n"
       << "\tOptions:\n"
       << "\t  --out-dir     = output directory (default: " LSH_ORDER_OUT_DIR ")\n"
       << "\t  --k           = MinHash signature length (default 64)\n"
       << "\t  --bands       = LSH bands, k % bands == 0 (default 16)\n"
       << "\t  --seed        = random seed (default 7)\n"
       << "\t  --hc_v        = 0|1|2 (0=C-LSH+fast v0 HC, 1=naive LSH+v1 HC, 2=naive LSH+fast v0 HC, default 0)\n"
       << "\t  --cluster_size= max cluster size (default 8)\n"
       << "\t  --save        = 0|1, write .perm/.offsets (default 1)\n"
       << "\t  --threads     = number of OMP threads (accepted, unused)\n"
       << "\t  --precision   = 32|64 (default 64)\n";
}

static void make_dirs(const string &dir) {
  string cur;
  for (size_t i = 0; i <= dir.size(); ++i) {
    if (i == dir.size() || dir[i] == '/') {
      if (!cur.empty()) mkdir(cur.c_str(), 0755);
    }
    if (i < dir.size()) cur += dir[i];
  }
}

template <typename I, typename V>
static int run(const char *matA, const string &out_dir, int k, int bands, uint64_t seed, int hc_v, I cs, bool save) {
  cout << "Reading A: " << matA << endl;
  CSR_Matrix<I, V> A = read_csr_matrix<I, V>(matA);
  cout << "A: " << A.num_rows << " x " << A.num_cols << ", nnz: " << A.num_nnzs << endl;
  if (k <= 0 || bands <= 0 || k % bands != 0) {
    cerr << "Error: k % bands == 0 required. k=" << k << " bands=" << bands << endl;
    delete_host_matrix(A);
    return EXIT_FAILURE;
  }
  cout << "LSH + HC (k=" << k << ", bands=" << bands << ", seed=" << seed << ", hc_v=" << hc_v
       << ", save=" << (save ? 1 : 0) << ")..." << endl;
  double t0 = now_ms();
  vector<CandidatePair<I, V>> pairs =
      lsh_candidate_pairs_vector<I, V>(A.row_offset, A.col_index, A.num_rows, k, bands, seed, hc_v == 0);
  const double t_lsh = now_ms() - t0;
  cout << "LSH candidate pairs: " << pairs.size() << " (genPairs time: " << t_lsh << " ms)" << endl;
  vector<I> perm, offsets;
  cout << "Hierarchical clustering (cluster_size=" << cs << ") [" << (hc_v == 1 ? "v1" : "fast v0") << "]..." << endl;
  t0 = now_ms();
  if (hc_v == 1) hierarchical_clustering_v1<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, cs, perm, offsets);
  else hierarchical_clustering_v0_fast<I, V>(A.row_offset, A.col_index, A.num_rows, pairs, cs, perm, offsets);
  const double t_hc = now_ms() - t0;
  cout << "Clusters: " << (offsets.empty() ? 0 : offsets.size() - 1) << " (HC time: " << t_hc << " ms)" << endl;
  const I n = A.num_rows;
  delete_host_matrix(A);
  if ((I)perm.size() != n) {
    cerr << "Error: permutation size (" << perm.size() << ") != A.num_rows (" << n << ")" << endl;
    return EXIT_FAILURE;
  }
  if (!save) {
    cout << "Save disabled (--save=0); no permutation or offsets files written." << endl;
    return EXIT_SUCCESS;
  }
  const string name = extractFileNameWithoutExtension(string(matA));
  make_dirs(out_dir);
  const string perm_path = out_dir + "/" + name + ".perm", off_path = out_dir + "/" + name + ".offsets";
  {
    ofstream f(perm_path);
    if (!f) { cerr << "Error: cannot write " << perm_path << endl; return EXIT_FAILURE; }
    for (I x : perm) f << x << "\n";
  }
  {
    ofstream f(off_path);
    if (!f) { cerr << "Error: cannot write " << off_path << endl; return EXIT_FAILURE; }
    for (I x : offsets) f << x << "\n";
  }
  cout << "Wrote " << perm_path << " (" << perm.size() << " entries)" << endl;
  cout << "Wrote " << off_path << " (" << offsets.size() << " entries)" << endl;
  return EXIT_SUCCESS;
}

int main(int argc, char **argv) {
  if (get_arg(argc, argv, "help") != nullptr || argc < 2) { usage(argv); return EXIT_SUCCESS; }
  const char *matA = nullptr;
  for (int i = 1; i < argc; i++)
    if (argv[i][0] != '-') { matA = argv[i]; break; }
  if (!matA) { cerr << "Error: need A.mtx\n"; usage(argv); return EXIT_SUCCESS; }
  char *p;
  int precision = (p = get_argval(argc, argv, "precision")) ? atoi(p) : 64;
  int threads = (p = get_argval(argc, argv, "threads")) ? atoi(p) : 1;
  string out_dir = (p = get_argval(argc, argv, "out-dir")) ? p : LSH_ORDER_OUT_DIR;
  long long cs = (p = get_argval(argc, argv, "cluster_size")) ? atoll(p) : 8;
  int k = (p = get_argval(argc, argv, "k")) ? atoi(p) : 64;
  int bands = (p = get_argval(argc, argv, "bands")) ? atoi(p) : 16;
  uint64_t seed = (p = get_argval(argc, argv, "seed")) ? (uint64_t)atoll(p) : 7;
  int hc_v = (p = get_argval(argc, argv, "hc_v")) ? atoi(p) : 0;
  if (hc_v != 0 && hc_v != 1 && hc_v != 2) hc_v = 0;
  bool save = (p = get_argval(argc, argv, "save")) ? atoi(p) != 0 : true;
  cout << "Precision: " << precision << "-bit, threads: " << threads << endl;
  if (precision == 32) return run<int64_t, float>(matA, out_dir, k, bands, seed, hc_v, cs, save);
  if (precision == 64) return run<int64_t, double>(matA, out_dir, k, bands, seed, hc_v, cs, save);
  usage(argv);
  return EXIT_FAILURE;
}
