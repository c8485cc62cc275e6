// benchmark_spmv_<fmt>: the reference's SpMV benchmark drivers (test/benchmark_spmv_csr.cpp,
// _ell.cpp, _sell_c_sigma.cpp, _sell_c_R.cpp, _bsr.cpp) on the drop-in headers, one source
// compiled once per format (-DSPOPS_BENCH_FORMAT_csr|ell|sell_c_sigma|sell_c_R|bsr).
//
//   benchmark_spmv_<fmt> my_matrix.mtx [--matID=0] [--Index=0|1] [--precision=32|64]
//       [--C=32] [--sigma=1024] [--blockDim=16] [--threads=N] [--sche=0..3]
//       [--perf_file=./performance/mat_perf.txt]
//
// Steps, as in the reference's utils/test_spmv_<fmt>.cpp: read the matrix, convert to the
// format (conversion time reported), check y = A*x against a serial host CSR product
// (max relative difference, the reference's tolerance check), then time the SpMV.  Two
// timings are printed: the drop-in call SpOps_<fmt> on host arrays (PCIe copies of x and y
// included, what a caller of the reference's API sees) and the device-resident kernel
// (spops_spmv on HBM-resident x and y, HIP events).  GFLOP/s = 2*nnz / time.  One line per
// run is appended to the performance file in the reference's column layout.
// --threads and --sche are accepted and ignored (the GPU kernels have no OpenMP schedule).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "SpOps.h"
#include "cmdline.h"
#include "sparse_io.h"

#if defined(SPOPS_BENCH_FORMAT_csr)
static const char *kFormat = "CSR";
#elif defined(SPOPS_BENCH_FORMAT_ell)
static const char *kFormat = "ELL";
#elif defined(SPOPS_BENCH_FORMAT_sell_c_sigma)
static const char *kFormat = "S-ELL-sigma";
#elif defined(SPOPS_BENCH_FORMAT_sell_c_R)
static const char *kFormat = "S-ELL-R";
#elif defined(SPOPS_BENCH_FORMAT_bsr)
static const char *kFormat = "BSR";
#else
#error "define SPOPS_BENCH_FORMAT_<csr|ell|sell_c_sigma|sell_c_R|bsr>"
#endif

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void usage(char **argv) {
  std::cout << "Usage:\n\t" << argv[0] << " my_matrix.mtx [options]\n"
            << "\t --matID     = m_num, the matrix ID number in the dataset (default 0)\n"
            << "\t --Index     = 0 (int, default) or 1 (long long)\n"
            << "\t --precision = 32 (default) or 64\n"
            << "\t --C         = SELL chunk height (default 32)\n"
            << "\t --sigma     = SELL-C-sigma sorting scope (default 1024)\n"
            << "\t --blockDim  = BSR block size (default 16)\n"
            << "\t --perf_file = file the result line is appended to (default ./performance/mat_perf.txt)\n"
            << "Note: my_matrix.mtx must be a real-valued sparse matrix in the MatrixMarket format.\n";
}

struct Opts {
  int matID = 0, C = 32, sigma = 1024, block = 16;
  std::string perf = "./performance/mat_perf.txt";
};

// Matrix in the benchmarked format plus the drop-in call on it.
template <typename I, typename V>
struct Formatted {
#if defined(SPOPS_BENCH_FORMAT_csr)
  CSR_Matrix<I, V> m;
  void make(const CSR_Matrix<I, V> &csr, const Opts &) {
    m = csr;
    m.row_offset = copy_array(csr.row_offset, (size_t)csr.num_rows + 1);
    m.col_index = copy_array(csr.col_index, (size_t)csr.num_nnzs);
    m.values = copy_array(csr.values, (size_t)csr.num_nnzs);
  }
  void run(V a, const V *x, V b, V *y) { SpOps_csr<I, V>(a, m, x, b, y); }
  const void *key() const { return m.row_offset; }
  void free_() { delete_host_matrix(m); }
#elif defined(SPOPS_BENCH_FORMAT_ell)
  ELL_Matrix<I, V> m;
  void make(const CSR_Matrix<I, V> &csr, const Opts &) { m = csr_to_ell(csr, RowMajor); }
  void run(V a, const V *x, V b, V *y) { SpOps_ell<I, V>(a, m, x, b, y); }
  const void *key() const { return m.col_index; }
  void free_() { delete_host_matrix(m); }
#elif defined(SPOPS_BENCH_FORMAT_sell_c_sigma)
  SELL_C_Sigma_Matrix<I, V> m;
  void make(const CSR_Matrix<I, V> &csr, const Opts &o) { m = csr_to_sell_c_sigma(csr, nullptr, o.sigma, o.C); }
  void run(V a, const V *x, V b, V *y) { SpOps_sell_c_sigma<I, V>(a, m, x, b, y); }
  const void *key() const { return m.reorder; }
  void free_() { delete_host_matrix(m); }
#elif defined(SPOPS_BENCH_FORMAT_sell_c_R)
  SELL_C_R_Matrix<I, V> m;
  void make(const CSR_Matrix<I, V> &csr, const Opts &o) { m = csr_to_sell_c_R(csr, nullptr, o.C); }
  void run(V a, const V *x, V b, V *y) { SpOps_sell_c_R<I, V>(a, m, x, b, y); }
  const void *key() const { return m.reorder; }
  void free_() { delete_host_matrix(m); }
#elif defined(SPOPS_BENCH_FORMAT_bsr)
  BSR_Matrix<I, V> m;
  void make(const CSR_Matrix<I, V> &csr, const Opts &o) { m = csr_to_bsr(csr, (I)o.block, (I)o.block); }
  void run(V a, const V *x, V b, V *y) { SpOps_bsr<I, V>(a, m, x, b, y); }
  const void *key() const { return m.row_ptr; }
  void free_() { delete_host_matrix(m); }
#endif
};

template <typename I, typename V>
static int run(const char *path, const Opts &o) {
  const std::string name = extractFileNameWithoutExtension(path);
  CSR_Matrix<I, V> csr = read_csr_matrix<I, V>(path);
  std::printf("Using %lld-by-%lld matrix with %lld nonzero values\n", (long long)csr.num_rows, (long long)csr.num_cols,
              (long long)csr.num_nnzs);
  std::fflush(stdout);
  std::cout << "=====  Testing " << kFormat << " Kernels (gfx950)  =====" << std::endl;

  double t0 = now_ms();
  Formatted<I, V> F;
  F.make(csr, o);
  const double convert_ms = now_ms() - t0;
  std::printf("\tFormat conversion: %8.4f ms\n", convert_ms);

  std::vector<V> x(csr.num_cols), y(csr.num_rows, V(0)), yref(csr.num_rows, V(0));
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (auto &v : x) v = (V)U(rng);
  // serial host CSR product in double, the check the reference runs against its CSR kernel
  double worst = 0.0;
  {
    F.run(V(1), x.data(), V(0), y.data());
    for (I r = 0; r < csr.num_rows; r++) {
      double s = 0.0, a = 0.0;
      for (I j = csr.row_offset[r]; j < csr.row_offset[r + 1]; j++) {
        s += (double)csr.values[j] * (double)x[csr.col_index[j]];
        a += std::fabs((double)csr.values[j] * (double)x[csr.col_index[j]]);
      }
      const double d = std::fabs((double)y[r] - s) / (a > 0.0 ? a : 1.0);
      worst = d > worst ? d : worst;
    }
  }
  const double tol = std::is_same<V, float>::value ? 1e-5 : 1e-12;
  std::printf("\tmax |y - y_csr| / sum|a*x| = %.3e  (%s, tolerance %.0e)\n", worst, worst <= tol ? "PASSED" : "FAILED",
              tol);

  // drop-in timing on host arrays: 1 warm-up call, then enough calls for about 2 s
  t0 = now_ms();
  F.run(V(1), x.data(), V(0), y.data());
  const double est = now_ms() - t0;
  const int iters = est < 1.0 ? 1000 : std::max(20, std::min(1000, (int)(2000.0 / est)));
  t0 = now_ms();
  for (int i = 0; i < iters; i++) F.run(V(1), x.data(), V(0), y.data());
  const double host_ms = (now_ms() - t0) / iters;

  // device-resident timing of the same cached device matrix
  double dev_ms = 0.0;
  {
    using namespace spops_dropin;
    std::lock_guard<std::mutex> g(state().mu);
    spops_matrix A = state().entries.at(F.key())->m;
    void *dx = nullptr, *dy = nullptr;
    check(spops_malloc(&dx, (size_t)(csr.num_cols ? csr.num_cols : 1) * sizeof(V)), "malloc x");
    check(spops_malloc(&dy, (size_t)(csr.num_rows ? csr.num_rows : 1) * sizeof(V)), "malloc y");
    check(spops_memcpy_htod(dx, x.data(), (size_t)csr.num_cols * sizeof(V), nullptr), "copy x");
    void *e0 = nullptr, *e1 = nullptr;
    check(spops_event_create(&e0), "event");
    check(spops_event_create(&e1), "event");
    for (int i = 0; i < 5; i++) check(spops_spmv(A, 1.0, dx, 0.0, dy, nullptr), "spmv");
    const int diters = 200;
    check(spops_event_record(e0, nullptr), "record");
    for (int i = 0; i < diters; i++) check(spops_spmv(A, 1.0, dx, 0.0, dy, nullptr), "spmv");
    check(spops_event_record(e1, nullptr), "record");
    check(spops_stream_synchronize(nullptr), "sync");
    float ms = 0.f;
    check(spops_event_elapsed_ms(&ms, e0, e1), "elapsed");
    dev_ms = ms / diters;
    spops_event_destroy(e0);
    spops_event_destroy(e1);
    spops_free(dx);
    spops_free(dy);
  }
  const double nnz2 = 2.0 * (double)csr.num_nnzs;
  std::printf("\tbenchmarking %-20s [%s]: %8.4f ms ( %5.4f GFLOP/s )  %d iterations\n", kFormat, "host arrays", host_ms,
              nnz2 / (host_ms * 1e6), iters);
  std::printf("\tbenchmarking %-20s [%s]: %8.4f ms ( %5.4f GFLOP/s )\n", kFormat, "gfx950 HBM", dev_ms,
              nnz2 / (dev_ms * 1e6));
  if (FILE *f = std::fopen(o.perf.c_str(), "a")) {
    // Mat Format Method Schedule c sigma Time Performance Convert
    std::fprintf(f, "%d %s %s %d %d %d %d %8.4f %5.4f %5.4f \n", o.matID, name.c_str(), kFormat, 2, 0, o.C, o.sigma, dev_ms,
                 nnz2 / (dev_ms * 1e6), convert_ms);
    std::fclose(f);
  } else {
    std::cout << "Unable to open perf-saved file: " << o.perf << std::endl;
  }
  spops_dropin_release(F.key());
  F.free_();
  delete_host_matrix(csr);
  return worst <= tol ? EXIT_SUCCESS : EXIT_FAILURE;
}

int main(int argc, char **argv) {
  if (get_arg(argc, argv, "help") != nullptr) {
    usage(argv);
    return EXIT_SUCCESS;
  }
  const char *path = nullptr;
  for (int i = 1; i < argc; i++)
    if (argv[i][0] != '-') {
      path = argv[i];
      break;
    }
  if (!path) {
    std::printf("You need to input a matrix file!\n");
    return EXIT_FAILURE;
  }
  char *p;
  Opts o;
  const int precision = (p = get_argval(argc, argv, "precision")) ? std::atoi(p) : 32;
  const int index = (p = get_argval(argc, argv, "Index")) ? std::atoi(p) : 0;
  if ((p = get_argval(argc, argv, "matID"))) o.matID = std::atoi(p);
  if ((p = get_argval(argc, argv, "C"))) o.C = std::atoi(p);
  if ((p = get_argval(argc, argv, "sigma"))) o.sigma = std::atoi(p);
  if ((p = get_argval(argc, argv, "blockDim"))) o.block = std::atoi(p);
  if ((p = get_argval(argc, argv, "perf_file"))) o.perf = p;
  std::printf("\nUsing %d-bit floating point precision, %d-bit Index, gfx950\n\n", precision, (index + 1) * 32);
  if (index == 0 && precision == 32) return run<int, float>(path, o);
  if (index == 0 && precision == 64) return run<int, double>(path, o);
  if (index == 1 && precision == 32) return run<long long, float>(path, o);
  if (index == 1 && precision == 64) return run<long long, double>(path, o);
  usage(argv);
  return EXIT_FAILURE;
}
