// Reads a Matrix Market file with the drop-in read_csr_matrix<int64_t, double>(path, compact)
// (include/sparse_io.h) and prints "rows cols nnz", then the row pointer, the columns and the
// values (%.17g), one array per line: the test compares them with the reference reader's output.
#include <cstdio>
#include <cstdlib>

#include "sparse_io.h"

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  CSR_Matrix<int64_t, double> A = read_csr_matrix<int64_t, double>(argv[1], std::atoi(argv[2]) != 0);
  std::printf("%lld %lld %lld\n", (long long)A.num_rows, (long long)A.num_cols, (long long)A.num_nnzs);
  for (int64_t i = 0; i <= A.num_rows; ++i) std::printf("%lld ", (long long)A.row_offset[i]);
  std::printf("\n");
  for (int64_t i = 0; i < A.num_nnzs; ++i) std::printf("%lld ", (long long)A.col_index[i]);
  std::printf("\n");
  for (int64_t i = 0; i < A.num_nnzs; ++i) std::printf("%.17g ", A.values[i]);
  std::printf("\n");
  delete_host_matrix(A);
  return 0;
}
