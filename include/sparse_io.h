/* MatrixMarket input/output for the drop-in API with the reference's semantics
 * (utils/sparse_io.cpp:26-214): 1-based indices become 0-based, pattern entries read as
 * 1.0, symmetric files mirror each off-diagonal entry, entries keep file order inside a
 * row, duplicates are kept unless compact.  Errors print and exit(1) as the reference does.
 * write_mtx_crd writes %.17g so values round-trip (the reference's %20.16g does not). */
#ifndef SPOPS_DROPIN_SPARSE_IO_H
#define SPOPS_DROPIN_SPARSE_IO_H

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "sparse_conversion.h"
#include "sparse_format.h"

inline std::string extractFileNameWithoutExtension(const std::string &path) {
  size_t s = path.find_last_of("/\\");
  std::string f = (s == std::string::npos) ? path : path.substr(s + 1);
  size_t d = f.find_last_of('.');
  return d == std::string::npos ? f : f.substr(0, d);
}

template <class I, class V>
COO_Matrix<I, V> read_coo_matrix(const char *mm_filename) {
  COO_Matrix<I, V> coo;
  FILE *f = std::fopen(mm_filename, "r");
  if (!f) {
    std::cout << "Unable to open file: " << mm_filename << std::endl;
    std::exit(1);
  }
  char line[1024];
  if (!std::fgets(line, sizeof line, f)) {
    std::cout << "Could not process Matrix Market banner." << std::endl;
    std::exit(1);
  }
  char b0[64], obj[64], fmt[64], field[64], sym[64];
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", b0, obj, fmt, field, sym) != 5 || std::strcmp(b0, "%%MatrixMarket")) {
    std::cout << "Could not process Matrix Market banner." << std::endl;
    std::exit(1);
  }
  for (char *p : {obj, fmt, field, sym})
    for (; *p; ++p) *p = (char)std::tolower(*p);
  const bool pattern = !std::strcmp(field, "pattern");
  if (std::strcmp(obj, "matrix") || std::strcmp(fmt, "coordinate") ||
      !(pattern || !std::strcmp(field, "real") || !std::strcmp(field, "integer"))) {
    std::printf("Sorry, this application does not support Market Market type: [%s %s %s]\n", fmt, field, sym);
    std::printf("Only sparse real-valued or pattern coordinate matrices are supported\n");
    std::exit(1);
  }
  long long nr = 0, nc = 0, nz = 0;
  do {
    if (!std::fgets(line, sizeof line, f)) {
      std::cout << "The line of rows, cols, nnzs is in wrong format" << std::endl;
      std::exit(1);
    }
  } while (line[0] == '%');
  if (std::sscanf(line, "%lld %lld %lld", &nr, &nc, &nz) != 3) {
    std::cout << "The line of rows, cols, nnzs is in wrong format" << std::endl;
    std::exit(1);
  }
  const bool symmetric = !std::strcmp(sym, "symmetric");
  std::vector<I> ri, ci;
  std::vector<V> va;
  ri.reserve(nz);
  ci.reserve(nz);
  va.reserve(nz);
  for (long long t = 0; t < nz; ++t) {
    long long r, c;
    double v = 1.0;
    int got = pattern ? std::fscanf(f, "%lld %lld", &r, &c) : std::fscanf(f, "%lld %lld %lf", &r, &c, &v);
    if (got != (pattern ? 2 : 3)) {
      std::cout << "Error reading entry " << t << " of " << mm_filename << std::endl;
      std::exit(1);
    }
    ri.push_back((I)(r - 1));
    ci.push_back((I)(c - 1));
    va.push_back((V)v);
    if (symmetric && r != c) {
      ri.push_back((I)(c - 1));
      ci.push_back((I)(r - 1));
      va.push_back((V)v);
    }
  }
  std::fclose(f);
  coo.num_rows = (I)nr;
  coo.num_cols = (I)nc;
  coo.num_nnzs = (I)ri.size();
  coo.tag = 0;
  coo.kernel_flag = 1;
  coo.partition = nullptr;
  coo.row_index = spops_checked(new_array<I>(ri.size()));
  coo.col_index = spops_checked(new_array<I>(ci.size()));
  coo.values = spops_checked(new_array<V>(va.size()));
  std::copy(ri.begin(), ri.end(), coo.row_index);
  std::copy(ci.begin(), ci.end(), coo.col_index);
  std::copy(va.begin(), va.end(), coo.values);
  return coo;
}

template <class I, class V>
CSR_Matrix<I, V> read_csr_matrix(const char *mm_filename, bool compact = false) {
  COO_Matrix<I, V> coo = read_coo_matrix<I, V>(mm_filename);
  CSR_Matrix<I, V> csr = coo_to_csr(coo, compact);
  delete_host_matrix(coo);
  csr.kernel_flag = 1;
  return csr;
}

template <class I, class V>
int write_mtx_crd(const char *path, const CSR_Matrix<I, V> &A) {
  FILE *f = std::fopen(path, "w");
  if (!f) return 1;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%lld %lld %lld\n", (long long)A.num_rows,
               (long long)A.num_cols, (long long)A.num_nnzs);
  for (I r = 0; r < A.num_rows; r++)
    for (I j = A.row_offset[r]; j < A.row_offset[r + 1]; j++)
      std::fprintf(f, "%lld %lld %.17g\n", (long long)r + 1, (long long)A.col_index[j] + 1, (double)A.values[j]);
  std::fclose(f);
  return 0;
}

#endif
