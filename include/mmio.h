/* Matrix Market I/O with the API of the public NIST Matrix Market library, which the
 * reference's callers use through include/mmio.h:1-135 (utils/mmio.cpp): the 4-character
 * MM_typecode (object 'M'; 'C' coordinate / 'A' array; 'R' real, 'C' complex, 'P' pattern,
 * 'I' integer; 'G' general, 'S' symmetric, 'K' skew, 'H' hermitian), its mm_is_* / mm_set_*
 * macros, the MM_* error codes, banner and size lines, and whole-file coordinate reads and
 * writes.  Written here as inline functions; the entry line format of mm_write_mtx_crd is the
 * library's ("%d %d %20.16g" for real values), so written files match the reference's byte
 * for byte.  (sparse_io.h holds the reference's read_csr_matrix on top of its own parser.) */
#ifndef SPOPS_DROPIN_MMIO_H
#define SPOPS_DROPIN_MMIO_H

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define MM_MAX_LINE_LENGTH 1025
#define MatrixMarketBanner "%%MatrixMarket"
#define MM_MAX_TOKEN_LENGTH 64

typedef char MM_typecode[4];

#define mm_is_matrix(typecode) ((typecode)[0] == 'M')
#define mm_is_sparse(typecode) ((typecode)[1] == 'C')
#define mm_is_coordinate(typecode) ((typecode)[1] == 'C')
#define mm_is_dense(typecode) ((typecode)[1] == 'A')
#define mm_is_array(typecode) ((typecode)[1] == 'A')
#define mm_is_complex(typecode) ((typecode)[2] == 'C')
#define mm_is_real(typecode) ((typecode)[2] == 'R')
#define mm_is_pattern(typecode) ((typecode)[2] == 'P')
#define mm_is_integer(typecode) ((typecode)[2] == 'I')
#define mm_is_symmetric(typecode) ((typecode)[3] == 'S')
#define mm_is_general(typecode) ((typecode)[3] == 'G')
#define mm_is_skew(typecode) ((typecode)[3] == 'K')
#define mm_is_hermitian(typecode) ((typecode)[3] == 'H')

#define mm_set_matrix(typecode) ((*typecode)[0] = 'M')
#define mm_set_coordinate(typecode) ((*typecode)[1] = 'C')
#define mm_set_array(typecode) ((*typecode)[1] = 'A')
#define mm_set_dense(typecode) mm_set_array(typecode)
#define mm_set_sparse(typecode) mm_set_coordinate(typecode)
#define mm_set_complex(typecode) ((*typecode)[2] = 'C')
#define mm_set_real(typecode) ((*typecode)[2] = 'R')
#define mm_set_pattern(typecode) ((*typecode)[2] = 'P')
#define mm_set_integer(typecode) ((*typecode)[2] = 'I')
#define mm_set_symmetric(typecode) ((*typecode)[3] = 'S')
#define mm_set_general(typecode) ((*typecode)[3] = 'G')
#define mm_set_skew(typecode) ((*typecode)[3] = 'K')
#define mm_set_hermitian(typecode) ((*typecode)[3] = 'H')
#define mm_clear_typecode(typecode) ((*typecode)[0] = (*typecode)[1] = (*typecode)[2] = ' ', (*typecode)[3] = 'G')
#define mm_initialize_typecode(typecode) mm_clear_typecode(typecode)

#define MM_COULD_NOT_READ_FILE 11
#define MM_PREMATURE_EOF 12
#define MM_NOT_MTX 13
#define MM_NO_HEADER 14
#define MM_UNSUPPORTED_TYPE 15
#define MM_LINE_TOO_LONG 16
#define MM_COULD_NOT_WRITE_FILE 17

#define MM_MTX_STR "matrix"
#define MM_ARRAY_STR "array"
#define MM_DENSE_STR "array"
#define MM_COORDINATE_STR "coordinate"
#define MM_SPARSE_STR "coordinate"
#define MM_COMPLEX_STR "complex"
#define MM_REAL_STR "real"
#define MM_INT_STR "integer"
#define MM_GENERAL_STR "general"
#define MM_SYMM_STR "symmetric"
#define MM_HERM_STR "hermitian"
#define MM_SKEW_STR "skew-symmetric"
#define MM_PATTERN_STR "pattern"

inline char *mm_strdup(const char *s) {
  const size_t n = std::strlen(s) + 1;
  char *d = static_cast<char *>(std::malloc(n));
  if (d) std::memcpy(d, s, n);
  return d;
}

// A typecode the format allows: a matrix; no dense pattern; no real hermitian; no hermitian
// or skew-symmetric pattern.
inline int mm_is_valid(MM_typecode matcode) {
  if (!mm_is_matrix(matcode)) return 0;
  if (mm_is_dense(matcode) && mm_is_pattern(matcode)) return 0;
  if (mm_is_real(matcode) && mm_is_hermitian(matcode)) return 0;
  if (mm_is_pattern(matcode) && (mm_is_hermitian(matcode) || mm_is_skew(matcode))) return 0;
  return 1;
}

namespace spops_dropin {
inline const char *mm_word(char c, int field) {
  switch (field) {
    case 1: return c == 'C' ? MM_SPARSE_STR : c == 'A' ? MM_DENSE_STR : nullptr;
    case 2: return c == 'R' ? MM_REAL_STR : c == 'C' ? MM_COMPLEX_STR : c == 'P' ? MM_PATTERN_STR
                 : c == 'I' ? MM_INT_STR : nullptr;
    case 3: return c == 'G' ? MM_GENERAL_STR : c == 'S' ? MM_SYMM_STR : c == 'H' ? MM_HERM_STR
                 : c == 'K' ? MM_SKEW_STR : nullptr;
    default: return c == 'M' ? MM_MTX_STR : nullptr;
  }
}
inline void mm_lower(char *s) {
  for (; *s; ++s) *s = (char)std::tolower((unsigned char)*s);
}
inline bool mm_blank(const char *s) {
  for (; *s; ++s)
    if (!std::isspace((unsigned char)*s)) return false;
  return true;
}
// next line that is not a comment; false at end of file
inline bool mm_data_line(FILE *f, char *line) {
  do {
    if (std::fgets(line, MM_MAX_LINE_LENGTH, f) == nullptr) return false;
  } while (line[0] == '%' || mm_blank(line));
  return true;
}
}  // namespace spops_dropin

// "matrix coordinate real general" and the like (a malloc'ed string), or NULL when invalid
inline char *mm_typecode_to_str(MM_typecode matcode) {
  using spops_dropin::mm_word;
  const char *w[4];
  for (int i = 0; i < 4; ++i)
    if ((w[i] = mm_word(matcode[i], i)) == nullptr) return nullptr;
  char buf[MM_MAX_LINE_LENGTH];
  std::snprintf(buf, sizeof buf, "%s %s %s %s", w[0], w[1], w[2], w[3]);
  return mm_strdup(buf);
}

inline int mm_read_banner(FILE *f, MM_typecode *matcode) {
  char line[MM_MAX_LINE_LENGTH], banner[MM_MAX_TOKEN_LENGTH], mtx[MM_MAX_TOKEN_LENGTH],
      crd[MM_MAX_TOKEN_LENGTH], data[MM_MAX_TOKEN_LENGTH], sym[MM_MAX_TOKEN_LENGTH];
  mm_clear_typecode(matcode);
  if (std::fgets(line, MM_MAX_LINE_LENGTH, f) == nullptr) return MM_PREMATURE_EOF;
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, data, sym) != 5) return MM_PREMATURE_EOF;
  if (std::strncmp(banner, MatrixMarketBanner, std::strlen(MatrixMarketBanner)) != 0) return MM_NO_HEADER;
  spops_dropin::mm_lower(mtx);
  spops_dropin::mm_lower(crd);
  spops_dropin::mm_lower(data);
  spops_dropin::mm_lower(sym);
  if (std::strcmp(mtx, MM_MTX_STR) != 0) return MM_UNSUPPORTED_TYPE;
  mm_set_matrix(matcode);
  if (!std::strcmp(crd, MM_SPARSE_STR)) mm_set_sparse(matcode);
  else if (!std::strcmp(crd, MM_DENSE_STR)) mm_set_dense(matcode);
  else return MM_UNSUPPORTED_TYPE;
  if (!std::strcmp(data, MM_REAL_STR)) mm_set_real(matcode);
  else if (!std::strcmp(data, MM_COMPLEX_STR)) mm_set_complex(matcode);
  else if (!std::strcmp(data, MM_PATTERN_STR)) mm_set_pattern(matcode);
  else if (!std::strcmp(data, MM_INT_STR)) mm_set_integer(matcode);
  else return MM_UNSUPPORTED_TYPE;
  if (!std::strcmp(sym, MM_GENERAL_STR)) mm_set_general(matcode);
  else if (!std::strcmp(sym, MM_SYMM_STR)) mm_set_symmetric(matcode);
  else if (!std::strcmp(sym, MM_HERM_STR)) mm_set_hermitian(matcode);
  else if (!std::strcmp(sym, MM_SKEW_STR)) mm_set_skew(matcode);
  else return MM_UNSUPPORTED_TYPE;
  return 0;
}

inline int mm_write_banner(FILE *f, MM_typecode matcode) {
  char *s = mm_typecode_to_str(matcode);
  if (s == nullptr) return MM_UNSUPPORTED_TYPE;
  const int r = std::fprintf(f, "%s %s\n", MatrixMarketBanner, s);
  std::free(s);
  return r < 0 ? MM_COULD_NOT_WRITE_FILE : 0;
}

// rows, columns and stored entries of a coordinate file (after the banner and comments)
template <typename IndexType>
int mm_read_mtx_crd_size(FILE *f, IndexType *M, IndexType *N, IndexType *nz) {
  char line[MM_MAX_LINE_LENGTH];
  *M = *N = *nz = 0;
  if (!spops_dropin::mm_data_line(f, line)) return MM_PREMATURE_EOF;
  long long m = 0, n = 0, z = 0;
  if (std::sscanf(line, "%lld %lld %lld", &m, &n, &z) != 3) return MM_PREMATURE_EOF;
  *M = (IndexType)m;
  *N = (IndexType)n;
  *nz = (IndexType)z;
  return 0;
}

inline int mm_read_mtx_array_size(FILE *f, int *M, int *N) {
  char line[MM_MAX_LINE_LENGTH];
  *M = *N = 0;
  if (!spops_dropin::mm_data_line(f, line)) return MM_PREMATURE_EOF;
  return std::sscanf(line, "%d %d", M, N) == 2 ? 0 : MM_PREMATURE_EOF;
}

inline int mm_write_mtx_crd_size(FILE *f, int M, int N, int nz) {
  return std::fprintf(f, "%d %d %d\n", M, N, nz) < 0 ? MM_COULD_NOT_WRITE_FILE : 0;
}

inline int mm_write_mtx_array_size(FILE *f, int M, int N) {
  return std::fprintf(f, "%d %d\n", M, N) < 0 ? MM_COULD_NOT_WRITE_FILE : 0;
}

// nz coordinate entries (1-based as in the file) into caller arrays; complex values take two
// doubles per entry, patterns none
inline int mm_read_mtx_crd_data(FILE *f, int M, int N, int nz, int I[], int J[], double val[], MM_typecode matcode) {
  (void)M;
  (void)N;
  for (int i = 0; i < nz; ++i) {
    int got;
    if (mm_is_complex(matcode)) got = std::fscanf(f, "%d %d %lg %lg", &I[i], &J[i], &val[2 * i], &val[2 * i + 1]) - 4;
    else if (mm_is_real(matcode) || mm_is_integer(matcode)) got = std::fscanf(f, "%d %d %lg", &I[i], &J[i], &val[i]) - 3;
    else if (mm_is_pattern(matcode)) got = std::fscanf(f, "%d %d", &I[i], &J[i]) - 2;
    else return MM_UNSUPPORTED_TYPE;
    if (got != 0) return MM_PREMATURE_EOF;
  }
  return 0;
}

inline int mm_read_mtx_crd_entry(FILE *f, int *I, int *J, double *real, double *imag, MM_typecode matcode) {
  int got;
  if (mm_is_complex(matcode)) got = std::fscanf(f, "%d %d %lg %lg", I, J, real, imag) - 4;
  else if (mm_is_real(matcode) || mm_is_integer(matcode)) got = std::fscanf(f, "%d %d %lg", I, J, real) - 3;
  else if (mm_is_pattern(matcode)) got = std::fscanf(f, "%d %d", I, J) - 2;
  else return MM_UNSUPPORTED_TYPE;
  return got == 0 ? 0 : MM_PREMATURE_EOF;
}

// whole coordinate file: sizes, typecode and malloc'ed 1-based I, J (and values unless pattern)
inline int mm_read_mtx_crd(char *fname, int *M, int *N, int *nz, int **I, int **J, double **val,
                           MM_typecode *matcode) {
  const bool std_in = !std::strcmp(fname, "stdin");
  FILE *f = std_in ? stdin : std::fopen(fname, "r");
  if (f == nullptr) return MM_COULD_NOT_READ_FILE;
  int r = mm_read_banner(f, matcode);
  if (r == 0 && !(mm_is_valid(*matcode) && mm_is_sparse(*matcode) && mm_is_matrix(*matcode))) r = MM_UNSUPPORTED_TYPE;
  if (r == 0) r = mm_read_mtx_crd_size(f, M, N, nz);
  if (r == 0) {
    *I = static_cast<int *>(std::malloc(sizeof(int) * (size_t)(*nz > 0 ? *nz : 1)));
    *J = static_cast<int *>(std::malloc(sizeof(int) * (size_t)(*nz > 0 ? *nz : 1)));
    *val = nullptr;
    const int per = mm_is_complex(*matcode) ? 2 : (mm_is_pattern(*matcode) ? 0 : 1);
    if (per) *val = static_cast<double *>(std::malloc(sizeof(double) * (size_t)per * (size_t)(*nz > 0 ? *nz : 1)));
    r = mm_read_mtx_crd_data(f, *M, *N, *nz, *I, *J, *val, *matcode);
  }
  if (!std_in) std::fclose(f);
  return r;
}

inline int mm_write_mtx_crd(char fname[], int M, int N, int nz, int I[], int J[], double val[], MM_typecode matcode) {
  const bool std_out = !std::strcmp(fname, "stdout");
  FILE *f = std_out ? stdout : std::fopen(fname, "w");
  if (f == nullptr) return MM_COULD_NOT_WRITE_FILE;
  int r = mm_write_banner(f, matcode);
  if (r == 0) r = mm_write_mtx_crd_size(f, M, N, nz);
  for (int i = 0; r == 0 && i < nz; ++i) {
    int w;
    if (mm_is_pattern(matcode)) w = std::fprintf(f, "%d %d\n", I[i], J[i]);
    else if (mm_is_complex(matcode)) w = std::fprintf(f, "%d %d %20.16g %20.16g\n", I[i], J[i], val[2 * i], val[2 * i + 1]);
    else if (mm_is_real(matcode) || mm_is_integer(matcode)) w = std::fprintf(f, "%d %d %20.16g\n", I[i], J[i], val[i]);
    else w = -1, r = MM_UNSUPPORTED_TYPE;
    if (w < 0 && r == 0) r = MM_COULD_NOT_WRITE_FILE;
  }
  if (!std_out) std::fclose(f);
  return r;
}

// a real, general coordinate file as 0-based row / column arrays and values (malloc'ed)
inline int mm_read_unsymmetric_sparse(const char *fname, int *M_, int *N_, int *nz_, double **val_, int **I_,
                                      int **J_) {
  MM_typecode code;
  int M, N, nz, *I = nullptr, *J = nullptr;
  double *val = nullptr;
  const int r = mm_read_mtx_crd(const_cast<char *>(fname), &M, &N, &nz, &I, &J, &val, &code);
  if (r != 0 || !(mm_is_real(code) && mm_is_general(code))) {
    std::free(I);
    std::free(J);
    std::free(val);
    return r != 0 ? r : MM_UNSUPPORTED_TYPE;
  }
  for (int i = 0; i < nz; ++i) --I[i], --J[i];
  *M_ = M;
  *N_ = N;
  *nz_ = nz;
  *val_ = val;
  *I_ = I;
  *J_ = J;
  return 0;
}

#endif
