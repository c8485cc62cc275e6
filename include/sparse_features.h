/* Matrix feature extractor with the reference's class and printout (include/sparse_features.h:
 * 13-326, src/sparse_features.cpp:121-1135), the features computed on the GPU
 * (spops_matrix_features) from the matrix as read_csr_matrix reads it:
 *
 *   MTX<I, V> m(matID);
 *   m.MtxLoad(path);                      banner, sizes, entries (symmetric files mirrored)
 *   m.CalculateFeatures();                the 61 features + pattern / value symmetry
 *   m.CalculateTilesExtraFeatures(path);  distance_per_row (the reference's default build
 *                                         computes nothing else there: its GrX_* block needs
 *                                         BSR_ANA, which is off, so they stay -1)
 *   m.FeaturesPrint(); m.ExtraFeaturesPrint();   the reference's text, label for label
 *   m.FeaturesWrite(MAT_FEATURES);        appends the feature vector as one line (below)
 *
 * Symmetry follows src/sparse_features.cpp:724-767 over the file's own entries: pattern_symm is
 * the share of off-diagonal entries (i, j) whose (j, i) is also in the file, value_symm the share
 * whose two values print alike with "%lf" (the first entry of a duplicated position is the one
 * compared), both 1 for files declared symmetric, 0 for rectangular ones.
 *
 * FeaturesWrite appends "<id> <name> <rows> <cols> <nnz> <nnz_ratio> <closed2diag>
 * <is_symmetric> <pattern_symm> <value_symm> <lower> <upper> <diagonal> <ndiags>" then the row
 * and the column statistics (NoneZero ratio %.3f, min, max, ave, var, std, P-ratio, Gini) and
 * distance_per_row -- the layout of the reference's feature-vector writer
 * (src/sparse_features.cpp:1150-1181, the reference's current build writes per-tile image files
 * to a fixed directory of its authors' machine instead; those are not written here).
 * PrintImage prints "image=" as the reference's does (its image is never filled). */
#ifndef SPOPS_DROPIN_SPARSE_FEATURES_H
#define SPOPS_DROPIN_SPARSE_FEATURES_H

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <limits>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "general_config.h"
#include "mmio.h"
#include "sparse_format.h"
#include "sparse_io.h"
#include "spops_dropin.h"
#include "timer.h"

template <typename IndexType, typename ValueType>
class MTX {
 public:
  explicit MTX(IndexType matID = 0) : matrixID_(matID) { std::cout << "== MTX Features Extraction ==\n" << std::endl; }
  IndexType getRowNum() { return num_rows; }
  IndexType getColNum() { return num_cols; }
  IndexType getTileSize() { return t_num_blocks; }
  std::string getMatName() { return matrixName; }

  bool MtxLoad(const char *mat_path);
  bool FeaturesWrite(const char *file_path);
  bool ConvertToCSR(CSR_Matrix<IndexType, ValueType> &csr);
  void Stringsplit(const std::string &s, const char split, std::vector<std::string> &res);
  bool CalculateFeatures();
  bool CalculateTilesFeatures() { return features_done_; }   // computed with the others on the GPU
  bool CalculateTilesExtraFeatures(const char *mat_path);
  bool PrintImage(std::string &outputpath) {
    (void)outputpath;
    std::cout << "image=" << std::endl;
    return true;
  }
  double MtxLoad_time_ = 0.0;
  double CalculateFeatures_time_ = 0.0;
  double ConvertToCSR_time_ = 0.0;

  void FeaturesPrint();
  void ExtraFeaturesPrint();

 private:
  struct Line {
    const char *label;
    double value;
    bool integer, percent;
  };
  void print_lines(const std::vector<Line> &lines);

  bool is_symmetric_ = false, features_done_ = false;
  std::string matrixName;
  IndexType matrixID_ = 0;
  IndexType num_rows = 0, num_cols = 0, num_nnzs = 0, nnz_mtx_ = 0;
  IndexType nnz_lower_ = 0, nnz_upper_ = 0, nnz_diagonal_ = 0, complete_ndiags = 0;
  IndexType min_nnz_each_row_ = 100000000, max_nnz_each_row_ = 0, min_nnz_each_col_ = 100000000,
            max_nnz_each_col_ = 0;
  ValueType pattern_symm_ = 0.0, value_symm_ = 0.0, nnz_ratio_ = 0.0, diag_close_ratio_ = 0.0,
            distance_per_row_ = 0.0;
  ValueType nz_row_ratio_ = 0.0, ave_nnz_each_row_ = 0.0, var_nnz_each_row_ = 0.0, standard_dev_row_ = 0.0,
            Gini_row_ = 0.0, P_ratio_row_ = 0.0;
  ValueType nz_col_ratio_ = 0.0, ave_nnz_each_col_ = 0.0, var_nnz_each_col_ = 0.0, standard_dev_col_ = 0.0,
            Gini_col_ = 0.0, P_ratio_col_ = 0.0;
  ValueType diagonal_dominant_ratio_ = 0.0;
  ValueType max_value_offdiag_ = -std::numeric_limits<ValueType>::max();
  ValueType max_value_diagonal_ = -std::numeric_limits<ValueType>::max();
  ValueType row_variability_ = -1.0, col_variability_ = -1.0;
  IndexType t_num_blocks = MAT_TILE_SIZE, t_num_RB = -1, t_num_CB = -1;
  ValueType t_ave_nnz_all_tiles = -1.0, t_ave_nnz_RB = -1.0, t_ave_nnz_CB = -1.0;
  ValueType t_var_nnz_all_tiles = -1.0, t_var_nnz_RB = -1.0, t_var_nnz_CB = -1.0;
  ValueType t_standard_dev_all_tiles = -1.0, t_standard_dev_RB = -1.0, t_standard_dev_CB = -1.0;
  IndexType t_min_nnz_all_tiles_ = 100000000, t_max_nnz_all_tiles_ = 0, t_min_nnz_each_RB_ = 100000000,
            t_max_nnz_each_RB_ = 0, t_min_nnz_each_CB_ = 100000000, t_max_nnz_each_CB_ = 0;
  ValueType t_Gini_all_tiles_ = -1.0, t_Gini_RB_ = -1.0, t_Gini_CB_ = -1.0;
  ValueType t_P_ratio_all_tiles_ = -1.0, t_P_ratio_RB_ = -1.0, t_P_ratio_CB_ = -1.0;
  ValueType t_nz_ratio_tiles_ = -1.0, t_nz_ratio_RB_ = -1.0, t_nz_ratio_CB_ = -1.0;
  ValueType uniqR = -1.0, uniqC = -1.0, GrX_uniqR = -1.0, GrX_uniqC = -1.0;
  ValueType potReuseR = -1.0, potReuseC = -1.0, GrX_potReuseR = -1.0, GrX_potReuseC = -1.0;

  // the file's entries (0-based, before mirroring) and the matrix as read_csr_matrix reads it
  std::vector<int64_t> file_row_, file_col_;
  std::vector<double> file_val_;
  std::vector<int64_t> rowptr_, col_;
  std::vector<double> val_;
};

template <typename IndexType, typename ValueType>
void MTX<IndexType, ValueType>::Stringsplit(const std::string &s, const char split, std::vector<std::string> &res) {
  res.clear();
  std::string cur;
  for (char ch : s) {
    if (ch == split) {
      if (!cur.empty()) res.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) res.push_back(cur);
}

template <typename IndexType, typename ValueType>
bool MTX<IndexType, ValueType>::MtxLoad(const char *mat_path) {
  timer t;
  FILE *f = std::fopen(mat_path, "r");
  if (f == nullptr) {
    std::cout << "Unable to open mtx file: " << mat_path << std::endl;
    std::exit(1);
  }
  matrixName = extractFileNameWithoutExtension(mat_path);
  MM_typecode code;
  if (mm_read_banner(f, &code) != 0) {
    std::cout << "Could not process Matrix Market banner." << std::endl;
    std::exit(1);
  }
  if (!mm_is_valid(code)) {
    std::cout << "Invalid Matrix" << std::endl;
    std::exit(1);
  }
  if (!((mm_is_real(code) || mm_is_integer(code) || mm_is_pattern(code)) && mm_is_coordinate(code))) {
    std::printf("Sorry, this application does not support Market Market type: [%s]\n", mm_typecode_to_str(code));
    std::printf("Only sparse real-valued or pattern coordinate matrices are supported\n");
    std::exit(1);
  }
  is_symmetric_ = mm_is_symmetric(code) || mm_is_hermitian(code);
  long long r = 0, c = 0, z = 0;
  if (mm_read_mtx_crd_size(f, &r, &c, &z) != 0) {
    std::cout << "The line of rows, cols, nnzs is in wrong format" << std::endl;
    std::exit(1);
  }
  num_rows = (IndexType)r;
  num_cols = (IndexType)c;
  nnz_mtx_ = (IndexType)z;
  std::cout << "- Reading sparse matrix from file: " << mat_path << std::endl;
  file_row_.resize((size_t)z);
  file_col_.resize((size_t)z);
  file_val_.assign((size_t)z, 1.0);
  const bool pattern = mm_is_pattern(code);
  for (long long e = 0; e < z; ++e) {
    long long i = 0, j = 0;
    double v = 1.0;
    const int got = pattern ? std::fscanf(f, "%lld %lld", &i, &j) : std::fscanf(f, "%lld %lld %lf", &i, &j, &v);
    if (got != (pattern ? 2 : 3)) {
      std::cout << "Error reading entry " << e << " of " << mat_path << std::endl;
      std::exit(1);
    }
    file_row_[(size_t)e] = i - 1;
    file_col_[(size_t)e] = j - 1;
    file_val_[(size_t)e] = (double)(ValueType)v;
  }
  std::fclose(f);
  // read_csr_matrix's matrix: each off-diagonal entry of a symmetric file followed by its mirror,
  // then rows in order with entries in that order
  std::vector<int64_t> ri, ci;
  std::vector<double> va;
  const bool mirror = mm_is_symmetric(code);
  for (size_t e = 0; e < file_row_.size(); ++e) {
    ri.push_back(file_row_[e]);
    ci.push_back(file_col_[e]);
    va.push_back(file_val_[e]);
    if (mirror && file_row_[e] != file_col_[e]) {
      ri.push_back(file_col_[e]);
      ci.push_back(file_row_[e]);
      va.push_back(file_val_[e]);
    }
  }
  rowptr_.assign((size_t)num_rows + 1, 0);
  for (int64_t x : ri) rowptr_[(size_t)x + 1]++;
  for (size_t x = 0; x < (size_t)num_rows; ++x) rowptr_[x + 1] += rowptr_[x];
  std::vector<int64_t> cur(rowptr_.begin(), rowptr_.end() - 1);
  col_.resize(ri.size());
  val_.resize(ri.size());
  for (size_t e = 0; e < ri.size(); ++e) {
    const int64_t p = cur[(size_t)ri[e]]++;
    col_[(size_t)p] = ci[e];
    val_[(size_t)p] = va[e];
  }
  num_nnzs = (IndexType)ri.size();
  MtxLoad_time_ = t.milliseconds_elapsed();
  return true;
}

template <typename IndexType, typename ValueType>
bool MTX<IndexType, ValueType>::CalculateFeatures() {
  using namespace spops_dropin;
  timer t;
  double out[SPOPS_NUM_FEATURES];
  {
    DeviceBuffer drp, dcol, dval;
    check(spops_memcpy_htod(drp.get(rowptr_.size() * 8), rowptr_.data(), rowptr_.size() * 8, nullptr), "upload rowptr");
    check(spops_memcpy_htod(dcol.get(std::max<size_t>(col_.size(), 1) * 8), col_.data(), col_.size() * 8, nullptr),
          "upload col");
    check(spops_memcpy_htod(dval.get(std::max<size_t>(val_.size(), 1) * 8), val_.data(), val_.size() * 8, nullptr),
          "upload values");
    check(spops_matrix_features(drp.p, dcol.p, dval.p, SPOPS_INDEX_I64, SPOPS_VALUE_F64, num_rows, num_cols, out,
                                nullptr),
          "spops_matrix_features");
  }
  const auto I = [&](int k) { return (IndexType)out[k]; };
  const auto V = [&](int k) { return (ValueType)out[k]; };
  num_nnzs = I(2);
  nnz_ratio_ = V(3);
  complete_ndiags = I(4);
  diag_close_ratio_ = V(5);
  nnz_lower_ = I(6);
  nnz_upper_ = I(7);
  nnz_diagonal_ = I(8);
  nz_row_ratio_ = V(9); min_nnz_each_row_ = I(10); max_nnz_each_row_ = I(11); ave_nnz_each_row_ = V(12);
  var_nnz_each_row_ = V(13); standard_dev_row_ = V(14); P_ratio_row_ = V(15); Gini_row_ = V(16);
  nz_col_ratio_ = V(17); min_nnz_each_col_ = I(18); max_nnz_each_col_ = I(19); ave_nnz_each_col_ = V(20);
  var_nnz_each_col_ = V(21); standard_dev_col_ = V(22); P_ratio_col_ = V(23); Gini_col_ = V(24);
  diagonal_dominant_ratio_ = V(25); max_value_offdiag_ = V(26); max_value_diagonal_ = V(27);
  row_variability_ = V(28); col_variability_ = V(29);
  t_num_blocks = I(30); t_num_RB = I(31); t_num_CB = I(32);
  t_nz_ratio_tiles_ = V(33); t_min_nnz_all_tiles_ = I(34); t_max_nnz_all_tiles_ = I(35); t_ave_nnz_all_tiles = V(36);
  t_var_nnz_all_tiles = V(37); t_standard_dev_all_tiles = V(38); t_P_ratio_all_tiles_ = V(39);
  t_Gini_all_tiles_ = V(40);
  t_nz_ratio_RB_ = V(41); t_min_nnz_each_RB_ = I(42); t_max_nnz_each_RB_ = I(43); t_ave_nnz_RB = V(44);
  t_var_nnz_RB = V(45); t_standard_dev_RB = V(46); t_P_ratio_RB_ = V(47); t_Gini_RB_ = V(48);
  t_nz_ratio_CB_ = V(49); t_min_nnz_each_CB_ = I(50); t_max_nnz_each_CB_ = I(51); t_ave_nnz_CB = V(52);
  t_var_nnz_CB = V(53); t_standard_dev_CB = V(54); t_P_ratio_CB_ = V(55); t_Gini_CB_ = V(56);
  uniqR = V(57); uniqC = V(58); potReuseR = V(59); potReuseC = V(60);
  // pattern / value symmetry over the file's entries; the diagonal count is the file's
  IndexType file_diag = 0;
  for (size_t e = 0; e < file_row_.size(); ++e) file_diag += file_row_[e] == file_col_[e] ? 1 : 0;
  if (is_symmetric_) {
    pattern_symm_ = 1.0;
    value_symm_ = 1.0;
  } else if (num_rows == num_cols) {
    struct PairKey {
      size_t operator()(const std::pair<int64_t, int64_t> &p) const {
        return std::hash<int64_t>()(p.first * 1000003 + p.second);
      }
    };
    std::unordered_map<std::pair<int64_t, int64_t>, std::string, PairKey> first;   // first value per position
    for (size_t e = 0; e < file_row_.size(); ++e) {
      char buf[64];
      std::snprintf(buf, sizeof buf, "%lf", (double)(ValueType)file_val_[e]);
      first.emplace(std::make_pair(file_row_[e], file_col_[e]), std::string(buf));
    }
    IndexType pat = 0, vals = 0;
    for (size_t e = 0; e < file_row_.size(); ++e) {
      if (file_row_[e] == file_col_[e]) continue;
      const auto mirror = first.find(std::make_pair(file_col_[e], file_row_[e]));
      if (mirror == first.end()) continue;
      ++pat;
      if (mirror->second == first.at(std::make_pair(file_row_[e], file_col_[e]))) ++vals;
    }
    pattern_symm_ = (ValueType)pat / (nnz_mtx_ - file_diag);
    value_symm_ = (ValueType)vals / (nnz_mtx_ - file_diag);
  } else {
    pattern_symm_ = 0.0;
    value_symm_ = 0.0;
  }
  is_symmetric_ = pattern_symm_ == 1.0 && value_symm_ == 1.0;
  features_done_ = true;
  CalculateFeatures_time_ = t.milliseconds_elapsed();
  return true;
}

template <typename IndexType, typename ValueType>
bool MTX<IndexType, ValueType>::CalculateTilesExtraFeatures(const char *mat_path) {
  // row by row, the column distance from one stored entry to the next, summed (it telescopes to
  // last column - first column), divided by the row's entry count; averaged over all rows
  CSR_Matrix<IndexType, ValueType> csr = read_csr_matrix<IndexType, ValueType>(mat_path);
  distance_per_row_ = 0.0;
  for (IndexType i = 0; i < csr.num_rows; i++) {
    const IndexType len = csr.row_offset[i + 1] - csr.row_offset[i];
    if (len == 0) continue;
    const IndexType dis = len > 1 ? csr.col_index[csr.row_offset[i + 1] - 1] - csr.col_index[csr.row_offset[i]] : 0;
    distance_per_row_ += (ValueType)dis / len;
  }
  distance_per_row_ = distance_per_row_ / csr.num_rows;
  delete_csr_matrix(csr);
  return true;
}

template <typename IndexType, typename ValueType>
bool MTX<IndexType, ValueType>::ConvertToCSR(CSR_Matrix<IndexType, ValueType> &csr) {
  timer t;
  csr.num_rows = num_rows;
  csr.num_cols = num_cols;
  csr.num_nnzs = (IndexType)col_.size();
  csr.tag = 0;
  csr.kernel_flag = KERNEL_FLAG;
  csr.partition = nullptr;
  csr.row_offset = new_array<IndexType>(rowptr_.size());
  csr.col_index = new_array<IndexType>(std::max<size_t>(col_.size(), 1));
  csr.values = new_array<ValueType>(std::max<size_t>(val_.size(), 1));
  for (size_t i = 0; i < rowptr_.size(); ++i) csr.row_offset[i] = (IndexType)rowptr_[i];
  for (size_t i = 0; i < col_.size(); ++i) {
    csr.col_index[i] = (IndexType)col_[i];
    csr.values[i] = (ValueType)val_[i];
  }
  ConvertToCSR_time_ = t.milliseconds_elapsed();
  return true;
}

template <typename IndexType, typename ValueType>
void MTX<IndexType, ValueType>::print_lines(const std::vector<Line> &lines) {
  for (const Line &l : lines) {
    if (l.label == nullptr) {
      std::cout << std::endl;
      continue;
    }
    std::cout << l.label;
    if (l.integer) std::cout << (IndexType)l.value;
    else std::cout << (ValueType)(l.percent ? l.value * 100 : l.value);
    if (l.percent) std::cout << "%";
    std::cout << std::endl;
  }
}

template <typename IndexType, typename ValueType>
void MTX<IndexType, ValueType>::FeaturesPrint() {
  std::cout << "[" << matrixID_ << "]  " << matrixName << "counting by " << 8 * sizeof(IndexType) << "-bit Index, "
            << 8 * sizeof(ValueType) << "-bit Value" << std::endl;
  std::cout << "is_symmetric " << is_symmetric_ << std::endl;
  const Line blank{nullptr, 0, false, false};
  auto in = [](const char *s, double v) { return Line{s, v, true, false}; };
  auto fl = [](const char *s, double v) { return Line{s, v, false, false}; };
  auto pc = [](const char *s, double v) { return Line{s, v, false, true}; };
  print_lines({pc("pattern_symm ", pattern_symm_), pc("value_symm   ", value_symm_), blank,
               in("num_rows   = ", num_rows), in("num_cols   = ", num_cols), in("nnz_mtx    = ", nnz_mtx_),
               in("real_nnz   = ", num_nnzs), pc("nnz_ratio  = ", nnz_ratio_), in("num_of_diags  = ", complete_ndiags),
               pc("nnz_closed2diag_ratio = ", diag_close_ratio_), blank,
               in("nnz_lower    : ", nnz_lower_), in("nnz_upper    : ", nnz_upper_),
               in("nnz_diagonal : ", nnz_diagonal_), blank,
               pc("NoneZero_Row_Ratio      : ", nz_row_ratio_), in("min_nnz_each_row        : ", min_nnz_each_row_),
               in("max_nnz_each_row        : ", max_nnz_each_row_), fl("ave_nnz_each_row        : ", ave_nnz_each_row_),
               fl("var_nnz_each_row        : ", var_nnz_each_row_), fl("standard_dev_row        : ", standard_dev_row_),
               fl("P-ratio_row             : ", P_ratio_row_), fl("Gini_coeff_row          : ", Gini_row_), blank,
               pc("NoneZero_Col_Ratio      : ", nz_col_ratio_), in("min_nnz_each_col        : ", min_nnz_each_col_),
               in("max_nnz_each_col        : ", max_nnz_each_col_), fl("ave_nnz_each_col        : ", ave_nnz_each_col_),
               fl("var_nnz_each_col        : ", var_nnz_each_col_), fl("standard_dev_col        : ", standard_dev_col_),
               fl("P-ratio_col             : ", P_ratio_col_), fl("Gini_coeff_col          : ", Gini_col_), blank,
               pc("diagonal_dominant_ratio : ", diagonal_dominant_ratio_), blank,
               fl("max_value_offdiag  = ", max_value_offdiag_), fl("max_value_diagonal = ", max_value_diagonal_),
               fl("row_variability    = ", row_variability_), fl("col_variability    = ", col_variability_), blank});
  std::cout << "==========   Tile Features   ==========" << std::endl;
  std::cout << "Number of Tiles  : " << t_num_blocks << std::endl;
  std::cout << "Tile inner size  : " << t_num_RB << " * " << t_num_CB << std::endl;
  print_lines({fl("t_ave_nnz_tiles     : ", t_ave_nnz_all_tiles), fl("t_ave_nnz_RB        : ", t_ave_nnz_RB),
               fl("t_ave_nnz_CB        : ", t_ave_nnz_CB), blank,
               fl("t_var_nnz_tiles     : ", t_var_nnz_all_tiles), fl("t_var_nnz_RB        : ", t_var_nnz_RB),
               fl("t_var_nnz_CB        : ", t_var_nnz_CB), blank,
               fl("t_stand_dev_tiles     : ", t_standard_dev_all_tiles), fl("t_stand_dev_RB        : ", t_standard_dev_RB),
               fl("t_stand_dev_CB        : ", t_standard_dev_CB), blank,
               in("min_nnz_each_tiles       : ", t_min_nnz_all_tiles_), in("max_nnz_each_tiles       : ", t_max_nnz_all_tiles_),
               in("min_nnz_each_RB          : ", t_min_nnz_each_RB_), in("max_nnz_each_RB          : ", t_max_nnz_each_RB_),
               in("min_nnz_each_CB          : ", t_min_nnz_each_CB_), in("max_nnz_each_CB          : ", t_max_nnz_each_CB_),
               blank,
               fl("Gini_nnz_tiles     : ", t_Gini_all_tiles_), fl("Gini_nnz_RB        : ", t_Gini_RB_),
               fl("Gini_nnz_CB        : ", t_Gini_CB_), blank,
               fl("P-ratio_tiles      : ", t_P_ratio_all_tiles_), fl("P-ratio_RB         : ", t_P_ratio_RB_),
               fl("P-ratio_CB         : ", t_P_ratio_CB_), blank,
               fl("NE_ratio_tiles_     : ", t_nz_ratio_tiles_), fl("NE_ratio_RB         : ", t_nz_ratio_RB_),
               fl("NE_ratio_CB         : ", t_nz_ratio_CB_), blank});
}

template <typename IndexType, typename ValueType>
void MTX<IndexType, ValueType>::ExtraFeaturesPrint() {
  std::cout << "==========   Tile Extra Features   ==========" << std::endl;
  const Line blank{nullptr, 0, false, false};
  auto fl = [](const char *s, double v) { return Line{s, v, false, false}; };
  print_lines({fl("uniq_R         = ", uniqR), fl("uniq_C         = ", uniqC), blank,
               fl("GrX_uniqR      = ", GrX_uniqR), fl("GrX_uniqC      = ", GrX_uniqC), blank,
               fl("potReuseR      = ", potReuseR), fl("potReuseC      = ", potReuseC), blank,
               fl("GrX_potReuseR  = ", GrX_potReuseR), fl("GrX_potReuseC  = ", GrX_potReuseC), blank,
               fl("ave_distance_= ", distance_per_row_), blank});
}

template <typename IndexType, typename ValueType>
bool MTX<IndexType, ValueType>::FeaturesWrite(const char *file_path) {
  FILE *f = std::fopen(file_path, "a");
  if (f == nullptr) {
    std::cout << "Unable to open features saved file: " << file_path << std::endl;
    return false;
  }
  const auto L = [](IndexType v) { return (long long)v; };
  const auto D = [](ValueType v) { return (double)v; };
  std::fprintf(f, "%lld %s %lld %lld %lld %lf %lf %d %lf %lf %lld %lld %lld %lld ", L(matrixID_), matrixName.c_str(),
               L(num_rows), L(num_cols), L(num_nnzs), D(nnz_ratio_), D(diag_close_ratio_), (int)is_symmetric_,
               D(pattern_symm_), D(value_symm_), L(nnz_lower_), L(nnz_upper_), L(nnz_diagonal_), L(complete_ndiags));
  std::fprintf(f, "%.3f %lld %lld %lf %lf %lf %lf %lf ", D(nz_row_ratio_), L(min_nnz_each_row_), L(max_nnz_each_row_),
               D(ave_nnz_each_row_), D(var_nnz_each_row_), D(standard_dev_row_), D(P_ratio_row_), D(Gini_row_));
  std::fprintf(f, "%.3f %lld %lld %lf %lf %lf %lf %lf ", D(nz_col_ratio_), L(min_nnz_each_col_), L(max_nnz_each_col_),
               D(ave_nnz_each_col_), D(var_nnz_each_col_), D(standard_dev_col_), D(P_ratio_col_), D(Gini_col_));
  std::fprintf(f, "%lf\n", D(distance_per_row_));
  std::fclose(f);
  return true;
}

#endif
